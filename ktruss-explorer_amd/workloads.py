"""Seeded synthetic workloads shared by the benchmark, the tests and the
autotuner.

batch_specs / make_batch: the many-graph workload.  Graph i uses seed
base + i; its target edge count is log-uniform in [2^10, 2^18]; even graphs
are Erdos-Renyi with mean degree log-uniform in [8, 64] (n = 2m / degree),
odd graphs are R-MAT (a, b, c, d) = (.57, .19, .19, .05) with
scale = clamp(round(log2(m / 16)), 8, 14) and edge factor
clamp(round(m / 2^scale), 2, 2^(scale - 3)).  Edge counts after cleaning
(self-loops and duplicates removed) fall somewhat below the targets.
"""
import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import ktruss_amd as K


def batch_specs(count=4096, base_seed=1, min_log2=10, max_log2=18):
    specs = []
    for i in range(count):
        rng = np.random.default_rng(base_seed + i)
        m = 2.0 ** rng.uniform(min_log2, max_log2)
        seed = int(rng.integers(1, 1 << 62))
        if i % 2 == 0:
            deg = 2.0 ** rng.uniform(3, 6)
            n = max(int(deg) + 2, int(round(2 * m / deg)))
            specs.append({"kind": "er", "n": n, "deg": float(deg), "seed": seed})
        else:
            scale = min(14, max(8, int(round(math.log2(m / 16)))))
            ef = min(1 << (scale - 3), max(2, int(round(m / (1 << scale)))))
            specs.append({"kind": "rmat", "scale": scale, "ef": ef, "seed": seed})
    return specs


def make_graph(spec):
    if spec["kind"] == "er":
        n, s, d = K.generate_erdos_renyi(spec["n"], spec["deg"], spec["seed"])
    else:
        n, s, d = K.generate_rmat(spec["scale"], spec["ef"], seed=spec["seed"])
    return n, s, d


def make_batch(count=4096, base_seed=1, threads=16, keep_pairs=False, **kw):
    """Graphs of the many-graph workload (generated on `threads` threads; the
    generators release the interpreter while they run)."""
    specs = batch_specs(count, base_seed, **kw)

    def one(spec):
        n, s, d = make_graph(spec)
        g = K.Graph(n, s, d, pairs=True)
        return (g, (n, s, d)) if keep_pairs else (g, None)

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        out = list(ex.map(one, specs))
    graphs = [g for g, _ in out]
    pairs = [p for _, p in out] if keep_pairs else None
    return specs, graphs, pairs
