"""Writes the committed truss-number fixtures tests/golden/rmat<scale>_tau.json.

For R-MAT(scale, edge factor 16, a/b/c = .57/.19/.19, seed 1) produced by the
engine's own generator, the CPU oracle decomposes the graph and the fixture
records a SHA-256 of the truss numbers (uint32, little-endian, in the
generator's sorted pair order), their histogram and k_max, plus a SHA-256 of
the generated pairs so a generator change is caught.

    python tests/golden/make_golden.py 14 16 18
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "ktruss-explorer_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import ktruss_amd as K  # noqa: E402
import oracle as O  # noqa: E402


def pairs_digest(s, d):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(s, "<u4").tobytes())
    h.update(np.ascontiguousarray(d, "<u4").tobytes())
    return h.hexdigest()


def tau_digest(tau):
    return hashlib.sha256(np.ascontiguousarray(tau, "<u4").tobytes()).hexdigest()


def histogram(tau):
    vals, counts = np.unique(tau, return_counts=True)
    return {str(int(v)): int(c) for v, c in zip(vals, counts)}


def make(scale):
    n, s, d = K.generate_rmat(scale, 16, seed=1)
    ss, dd = K.symmetrize(s, d)
    t0 = time.time()
    tau, st, _ = O.decompose(n, ss, dd)
    dt = time.time() - t0
    tau = tau[: len(s)]
    out = {
        "graph": f"R-MAT scale {scale}, edge factor 16, a/b/c = .57/.19/.19, seed 1",
        "n": int(n),
        "m": int(len(s)),
        "pairs_sha256": pairs_digest(s, d),
        "tau_sha256": tau_digest(tau),
        "tau_histogram": histogram(tau),
        "k_max": int(st["k_max"]),
        "triangles": int(st["triangles"]),
        "oracle_seconds": round(dt, 1),
    }
    path = os.path.join(HERE, f"rmat{scale}_tau.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path, out["k_max"], out["oracle_seconds"], flush=True)


if __name__ == "__main__":
    for a in sys.argv[1:]:
        make(int(a))
