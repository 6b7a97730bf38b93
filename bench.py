#!/usr/bin/env python3
"""Full k-truss decomposition throughput on one MI355X.

A step is one full decomposition (k++ loop over the peel) of a synthetic
R-MAT graph whose oriented COO+CSR layout is already resident in HBM; the
step ends with the truss number of every edge resident in HBM.  Workload:
R-MAT scale 20, edge factor 16, (a, b, c, d) = (.57, .19, .19, .05), seed 1,
symmetrised, self-loops and duplicates removed; default variant
(directed-by-degree / coocsr / recount all).

Prints ONE JSON line.  value = undirected edges / seconds per step (wall time
of K steps between two device synchronisations).  The roofline object is for
the support-count kernel: algorithmic bytes of all its launches in the timed
steps divided by their summed HIP-event durations on the library's stream.
cpu_baseline times the CPU oracle (single thread) on a bounded sample of the
same workload, R-MAT scale 14 with the same parameters.
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ktruss-explorer_amd"))

METRIC = ("full k-truss edges/s at 1 GPU (learned-pick vs default variant), bit-exact; HBM GB/s")
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "count_kernel_traffic.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(scale, edge_factor, seed):
    """The oracle (C restatement of the reference's CPU path plus the k++
    loop), one thread, on R-MAT `scale`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ktruss_amd as K
    import oracle as O

    n, s, d = K.generate_rmat(scale, edge_factor, seed=seed)
    ss, dd = K.symmetrize(s, d)
    t0 = time.perf_counter()
    O.decompose(n, ss, dd, O.make_variant(), O.SCHED_JUMP | O.SCHED_REUSE)
    dt = time.perf_counter() - t0
    return {
        "value": len(s) / dt,
        "unit": "edges/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"R-MAT scale {scale}, edge factor {edge_factor}, seed {seed} ({len(s)} edges): "
                   f"full decomposition by the CPU oracle, default variant, {dt:.1f} s on one "
                   f"thread of {cpu_model()} ({os.cpu_count()} visible CPUs)"),
    }


def cpu_baseline_batch(specs, budget_s=20.0):
    """The oracle, one thread, on the first graphs of the many-graph workload
    until about budget_s seconds of CPU work are spent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ktruss_amd as K
    import oracle as O
    import workloads as W

    edges, spent, done = 0, 0.0, 0
    for spec in specs:
        n, s, d = W.make_graph(spec)
        ss, dd = K.symmetrize(s, d)
        t0 = time.perf_counter()
        O.decompose(n, ss, dd, O.make_variant(), O.SCHED_JUMP | O.SCHED_REUSE)
        spent += time.perf_counter() - t0
        edges += len(s)
        done += 1
        if spent >= budget_s:
            break
    return {
        "value": edges / spent,
        "unit": "edges/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"the first {done} graphs of the workload ({edges} edges): full decomposition "
                   f"of each by the CPU oracle, default variant, {spent:.1f} s on one thread of "
                   f"{cpu_model()} ({os.cpu_count()} visible CPUs)"),
    }


def bench_batch(args, K):
    """Many-graph workload: every graph of the batch decomposed together."""
    import workloads as W

    t0 = time.perf_counter()
    specs, graphs, _ = W.make_batch(args.batch_graphs, args.seed, threads=16)
    m_total = sum(g.m for g in graphs)
    log(f"generated {len(graphs)} graphs, {m_total} edges in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    b = K.Batch(graphs)
    log(f"batch layout: {b.entries} entries, longest row {b.max_row}, "
        f"{time.perf_counter() - t0:.1f} s")
    if not args.no_autotune:
        t0 = time.perf_counter()
        best = b.autotune(trial_steps=args.tune_steps, repeats=1)
        log(f"autotune: {len(b.tune_log())} trials in {time.perf_counter() - t0:.1f} s -> {best}")
    for i in range(args.warmup):
        st = b.decompose()
        log(f"warmup {i}: {st['ms_total']:.1f} ms device, {st['steps']} steps, "
            f"{st['replays']} replays, k_max {st['k_max']}")
    K.device_synchronize()
    per_step = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = b.decompose()
        per_step.append(st["ms_total"])
        log(f"step {i}: {st['ms_total']:.1f} ms device")
    K.device_synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / max(args.steps, 1)
    # count-kernel roofline from one extra run in timing mode (outside the timed steps)
    b.set_timing(True)
    stt = b.decompose()
    b.set_timing(False)
    achieved = stt["balg_count_bytes"] / (stt["ms_count"] * 1e-3) / 1e9 if stt["ms_count"] else 0.0
    out = {
        "metric": "full k-truss edges/s at 1 GPU, many-graph batch, bit-exact; HBM GB/s",
        "value": m_total / (ms_per_step * 1e-3),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": (f"{len(graphs)} seeded graphs (Erdos-Renyi mean degree 8-64 and R-MAT "
                         f"scale 8-14, 1k-256k edges log-uniform, base seed {args.seed}): full "
                         f"k-truss decomposition of every graph in one batch"),
            "variant": "directed-by-degree/coocsr/all, batched",
            "graphs": len(graphs),
            "m": m_total,
            "k_max": st["k_max"],
            "steps": st["steps"],
            "tuning": b.tuning(),
            "median_device_ms": statistics.median(per_step) if per_step else None,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "k_count_dir<true> (batched support count)",
            "alg_bytes_per_launch": stt["balg_count_bytes"] / max(stt["count_launches"], 1),
            "avg_launch_ms": stt["ms_count"] / max(stt["count_launches"], 1),
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batch(specs)
    print(json.dumps(out), flush=True)


def parse_variant(spec, K):
    if spec == "default":
        return K.default_variant()
    parts = spec.split("/")
    return K.make_variant(*parts[:3])


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--variant", default="default",
                    help="default or directedness/format/recount, e.g. undirected/coocsr/affected")
    ap.add_argument("--cpu-scale", type=int, default=14)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["rmat", "batch"], default="rmat",
                    help="rmat: one R-MAT graph (default); batch: the many-graph workload")
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--tune-steps", type=int, default=6)
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one MI355X, and the k-truss path "
              "is built and measured on a single device (see DESIGN.md for how it would shard).",
              file=sys.stderr)
        sys.exit(2)

    import ktruss_amd as K

    K.init(0)
    if args.workload == "batch":
        bench_batch(args, K)
        return
    log(f"generating R-MAT scale {args.scale} ef {args.edge_factor} seed {args.seed}")
    n, s, d = K.generate_rmat(args.scale, args.edge_factor, seed=args.seed)
    g = K.Graph(n, s, d, pairs=True)
    v = parse_variant(args.variant, K)
    log(f"graph: n={n} m={g.m} max degree={g.max_degree}; building plan {v.describe()}")
    plan = K.Plan(g, v)

    for i in range(args.warmup):
        st = plan.decompose()
        log(f"warmup {i}: {st['ms_total']:.1f} ms device, k_max {st['k_max']}, "
            f"{st['iterations']} iterations")

    K.device_synchronize()
    per_step = []
    ms_count = 0.0
    launches = 0
    balg = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = plan.decompose()
        per_step.append(st["ms_total"])
        ms_count += st["ms_count"]
        launches += st["count_launches"]
        balg += st["balg_count_bytes"]
        log(f"step {i}: {st['ms_total']:.1f} ms device, count kernel {st['ms_count']:.1f} ms "
            f"in {st['count_launches']} launches")
    K.device_synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / max(args.steps, 1)

    achieved = balg / (ms_count * 1e-3) / 1e9 if ms_count > 0 else 0.0
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        if t.get("workload_scale") == args.scale and t.get("variant") == v.describe():
            traffic = t.get("hbm_bytes_per_launch")
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": "k_count (support count)",
        "alg_bytes_per_launch": balg / max(launches, 1),
        "avg_launch_ms": ms_count / max(launches, 1),
    }
    out = {
        "metric": METRIC,
        "value": g.m / (ms_per_step * 1e-3),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": (f"R-MAT scale {args.scale}, edge factor {args.edge_factor}, a/b/c/d = "
                         f".57/.19/.19/.05, seed {args.seed}, cleaned: full k-truss decomposition"),
            "variant": v.describe(),
            "n": n,
            "m": g.m,
            "k_max": st["k_max"],
            "iterations": st["iterations"],
            "median_device_ms": statistics.median(per_step) if per_step else None,
        },
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        log(f"cpu baseline: oracle on R-MAT scale {args.cpu_scale}")
        out["cpu_baseline"] = cpu_baseline(args.cpu_scale, args.edge_factor, args.seed)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
