#!/usr/bin/env python3
"""Full k-truss decomposition throughput on one MI355X.

A step is one full decomposition (k++ loop over the peel) of a synthetic
R-MAT graph whose oriented COO+CSR layout is already resident in HBM; the
step ends with the truss number of every edge resident in HBM.  Workload:
R-MAT scale 22 (the held-out graph the cost model is evaluated on), edge
factor 16, (a, b, c, d) = (.57, .19, .19, .05), seed 1, symmetrised,
self-loops and duplicates removed.

The measured steps use the learned per-phase pick (the cost model in
ktruss-explorer_amd/picker/ chooses, at every phase, among recount
configurations and DELTA support maintenance); one run of the fixed default
variant (directed-by-degree / coocsr / recount all, the reference's method)
is timed beside it, and the two are checked bit-exact against each other.

Prints ONE JSON line.  value = undirected edges / seconds per step (wall time
of K steps between two device synchronisations).  The roofline object is for
the kernel with the most device time in the measured steps: its algorithmic
bytes divided by the summed HIP-event durations of its launches on the
library's stream.  cpu_baseline times the CPU oracle (single thread) on a
bounded sample of the same workload, R-MAT scale 14 with the same parameters.
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time
import numpy as np

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
    bv = K.make_variant(count_kernel="delta") if args.batch_mode == "delta" else K.default_variant()
    b = K.Batch(graphs, bv)
    log(f"batch layout ({args.batch_mode}): {b.entries} entries, longest row {b.max_row}, "
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
    if args.batch_mode == "delta":
        roofline = roofline_obj(stt["balg_delta_bytes"], stt["ms_delta"], max(stt["steps"], 1),
                                "k_delta (DELTA support update, batched)")
        other = roofline_obj(stt["balg_count_bytes"], stt["ms_count"], 1,
                             count_kernel_name(b.entries).replace(" (", "<true> (first full count, batched, ", 1))
    else:
        roofline = roofline_obj(stt["balg_count_bytes"], stt["ms_count"], stt["count_launches"],
                                count_kernel_name(b.entries).replace(" (", "<true> (batched, ", 1))
        other = None
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
            "variant": ("directed-by-degree/coocsr/all, batched, supports maintained by DELTA"
                        if args.batch_mode == "delta" else "directed-by-degree/coocsr/all, batched"),
            "graphs": len(graphs),
            "m": m_total,
            "k_max": st["k_max"],
            "steps": st["steps"],
            "tuning": b.tuning(),
            "median_device_ms": statistics.median(per_step) if per_step else None,
        },
        "roofline": roofline,
    }
    if other:
        out["config"]["other_kernel_roofline"] = other
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batch(specs)
    print(json.dumps(out), flush=True)


def parse_variant(spec, K):
    if spec == "default":
        return K.default_variant()
    if spec == "delta":
        return K.make_variant(count_kernel="delta")
    parts = spec.split("/")
    return K.make_variant(*parts[:3])


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--mode", choices=["learned", "fixed"], default="learned",
                    help="learned: the cost model picks per phase (default); fixed: --variant")
    ap.add_argument("--variant", default="default",
                    help="with --mode fixed: default, delta, or directedness/format/recount")
    ap.add_argument("--skip-default", action="store_true",
                    help="skip the comparison run of the fixed default variant")
    ap.add_argument("--cpu-scale", type=int, default=14)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["rmat", "batch", "skewed"], default="rmat",
                    help="rmat: one R-MAT graph (default); batch: the many-graph workload; "
                         "skewed: the Chung-Lu graph with planted cliques, every variant swept")
    ap.add_argument("--skew-scale", type=int, default=22,
                    help="with --workload skewed: 2^scale vertices")
    ap.add_argument("--probe-k", type=int, default=8,
                    help="with --workload skewed: the k every variant's k-truss run finds")
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--batch-mode", choices=["delta", "recount"], default="delta")
    ap.add_argument("--tune-steps", type=int, default=6)
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: the k-truss path runs on a single MI355X, so only "
              "--gpus 1 is accepted.", file=sys.stderr)
        sys.exit(2)

    import ktruss_amd as K

    K.init(0)
    if args.workload == "batch":
        bench_batch(args, K)
        return
    if args.workload == "skewed":
        bench_skewed(args, K)
        return
    bench_rmat(args, K)


def count_kernel_name(entries):
    """The directed COO+CSR support count the library launches for a layout of
    `entries` stored entries: window-staged while the layout and supports (8
    bytes per entry) fit 192 MiB of the last-level cache, row-staged beyond."""
    if entries * 8 <= (192 << 20):
        return "k_count_win (support count, window-staged)"
    return "k_count_dir (support count, row-staged)"


def roofline_obj(balg, ms, launches, kernel, traffic=None):
    achieved = balg / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": kernel,
        "alg_bytes_per_launch": balg / max(launches, 1),
        "avg_launch_ms": ms / max(launches, 1),
    }


def bench_rmat(args, K):
    """One R-MAT graph: the learned per-phase pick (the measured steps) beside
    one run of the fixed default variant (the reference's method)."""
    log(f"generating R-MAT scale {args.scale} ef {args.edge_factor} seed {args.seed}")
    n, s, d = K.generate_rmat(args.scale, args.edge_factor, seed=args.seed)
    g = K.Graph(n, s, d, pairs=True)
    log(f"graph: n={n} m={g.m} max degree={g.max_degree}")

    default = None
    tau_default = None
    if not args.skip_default:
        pd = K.Plan(g, K.default_variant())
        K.device_synchronize()
        t0 = time.perf_counter()
        st = pd.decompose()
        K.device_synchronize()
        wall = time.perf_counter() - t0
        tau_default = pd.truss_numbers()
        default = {"ms_per_step": 1e3 * wall, "device_ms": st["ms_total"],
                   "edges_per_s": g.m / wall, "k_max": st["k_max"],
                   "iterations": st["iterations"],
                   "count_roofline": roofline_obj(st["balg_count_bytes"], st["ms_count"],
                                                  st["count_launches"], "support count (k_count_win while "
                                                  "the layout fits the cache, else k_count_dir)")}
        log(f"default variant: {1e3 * wall:.1f} ms, k_max {st['k_max']}, {st['iterations']} iterations")
        del pd

    if args.mode == "learned":
        import learn
        model, configs = learn.load_model()
        plan = K.Plan(g, K.default_variant())
        plan.set_candidates(configs)
        plan.set_pick(K.PICK_MODEL, model)
        label = "learned per-phase pick over " + ", ".join(c.describe() for c in configs)
    else:
        v = parse_variant(args.variant, K)
        plan = K.Plan(g, v)
        label = "fixed " + v.describe()
    log(f"measured: {label}")

    # the timed steps run without per-pass events (they cost the device gaps
    # between dependent kernels); one more decomposition with them gives the
    # kernels' own durations for the roofline
    plan.set_timing(False)
    setup_ms = 0.0  # DELTA adjacency lists, built by the plan's first run (outside ms_total)
    for i in range(args.warmup):
        st = plan.decompose()
        setup_ms += st["ms_setup"]
        log(f"warmup {i}: {st['ms_total']:.1f} ms device, k_max {st['k_max']}"
            + (f", set-up {st['ms_setup']:.1f} ms" if st["ms_setup"] else ""))
    K.device_synchronize()
    per_step = []
    setup_in_timed = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = plan.decompose()
        per_step.append(st["ms_total"])
        setup_in_timed += st["ms_setup"]
        log(f"step {i}: {st['ms_total']:.1f} ms device")
    K.device_synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / max(args.steps, 1)
    tau = plan.truss_numbers()
    plan.set_timing(True)
    st = plan.decompose()
    acc = {key: st[key] for key in ("ms_count", "count_launches", "balg_count_bytes", "ms_delta",
                                    "delta_passes", "balg_delta_bytes")}
    log(f"timing run: {st['ms_total']:.1f} ms device (count {st['ms_count']:.1f} ms, "
        f"DELTA {st['ms_delta']:.1f} ms)")
    if not np.array_equal(plan.truss_numbers(), tau):
        raise SystemExit("timing run changed the truss numbers")

    count_rl = roofline_obj(acc["balg_count_bytes"], acc["ms_count"], acc["count_launches"],
                            count_kernel_name(plan.entries))
    delta_rl = roofline_obj(acc["balg_delta_bytes"], acc["ms_delta"], acc["delta_passes"],
                            "k_delta (DELTA support update)")
    if os.path.exists(TRAFFIC_FILE):
        # HBM bytes per launch measured by rocprof counters on this workload
        with open(TRAFFIC_FILE) as f:
            t = json.load(f).get("by_scale", {}).get(str(args.scale), {})
        for rl in (count_rl, delta_rl):
            key = rl["kernel"].split()[0]
            if args.edge_factor == 16 and args.seed == 1 and key in t.get("hbm_bytes_per_launch", {}):
                rl["traffic"] = t["hbm_bytes_per_launch"][key]
    dominant = delta_rl if acc["ms_delta"] > acc["ms_count"] else count_rl
    other = count_rl if dominant is delta_rl else delta_rl
    picks = None
    if args.mode == "learned":
        ph = plan.phase_log()
        names = [c.describe() for c in plan.candidates]
        counts = {}
        for c in ph[:, 2].astype(int):
            counts[names[c]] = counts.get(names[c], 0) + 1
        picks = counts
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
            "measured": label,
            "n": n,
            "m": g.m,
            "k_max": st["k_max"],
            "iterations": st["iterations"],
            "median_device_ms": statistics.median(per_step) if per_step else None,
            "delta_setup_ms": round(setup_ms + setup_in_timed, 2),
            "delta_setup_in_timed_steps": setup_in_timed > 0,
            "phase_picks": picks,
            "default_variant": default,
            "speedup_vs_default": (default["ms_per_step"] / ms_per_step) if default else None,
            "bit_exact_vs_default": (bool((tau == tau_default).all())
                                     if tau_default is not None else None),
            "other_kernel_roofline": other,
        },
        "roofline": dominant,
    }
    if not args.no_cpu_baseline:
        log(f"cpu baseline: oracle on R-MAT scale {args.cpu_scale}")
        out["cpu_baseline"] = cpu_baseline(args.cpu_scale, args.edge_factor, args.seed)
    print(json.dumps(out), flush=True)


def undirected_keys(src, dst):
    """Sorted unique (min, max) pairs of a COO, packed into uint64."""
    a = np.minimum(src, dst).astype(np.uint64)
    b = np.maximum(src, dst).astype(np.uint64)
    return np.unique((a << np.uint64(32)) | b)


def bench_skewed(args, K):
    """The skewed graph with planted cliques: Chung-Lu with 2^skew_scale
    vertices, exponent 2.1, mean degree 16, largest expected degree 2^15, and
    64 planted cliques of 20-200 vertices (deep k).  Every variant of the
    reference's design space (3 orientations x COO+CSR / tiled x recount
    all / affected), tiles of 4096 vertices, finds the k-truss at --probe-k
    from scratch; their edge sets must agree with each other and with the
    edges whose truss number is at least --probe-k.  Then the learned
    per-phase pick decomposes the graph (the measured steps) beside one
    decomposition by the default variant (bit-exact)."""
    log(f"generating Chung-Lu 2^{args.skew_scale}, exponent 2.1, 64 cliques, seed {args.seed}")
    n, s, d = K.generate_chung_lu(1 << args.skew_scale, 2.1, 16.0, max_weight=float(1 << 15),
                                  num_cliques=64, clique_min=20, clique_max=200, seed=args.seed)
    g = K.Graph(n, s, d, pairs=True)
    log(f"graph: n={n} m={g.m} max degree={g.max_degree}")

    sweep = []
    ref = None
    for v in K.all_variants():
        v.tile_size = 4096
        p = K.Plan(g, v)
        K.device_synchronize()
        t0 = time.perf_counter()
        es, ed, st = p.ktruss(args.probe_k)
        K.device_synchronize()
        wall = time.perf_counter() - t0
        keys = undirected_keys(es, ed)
        if ref is None:
            ref = keys
        elif not np.array_equal(keys, ref):
            raise SystemExit(f"{v.describe()}: the {args.probe_k}-truss differs from the first variant's")
        sweep.append({"variant": v.describe(), "ms": round(1e3 * wall, 2),
                      "device_ms": round(st["ms_total"], 2), "iterations": st["iterations"],
                      "truss_edges": int(len(keys))})
        log(f"{v.describe()}: {args.probe_k}-truss of {len(keys)} edges in {1e3 * wall:.1f} ms "
            f"({st['iterations']} iterations)")
        del p

    import learn
    model, configs = learn.load_model()
    plan = K.Plan(g, K.default_variant())
    plan.set_candidates(configs)
    plan.set_pick(K.PICK_MODEL, model)
    plan.set_timing(False)
    for i in range(args.warmup):
        st = plan.decompose()
        log(f"warmup {i}: {st['ms_total']:.1f} ms device, k_max {st['k_max']}")
    K.device_synchronize()
    per_step = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = plan.decompose()
        per_step.append(st["ms_total"])
        log(f"step {i}: {st['ms_total']:.1f} ms device")
    K.device_synchronize()
    ms_per_step = 1e3 * (time.perf_counter() - t0) / max(args.steps, 1)
    tau = plan.truss_numbers()
    if not np.array_equal(undirected_keys(s[tau >= args.probe_k], d[tau >= args.probe_k]), ref):
        raise SystemExit(f"the edges of truss number >= {args.probe_k} are not the swept {args.probe_k}-truss")
    ph = plan.phase_log()
    names = [c.describe() for c in plan.candidates]
    picks = {}
    for c in ph[:, 2].astype(int):
        picks[names[c]] = picks.get(names[c], 0) + 1
    del plan

    default = None
    bit_exact = None
    if not args.skip_default:
        pd = K.Plan(g, K.default_variant())
        K.device_synchronize()
        t0 = time.perf_counter()
        sd = pd.decompose()
        K.device_synchronize()
        wall = time.perf_counter() - t0
        bit_exact = bool(np.array_equal(pd.truss_numbers(), tau))
        default = {"ms_per_step": 1e3 * wall, "device_ms": sd["ms_total"], "edges_per_s": g.m / wall,
                   "k_max": sd["k_max"], "iterations": sd["iterations"]}
        log(f"default variant: {1e3 * wall:.1f} ms, k_max {sd['k_max']}, bit-exact {bit_exact}")
        del pd
    out = {
        "metric": "full k-truss edges/s at 1 GPU, skewed graph with planted cliques "
                  "(learned pick vs default variant), bit-exact",
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
            "workload": (f"Chung-Lu 2^{args.skew_scale} vertices, exponent 2.1, mean degree 16, "
                         f"largest weight 2^15, 64 planted cliques of 20-200 vertices, seed {args.seed}: "
                         "full k-truss decomposition; every variant's k-truss at the probe k"),
            "n": n,
            "m": g.m,
            "k_max": st["k_max"],
            "iterations": st["iterations"],
            "median_device_ms": statistics.median(per_step) if per_step else None,
            "phase_picks": picks,
            "default_variant": default,
            "speedup_vs_default": (default["ms_per_step"] / ms_per_step) if default else None,
            "bit_exact_vs_default": bit_exact,
            "probe_k": args.probe_k,
            "variant_sweep": sweep,
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
