"""Learned per-phase pick of the support-count configuration.

The peel's results do not depend on how supports are counted, only its speed
does, and the fastest count configuration changes as the graph shrinks from
a sparse skewed graph to dense cores.  This module

  collect()   decomposes seeded synthetic graphs in SWEEP mode: every count
              pass runs every candidate configuration and is timed, and each
              phase yields one sample (16 features, per-candidate count ms);
  train()     fits the cost model (16 -> 64 -> 64 -> C MLP, trained on the GPU
              by the engine's own forward/backward/Adam kernels) to predict
              each candidate's log cost relative to the others;
  evaluate()  decomposes a held-out graph with the fixed default variant, with
              the model picking per phase, with single candidates for the whole
              run, and in SWEEP mode (the exhaustive per-phase judge), and
              reports speed-up, regret and parity.

Command line:
  python learn.py collect OUT.npz [--graphs N] [--seed S] [--max-scale X]
  python learn.py train DATA.npz [DATA.npz ...] [--out MODEL.npz] [--epochs E]
  python learn.py eval [--model MODEL.npz] [--scale 20] [--sweep] [--singles I ...]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import ktruss_amd as K  # noqa: E402

MODEL_PATH = os.path.join(HERE, "picker", "count_picker.npz")
HELDOUT_PATH = os.path.join(HERE, "picker", "heldout_phases.npz")
# graphs held out of training: index % HOLDOUT_MOD == HOLDOUT_REM in the seeded list
HOLDOUT_MOD, HOLDOUT_REM = 7, 3
MODEL_BIN_PATH = os.path.join(HERE, "picker", "count_picker.ktm")

# Candidate support configurations; index 0 is the default variant's own.
# "delta" maintains the supports across deletions instead of recounting
# (its first full count uses the ROW kernel with the same block and slice).
CANDIDATES = [
    ("row", 512, 0),
    ("row", 256, 1024),
    ("row", 1024, 512),
    ("row", 128, 2048),
    ("search", 256, 1024),
    ("merge", 256, 1024),
    ("delta", 512, 0),
]


def candidate_configs():
    return [K.count_config(k, b, c) for k, b, c in CANDIDATES]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# graphs

def graph_specs(count, seed, max_scale=17, min_scale=10):
    """Seeded mix of R-MAT, Chung-Lu (with planted cliques) and Erdos-Renyi
    graphs, sizes log-uniform in [2^min_scale, 2^max_scale] vertices."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        kind = ("rmat", "cl", "er")[i % 3]
        scale = int(rng.integers(min_scale, max_scale + 1))
        gseed = int(rng.integers(1, 1 << 30))
        if kind == "rmat":
            a = float(rng.uniform(0.45, 0.6))
            bc = (1 - a) * float(rng.uniform(0.35, 0.45))
            specs.append({"kind": "rmat", "scale": scale, "ef": int(rng.choice([8, 16, 24])),
                          "a": a, "b": bc, "c": bc, "seed": gseed})
        elif kind == "cl":
            specs.append({"kind": "cl", "n": 1 << scale, "gamma": float(rng.uniform(2.0, 2.8)),
                          "deg": float(rng.uniform(8, 32)), "maxw": float(2 ** rng.uniform(8, 13)),
                          "cliques": int(rng.integers(0, 9)), "seed": gseed})
        else:
            specs.append({"kind": "er", "n": 1 << scale, "deg": float(2 ** rng.uniform(3, 6)),
                          "seed": gseed})
    return specs


def graph_specs_c5(count, seed, min_scale=12, max_scale=20):
    """The training set of the learned pick: seeded R-MAT and Chung-Lu graphs
    (with planted cliques), alternating, scales uniform in [min_scale,
    max_scale]."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        scale = int(rng.integers(min_scale, max_scale + 1))
        gseed = int(rng.integers(1, 1 << 30))
        if i % 2 == 0:
            a = float(rng.uniform(0.45, 0.6))
            bc = (1 - a) * float(rng.uniform(0.35, 0.45))
            specs.append({"kind": "rmat", "scale": scale, "ef": int(rng.choice([8, 16, 24])),
                          "a": a, "b": bc, "c": bc, "seed": gseed})
        else:
            specs.append({"kind": "cl", "n": 1 << scale, "gamma": float(rng.uniform(2.0, 2.8)),
                          "deg": float(rng.uniform(8, 32)), "maxw": float(2 ** rng.uniform(8, 14)),
                          "cliques": int(rng.integers(0, 17)), "seed": gseed, "scale": scale})
    return specs


def sweep_every_for(scale):
    """Phases between timed sweeps: every phase up to scale 15, then every
    2nd, 4th, 8th as a full sweep of the larger graphs grows too costly."""
    return 1 if scale <= 15 else 2 if scale <= 17 else 4 if scale <= 18 else 8


def make_graph(spec):
    if spec["kind"] == "rmat":
        n, s, d = K.generate_rmat(spec["scale"], spec["ef"], spec["a"], spec["b"], spec["c"],
                                  spec["seed"])
    elif spec["kind"] == "cl":
        n, s, d = K.generate_chung_lu(spec["n"], spec["gamma"], spec["deg"], spec["maxw"],
                                      spec["cliques"], 20, 200, spec["seed"])
    else:
        n, s, d = K.generate_erdos_renyi(spec["n"], spec["deg"], spec["seed"])
    return K.Graph(n, s, d, pairs=True)


# ---------------------------------------------------------------------------
# data

def sweep(graph, variant=None, every=1):
    """SWEEP decomposition (every candidate timed in every `every`-th phase,
    the others run by DELTA): (truss numbers, phase log)."""
    p = K.Plan(graph, variant or K.default_variant())
    p.set_candidates(candidate_configs())
    p.set_pick(K.PICK_SWEEP)
    if every > 1:
        p.set_sweep_sampling(every, len(CANDIDATES) - 1, 0)
    p.decompose()
    return p.truss_numbers(), p.phase_log()


def tuner_sample(graph, k=3):
    """The autotuner's log as a training sample: kt_plan_autotune first times
    every candidate's count on the whole graph (the phase that opens the
    decomposition); their costs, or None when one is missing."""
    p = K.Plan(graph, K.default_variant())
    p.set_candidates(candidate_configs())
    p.autotune(k=k, repeats=1)
    lg = p.tune_log()[1:1 + len(CANDIDATES)]
    ms = np.array([t["ms"] for t in lg], np.float32)
    return ms if len(ms) == len(CANDIDATES) and np.all(ms > 0) else None


def tuner_samples(count, seed, min_scale, max_scale, start, stop):
    """Autotuner samples (features of the opening phase from a sweep of it
    alone) of graphs [start, stop) of the seeded C5 list."""
    feats, costs, meta = [], [], []
    specs = graph_specs_c5(count, seed, min_scale, max_scale)
    t0 = time.time()
    for i in range(start, min(stop, len(specs))):
        g = make_graph(specs[i])
        _, ph = sweep(g, every=1 << 30)
        ts = tuner_sample(g)
        if len(ph) and ts is not None:
            feats.append(ph[0, 4:20])
            costs.append(ts)
            meta.append((i, -1, 0))
        log(f"[{i + 1}/{len(specs)}] tuner sample, m={g.m}, elapsed={time.time() - t0:.0f}s")
    return np.array(feats, np.float32), np.array(costs, np.float32), np.array(meta, np.float32)


def collect(count, seed, max_scale=17, min_scale=10, c5=False, start=0, stop=None):
    """Samples of graphs [start, stop) of the seeded list; meta holds (graph
    index, k, iterations) per sample."""
    feats, costs, meta = [], [], []
    specs = (graph_specs_c5(count, seed, min_scale, max_scale) if c5
             else graph_specs(count, seed, max_scale, min_scale))
    t0 = time.time()
    for i in range(start, len(specs) if stop is None else min(stop, len(specs))):
        spec = specs[i]
        g = make_graph(spec)
        every = sweep_every_for(spec.get("scale", 0)) if c5 else 1
        _, ph = sweep(g, every=every)
        C = len(CANDIDATES)
        for r in ph:
            ms = r[20:20 + C]
            if np.all(ms > 0):
                feats.append(r[4:20])
                costs.append(ms)
                meta.append((i, r[0], r[1]))
        if c5 and len(ph):
            ts = tuner_sample(g)
            if ts is not None:
                feats.append(ph[0, 4:20])
                costs.append(ts)
                meta.append((i, -1, 0))  # k = -1: the autotuner's sample
        log(f"[{i + 1}/{len(specs)}] {spec['kind']} scale={spec.get('scale', '-')} m={g.m} phases={len(ph)} "
            f"samples={len(feats)} elapsed={time.time() - t0:.0f}s")
    return (np.array(feats, np.float32), np.array(costs, np.float32),
            np.array(meta, np.float32), specs)


def targets(costs):
    """Per-candidate log2 cost relative to the sample's mean log2 cost."""
    lc = np.log2(np.maximum(costs, 1e-4))
    return (lc - lc.mean(1, keepdims=True)).astype(np.float32)


def weights(costs):
    """Samples weighted by the square root of their default-variant cost, so
    the phases that dominate the run dominate the fit without a handful of
    huge phases drowning the rest."""
    w = np.sqrt(np.maximum(costs[:, 0], 1e-4))
    return (w / w.mean()).astype(np.float32)


def train(feats, costs, epochs=300, batch=256, lr=2e-3, seed=1):
    """Fit the cost model; its inference inputs are then clamped to the range
    of the training features, so graphs beyond that range (larger, deeper)
    are priced as the most extreme graphs it was trained on."""
    C = costs.shape[1]
    m = K.Mlp(C, seed=seed)
    t = targets(costs)
    w = weights(costs)
    loss = m.train(feats, t, None, w, epochs=epochs, batch=min(batch, len(feats)), lr=lr, seed=seed)
    m.set_input_range(feats.min(0), feats.max(0))
    return m, loss


def picks_regret(model, feats, costs):
    """On a sample set: cost of the model's picks vs the per-phase best and
    vs candidate 0 (the default)."""
    pred = model.forward(feats)
    choice = pred.argmin(1)
    picked = costs[np.arange(len(costs)), choice].sum()
    best = costs.min(1).sum()
    default = costs[:, 0].sum()
    single = costs.sum(0).min()
    return {"picked_ms": float(picked), "best_ms": float(best), "default_ms": float(default),
            "best_single_ms": float(single), "regret_vs_best": float(picked / best - 1),
            "speedup_vs_default": float(default / picked)}


def save_model(model, path=MODEL_PATH, info=None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    cands = np.array([[K.COUNT_KERNELS[k], b, c] for k, b, c in CANDIDATES], np.uint32)
    lo, hi = model.get_input_range()
    np.savez(path, params=model.get_params(), candidates=cands, lo=lo, hi=hi,
             info=np.array(json.dumps(info or {})))


def export_model_bin(npz_path=MODEL_PATH, out_path=MODEL_BIN_PATH):
    """The model as the ktruss command line reads it (-M), little-endian:
    "KTMODEL1", u32 candidates C, u32 parameter count P, C x (kernel, block,
    cap, chunk) u32, P float32 parameters, then the KT_MLP_IN lower and upper
    input bounds as float32."""
    with np.load(npz_path, allow_pickle=False) as z:
        params = z["params"].astype("<f4")
        cands = np.asarray(z["candidates"]).astype("<u4")
        lo = z["lo"].astype("<f4")
        hi = z["hi"].astype("<f4")
    c4 = np.zeros((len(cands), 4), "<u4")
    c4[:, :cands.shape[1]] = cands
    with open(out_path, "wb") as f:
        f.write(b"KTMODEL1")
        f.write(np.array([len(cands), len(params)], "<u4").tobytes())
        f.write(c4.tobytes())
        f.write(params.tobytes())
        f.write(lo.tobytes())
        f.write(hi.tobytes())


def load_model(path=MODEL_PATH):
    with np.load(path, allow_pickle=False) as z:
        params = z["params"]
        cands = z["candidates"]
        lo = z["lo"] if "lo" in z.files else None
        hi = z["hi"] if "hi" in z.files else None
    configs = [K.count_config(int(k), int(b), int(c)) for k, b, c in cands]
    m = K.Mlp(len(configs))
    m.set_params(params)
    if lo is not None:
        m.set_input_range(lo, hi)
    return m, configs


def evaluate(graph, model, configs, sweep=False, singles=(), default_ms=None):
    """The fixed default variant, the model's per-phase pick, each candidate
    in `singles` used alone for the whole run, and (sweep) the exhaustive
    per-phase best; wall-clock milliseconds of full decompositions."""
    def timed(plan):
        K.device_synchronize()
        t0 = time.perf_counter()
        st = plan.decompose()
        K.device_synchronize()
        return 1e3 * (time.perf_counter() - t0), st

    out = {"m": graph.m}
    log(f"evaluating on m={graph.m}")
    if default_ms is None:
        p = K.Plan(graph, K.default_variant())
        out["default_ms"], st = timed(p)
        tau_default = p.truss_numbers()
        out["k_max"] = st["k_max"]
        del p
    else:  # a recorded time of the default variant; parity against the first single
        out["default_ms"] = default_ms
        out["default_recorded"] = True
        cfg = configs[singles[0]]
        q = K.Plan(graph, K.make_variant(count_kernel=cfg.count_kernel, block_size=cfg.block_size,
                                         cap_lds=cfg.cap_lds, chunk=cfg.chunk))
        q.decompose()
        tau_default = q.truss_numbers()
        del q
    p = K.Plan(graph, K.default_variant())
    p.set_candidates(configs)
    p.set_pick(K.PICK_MODEL, model)
    timed(p)  # warm
    out["model_ms"], st = timed(p)
    log(f"model: {out['model_ms']:.1f} ms")
    out["model_support_ms"] = st["ms_count"] + st["ms_delta"]
    out["model_parity"] = bool(np.array_equal(p.truss_numbers(), tau_default))
    ph = p.phase_log()
    model_choice = ph[:, 2].astype(int)
    out["model_picks"] = np.bincount(model_choice, minlength=len(configs)).tolist()
    singles_ms = {}
    for c in singles:
        cfg = configs[c]
        v = K.make_variant(count_kernel=cfg.count_kernel, block_size=cfg.block_size,
                           cap_lds=cfg.cap_lds, chunk=cfg.chunk)
        q = K.Plan(graph, v)
        timed(q)
        ms, _ = timed(q)
        singles_ms[cfg.describe()] = ms
        log(f"single {cfg.describe()}: {ms:.1f} ms")
        assert np.array_equal(q.truss_numbers(), tau_default), cfg.describe()
        del q
    out["singles_ms"] = singles_ms
    out["speedup_vs_default"] = out["default_ms"] / out["model_ms"]
    if singles_ms:
        best = min(singles_ms.values())
        out["best_single_ms"] = best
        out["regret_vs_best_single"] = out["model_ms"] / best - 1
    if sweep:
        # the exhaustive per-phase sweep, as `sweep` sampled runs (phases of
        # index r modulo sweep in run r) merged; every run peels the same phases
        p.set_pick(K.PICK_SWEEP)
        ms = None
        parity = True
        for r in range(sweep):
            p.set_sweep_sampling(sweep, len(configs) - 1, r)
            p.decompose()
            parity = parity and bool(np.array_equal(p.truss_numbers(), tau_default))
            ph = p.phase_log()
            if ms is None:
                ms = np.full((len(ph), len(configs)), -1.0)
            rows = np.arange(len(ph)) % sweep == r
            ms[rows] = ph[rows, 20:20 + len(configs)]
            log(f"sweep run {r + 1}/{sweep}: {int(rows.sum())} phases timed")
        out["sweep_parity"] = parity
        ms = np.maximum(ms, 0.0)  # a phase without a support pass costs nothing
        out["phases"] = int(len(ms))
        out["phases_best_per_candidate"] = np.bincount(ms.argmin(1), minlength=len(configs)).tolist()
        if len(model_choice) == len(ms):  # the same peel: the model's picks priced by the sweep
            out["model_on_sweep_ms"] = float(ms[np.arange(len(ms)), model_choice].sum())
        out["exhaustive_support_ms"] = float(ms.min(1).sum())
        out["best_single_support_ms"] = float(ms.sum(0).min())
        out["best_single_candidate"] = configs[int(ms.sum(0).argmin())].describe()
        if "model_on_sweep_ms" in out:
            out["regret_vs_exhaustive"] = out["model_on_sweep_ms"] / out["exhaustive_support_ms"] - 1
    return out


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("out")
    c.add_argument("--graphs", type=int, default=300)
    c.add_argument("--seed", type=int, default=7)
    c.add_argument("--max-scale", type=int, default=17)
    c.add_argument("--min-scale", type=int, default=10)
    c.add_argument("--c5", action="store_true",
                   help="R-MAT / Chung-Lu graphs, scales uniform in [min, max], sampled sweeps above scale 15")
    c.add_argument("--start", type=int, default=0)
    c.add_argument("--stop", type=int, default=None)
    u = sub.add_parser("tune", help="autotuner samples of the C5 graphs [start, stop)")
    u.add_argument("out")
    u.add_argument("--graphs", type=int, default=2000)
    u.add_argument("--seed", type=int, default=2026)
    u.add_argument("--min-scale", type=int, default=12)
    u.add_argument("--max-scale", type=int, default=20)
    u.add_argument("--start", type=int, default=0)
    u.add_argument("--stop", type=int, default=2000)
    t = sub.add_parser("train")
    t.add_argument("data", nargs="+")
    t.add_argument("--out", default=MODEL_PATH)
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--holdout-by-graph", action="store_true",
                   help="validate on whole held-out graphs instead of a random 15%% of the phases")
    t.add_argument("--heldout-out", default=None, help="write the held-out phases here (npz)")
    t.add_argument("--data-out", default=None, help="write all samples here (npz)")
    e = sub.add_parser("eval")
    e.add_argument("--model", default=MODEL_PATH)
    e.add_argument("--scale", type=int, default=20)
    e.add_argument("--sweep", type=int, default=0,
                   help="exhaustive per-phase sweep in this many sampled runs (0: none)")
    e.add_argument("--singles", type=int, nargs="*", default=[])
    e.add_argument("--default-ms", type=float, default=None,
                   help="use this recorded default-variant time instead of running it")
    args = ap.parse_args()
    K.init(0)
    if args.cmd == "collect":
        f, c_, meta, specs = collect(args.graphs, args.seed, args.max_scale, args.min_scale, args.c5,
                                     args.start, args.stop)
        np.savez(args.out, feats=f, costs=c_, meta=meta, specs=np.array(json.dumps(specs)))
        log(f"wrote {len(f)} samples to {args.out}")
    elif args.cmd == "tune":
        f, c_, meta = tuner_samples(args.graphs, args.seed, args.min_scale, args.max_scale, args.start, args.stop)
        np.savez(args.out, feats=f, costs=c_, meta=meta)
        log(f"wrote {len(f)} tuner samples to {args.out}")
    elif args.cmd == "train":
        fs, cs, ms = [], [], []
        for path in args.data:
            with np.load(path, allow_pickle=False) as z:
                fs.append(z["feats"])
                cs.append(z["costs"])
                ms.append(z["meta"] if "meta" in z.files else np.full((len(z["feats"]), 3), -2, np.float32))
        f, c_, meta = np.concatenate(fs), np.concatenate(cs), np.concatenate(ms)
        if args.holdout_by_graph:
            # whole graphs held out (every HOLDOUT_MOD-th of the seeded list):
            # the validation phases come from graphs the model never saw
            held = (meta[:, 0].astype(np.int64) % HOLDOUT_MOD) == HOLDOUT_REM
            tr = np.flatnonzero(~held)
            va = np.flatnonzero(held & (meta[:, 1] >= 0))  # the sweeps' phases
        else:
            rng = np.random.default_rng(0)
            idx = rng.permutation(len(f))
            cut = int(0.85 * len(f))
            tr, va = idx[:cut], idx[cut:]
        model, loss = train(f[tr], c_[tr], epochs=args.epochs)
        info = {"train_samples": int(len(tr)), "train_graphs": int(len(np.unique(meta[tr, 0]))),
                "tuner_samples": int(np.sum(meta[tr, 1] == -1)), "loss": loss,
                "train": picks_regret(model, f[tr], c_[tr]),
                "validation": picks_regret(model, f[va], c_[va]),
                "validation_graphs": int(len(np.unique(meta[va, 0]))),
                "candidates": [c.describe() for c in candidate_configs()]}
        save_model(model, args.out, info)
        if args.heldout_out:
            np.savez_compressed(args.heldout_out, feats=f[va], costs=c_[va], meta=meta[va])
        if args.data_out:
            np.savez_compressed(args.data_out, feats=f, costs=c_, meta=meta)
        print(json.dumps(info))
    else:
        model, configs = load_model(args.model)
        n, s, d = K.generate_rmat(args.scale, 16, seed=1)
        g = K.Graph(n, s, d, pairs=True)
        res = evaluate(g, model, configs, args.sweep, args.singles, args.default_ms)
        res["scale"] = args.scale
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
