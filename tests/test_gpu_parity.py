"""GPU parity: every HIP variant against the CPU oracle, bit-exact.

Integer work, so the bar is exact equality of truss numbers, k-truss edge
sets, per-iteration deletion traces and the algorithmic-byte counters."""
import numpy as np
import pytest

import ktruss_amd as K
import oracle as O

pytestmark = pytest.mark.gpu

VARIANTS = [(d, f, r) for d in ("undirected", "directed-by-index", "directed-by-degree")
            for f in ("coocsr", "tiled-coocsr") for r in ("all", "affected")]


def vid(v):
    return "/".join(v)


def oracle_variant(d, f, r, e=0, tiles=3):
    return O.make_variant(d, f, r, e, tiles)


def gpu_variant(d, f, r, e=0, tiles=3, y=2):
    return K.make_variant(d, f, r, e_iters=e, tiles=tiles, parallel_tiles=y)


@pytest.fixture(scope="module", autouse=True)
def _device():
    K.init(0)


def edge_set(src, dst):
    return set(zip(src.tolist(), dst.tolist()))


@pytest.mark.parametrize("v", VARIANTS, ids=vid)
def test_decompose_matches_oracle(small_graphs, v):
    for name, (n, s, d) in small_graphs.items():
        tau_o, st_o, tr_o = O.decompose(n, s, d, oracle_variant(*v), O.SCHED_JUMP | O.SCHED_REUSE)
        g = K.Graph(n, s, d)
        p = K.Plan(g, gpu_variant(*v))
        st = p.decompose()
        tau = p.truss_numbers()
        assert np.array_equal(tau, tau_o), name
        assert st["k_max"] == st_o["k_max"], name
        tr = [(k, it, dl, rc) for k, it, dl, rc in p.trace()]
        assert tr == [tuple(x) for x in tr_o], name
        for key in ("count_passes", "recount_edges", "merge_elems", "live_in_passes",
                    "compactions", "compact_elems", "iterations", "phases"):
            assert st[key] == st_o[key], (name, key)


@pytest.mark.parametrize("v", VARIANTS, ids=vid)
def test_ktruss_matches_oracle(small_graphs, v):
    for name, (n, s, d) in small_graphs.items():
        g = K.Graph(n, s, d)
        p = K.Plan(g, gpu_variant(*v))
        for k in (3, 4, 5, 7):
            os_, od_, st_o, tr_o = O.ktruss(n, s, d, k, oracle_variant(*v))
            gs, gd, st = p.ktruss(k)
            assert edge_set(gs, gd) == edge_set(os_, od_), (name, k)
            assert len(gs) == len(os_), (name, k)
            assert [tuple(x) for x in tr_o] == [tuple(x) for x in p.trace()], (name, k)
            for key in ("recount_edges", "merge_elems", "live_in_passes", "compact_elems"):
                assert st[key] == st_o[key], (name, k, key)


@pytest.mark.parametrize("e_iters", [1, 3, 100])
@pytest.mark.parametrize("v", [("directed-by-degree", "coocsr", "affected"),
                               ("undirected", "tiled-coocsr", "all"),
                               ("directed-by-index", "tiled-coocsr", "affected")], ids=vid)
def test_edge_removal_iterations(small_graphs, v, e_iters):
    n, s, d = small_graphs["plc300"]
    tau_o, st_o, tr_o = O.decompose(n, s, d, oracle_variant(*v, e=e_iters), 3)
    p = K.Plan(K.Graph(n, s, d), gpu_variant(*v, e=e_iters))
    st = p.decompose()
    assert np.array_equal(p.truss_numbers(), tau_o)
    assert st["compactions"] == st_o["compactions"]
    assert st["compact_elems"] == st_o["compact_elems"]
    assert st["merge_elems"] == st_o["merge_elems"]


@pytest.mark.parametrize("sched", [0, 1, 2, 3])
def test_schedules_agree(small_graphs, sched):
    n, s, d = small_graphs["plc300"]
    tau_o, st_o, tr_o = O.decompose(n, s, d, oracle_variant("directed-by-degree", "coocsr", "all"),
                                    sched)
    p = K.Plan(K.Graph(n, s, d), K.make_variant(schedule=sched))
    st = p.decompose()
    assert np.array_equal(p.truss_numbers(), tau_o)
    assert st["count_passes"] == st_o["count_passes"]
    assert st["phases"] == st_o["phases"]


def test_rmat12_all_variants_agree():
    n, s, d = K.generate_rmat(12, 16, seed=3)
    ss, dd = K.symmetrize(s, d)
    tau_o, st_o, _ = O.decompose(n, ss, dd)
    g = K.Graph(n, s, d, pairs=True)
    for v in K.all_variants():
        p = K.Plan(g, v)
        p.decompose()
        tau = p.truss_numbers()
        assert np.array_equal(tau, tau_o[: len(s)]), v.describe()


def test_skewed_graph_with_cliques_all_variants_4096_tiles():
    """The skewed-graph sweep's shape at a size the oracle finishes in
    seconds: Chung-Lu, exponent 2.1, planted cliques of 20-60 vertices (deep
    k: a clique of c vertices has truss number c, so the largest is 20 or
    more), 4096-vertex tiles (4 per dimension); every variant and DELTA give
    the oracle's truss numbers."""
    n, s, d = K.generate_chung_lu(1 << 14, 2.1, 16.0, max_weight=2048.0, num_cliques=8,
                                  clique_min=20, clique_max=60, seed=5)
    ss, dd = K.symmetrize(s, d)
    tau_o, _, _ = O.decompose(n, ss, dd)
    assert tau_o.max() >= 20
    g = K.Graph(n, s, d, pairs=True)
    variants = K.all_variants()
    for v in variants:
        v.tile_size = 4096
    variants.append(K.make_variant(count_kernel="delta"))
    for v in variants:
        p = K.Plan(g, v)
        p.decompose()
        assert np.array_equal(p.truss_numbers(), tau_o[: len(s)]), v.describe()


def _golden(scale):
    import hashlib
    import json
    import os
    from conftest import GOLDEN
    path = os.path.join(GOLDEN, f"rmat{scale}_tau.json")
    if not os.path.exists(path):
        pytest.skip(f"fixture rmat{scale}_tau.json not generated")
    with open(path) as f:
        return json.load(f), hashlib


@pytest.mark.parametrize("scale", [14, 16, 18, 20])
def test_rmat_matches_golden_fixture(scale):
    gold, hashlib = _golden(scale)
    n, s, d = K.generate_rmat(scale, 16, seed=1)
    h = hashlib.sha256(np.ascontiguousarray(s, "<u4").tobytes() +
                       np.ascontiguousarray(d, "<u4").tobytes()).hexdigest()
    assert h == gold["pairs_sha256"], "generator output changed"
    g = K.Graph(n, s, d, pairs=True)
    variants = [K.default_variant(), K.make_variant("directed-by-degree", "coocsr", "affected"),
                K.make_variant("directed-by-degree", "tiled-coocsr", "all"),
                K.make_variant(count_kernel="delta")]
    if scale >= 20:  # the default variant (11 s), DELTA and the learned pick below
        variants = [variants[0], variants[3]]
    if scale <= 16:
        variants += [K.make_variant("undirected", "coocsr", "all"),
                     K.make_variant("directed-by-index", "coocsr", "affected", e_iters=4)]
    for v in variants:
        p = K.Plan(g, v)
        st = p.decompose()
        tau = p.truss_numbers()
        assert hashlib.sha256(np.ascontiguousarray(tau, "<u4").tobytes()).hexdigest() == \
            gold["tau_sha256"], v.describe()
        assert st["k_max"] == gold["k_max"]
        if v.count_kernel == K.COUNT_DELTA:
            # every triangle is destroyed exactly once over the decomposition
            # (a lost or doubled task list entry would miss or repeat some)
            assert st["delta_triangles"] == gold["triangles"], v.describe()
    # the committed cost model's per-phase pick
    import learn
    model, configs = learn.load_model()
    p = K.Plan(g, K.default_variant())
    p.set_candidates(configs)
    p.set_pick(K.PICK_MODEL, model)
    p.decompose()
    assert hashlib.sha256(np.ascontiguousarray(p.truss_numbers(), "<u4").tobytes()).hexdigest() == \
        gold["tau_sha256"], "learned pick"


TINY = {"edgeless": (6, []), "no_vertices": (0, []), "one_edge": (10, [(3, 7)]),
        "triangle_high_ids": (50, [(40, 45), (45, 49), (40, 49), (0, 49)])}


@pytest.mark.parametrize("v", VARIANTS + [("directed-by-degree", "coocsr", "all", "delta")],
                         ids=lambda v: "/".join(v))
def test_empty_and_tiny_graphs_match_oracle(v):
    """Graphs with no entries, no vertices, one edge, and a triangle on the
    highest ids: every variant (and DELTA) gives the oracle's truss numbers,
    k_max and k-truss edge sets."""
    d, f, r = v[:3]
    for name, (n, e) in TINY.items():
        s = np.array([a for a, b in e] + [b for a, b in e], np.uint32)
        t = np.array([b for a, b in e] + [a for a, b in e], np.uint32)
        g = K.Graph(n, s, t)
        gv = gpu_variant(d, f, r)
        if len(v) > 3:
            gv = K.make_variant(d, f, r, count_kernel=K.COUNT_KERNELS["delta"])
        p = K.Plan(g, gv)
        st = p.decompose()
        tau_o, st_o, _ = O.decompose(n, s, t, oracle_variant(d, f, r), O.SCHED_JUMP | O.SCHED_REUSE)
        assert np.array_equal(p.truss_numbers(), tau_o), name
        assert st["k_max"] == st_o["k_max"], name
        for k in (2, 3, 4):
            os_, od_, _, _ = O.ktruss(n, s, t, k, oracle_variant(d, f, r))
            gs, gd, _ = p.ktruss(k)
            assert edge_set(gs, gd) == edge_set(os_, od_), (name, k)


@pytest.mark.parametrize("v", [x for x in VARIANTS if x[1] == "tiled-coocsr"], ids=vid)
def test_many_small_tiles_match_oracle(small_graphs, v):
    """Tiles of a few vertices (tiles_per_dim = n / 4) and 8 parallel tile
    classes: a row meets many tiles of each class and its block rows are a
    few entries long, so items take the register, lane-merge and wave paths;
    truss numbers, traces and the byte counters equal the oracle's."""
    for name in ("plc300", "gnp80", "mixed"):
        n, s, d = small_graphs[name]
        tiles = max(1, n // 4)
        tau_o, st_o, tr_o = O.decompose(n, s, d, oracle_variant(*v, tiles=tiles), O.SCHED_JUMP | O.SCHED_REUSE)
        p = K.Plan(K.Graph(n, s, d), gpu_variant(*v, tiles=tiles, y=8))
        st = p.decompose()
        assert np.array_equal(p.truss_numbers(), tau_o), name
        assert [tuple(x) for x in tr_o] == [(k, it, dl, rc) for k, it, dl, rc in p.trace()], name
        for key in ("count_passes", "recount_edges", "merge_elems", "live_in_passes"):
            assert st[key] == st_o[key], (name, key)


@pytest.mark.parametrize("v", [x for x in VARIANTS if x[1] == "tiled-coocsr"], ids=vid)
@pytest.mark.parametrize("tiles,y", [(3, 1), (4, 4), (6, 4)])
def test_tile_classes_match_oracle(small_graphs, v, tiles, y):
    """Block layouts with one destination class (each source tile one block),
    with as many classes as tiles (the reference's own (tile, tile) blocks,
    graph.h:67-76) and with tiles grouped into fewer classes: truss numbers,
    traces and the byte counters equal the oracle's."""
    for name in ("plc300", "mixed", "known_answer"):
        n, s, d = small_graphs[name]
        tau_o, st_o, tr_o = O.decompose(n, s, d, oracle_variant(*v, tiles=tiles), O.SCHED_JUMP | O.SCHED_REUSE)
        p = K.Plan(K.Graph(n, s, d), gpu_variant(*v, tiles=tiles, y=y))
        st = p.decompose()
        assert np.array_equal(p.truss_numbers(), tau_o), name
        assert [tuple(x) for x in tr_o] == [(k, it, dl, rc) for k, it, dl, rc in p.trace()], name
        for key in ("count_passes", "recount_edges", "merge_elems", "live_in_passes", "compact_elems"):
            assert st[key] == st_o[key], (name, key)
        for k in (3, 5):
            os_, od_, _, _ = O.ktruss(n, s, d, k, oracle_variant(*v, e=1, tiles=tiles))
            q = K.Plan(K.Graph(n, s, d), gpu_variant(*v, e=1, tiles=tiles, y=y))
            gs, gd, _ = q.ktruss(k)
            assert edge_set(gs, gd) == edge_set(os_, od_), (name, k)


@pytest.mark.parametrize("v", [("directed-by-degree", "coocsr", "all"), ("directed-by-index", "tiled-coocsr", "all"),
                               ("undirected", "tiled-coocsr", "affected"), ("undirected", "coocsr", "all")], ids=vid)
def test_autotune_keeps_results(small_graphs, v):
    """kt_plan_autotune tries tile shapes (tiled: tile size, classes, workgroup
    size, the layout rebuilt per shape) or count shapes (COO+CSR), applies the
    fastest, and the decomposition still equals the oracle's."""
    n, s, d = small_graphs["plc300"]
    tau_o, st_o, _ = O.decompose(n, s, d, oracle_variant(*v), O.SCHED_JUMP | O.SCHED_REUSE)
    p = K.Plan(K.Graph(n, s, d), gpu_variant(*v))
    best = p.autotune(k=3, repeats=1)
    log = p.tune_log()
    assert len(log) >= 5 and all(t["ms"] > 0 for t in log)
    if v[1] == "tiled-coocsr":
        assert len({(t["tile_size"], t["num_parallel_tiles"]) for t in log}) >= 3
        assert (best.tile_size, best.num_parallel_tiles) in {(t["tile_size"], t["num_parallel_tiles"]) for t in log}
    assert min(t["ms"] for t in log) <= log[0]["ms"]
    p.decompose()
    assert np.array_equal(p.truss_numbers(), tau_o)
    es, ed, _ = p.ktruss(4)
    os_, od_, _, _ = O.ktruss(n, s, d, 4, oracle_variant(*v))
    assert edge_set(es, ed) == edge_set(os_, od_)
