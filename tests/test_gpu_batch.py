"""GPU parity of the batched decomposition.

Many graphs decomposed in one batch must give, graph by graph, exactly the
truss numbers of the CPU oracle (and of the single-graph engine), whatever
the tuning, the schedule or the number of steps captured per replay."""
import numpy as np
import pytest

import ktruss_amd as K
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    K.init(0)


def _mixed_batch(count=24, seed=5):
    """Seeded mix of tiny hand-made graphs, R-MAT, Erdos-Renyi and Chung-Lu
    graphs with planted cliques, uneven in size."""
    from conftest import clique, known_answer_graph
    rng = np.random.default_rng(seed)
    out = []
    n, e = known_answer_graph()
    out.append((n, np.array([a for a, b in e], np.uint32), np.array([b for a, b in e], np.uint32)))
    out.append((3, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32)))  # path: no triangle
    e = clique(range(9))
    out.append((9, np.array([a for a, b in e], np.uint32), np.array([b for a, b in e], np.uint32)))
    for i in range(count - len(out)):
        kind = i % 3
        if kind == 0:
            n, s, d = K.generate_rmat(int(rng.integers(6, 11)), int(rng.choice([4, 8, 16])),
                                      seed=int(rng.integers(1, 1 << 30)))
        elif kind == 1:
            n, s, d = K.generate_erdos_renyi(int(rng.integers(50, 1500)), float(rng.uniform(3, 20)),
                                             int(rng.integers(1, 1 << 30)))
        else:
            n, s, d = K.generate_chung_lu(int(rng.integers(200, 2000)), 2.3, 8.0, 0.0,
                                          int(rng.integers(0, 4)), 5, 25, int(rng.integers(1, 1 << 30)))
        out.append((n, s, d))
    return out


@pytest.fixture(scope="module")
def batch_data():
    raw = _mixed_batch()
    graphs = [K.Graph(n, s, d, pairs=True) for n, s, d in raw]
    taus = []
    for n, s, d in raw:
        ss, dd = K.symmetrize(s, d)
        tau_o, st_o, tr_o = O.decompose(n, ss, dd)
        st_o = dict(st_o)
        st_o["k_last"] = int(tr_o[-1][0]) if len(tr_o) else 0
        taus.append((tau_o[: len(s)], st_o))
    return raw, graphs, taus


def _check_all(b, graphs, taus):
    for i, g in enumerate(graphs):
        tau = b.truss_numbers(i)
        assert np.array_equal(tau, taus[i][0]), f"graph {i}"


def test_batch_matches_oracle(batch_data):
    raw, graphs, taus = batch_data
    b = K.Batch(graphs)
    st = b.decompose()
    _check_all(b, graphs, taus)
    assert st["graphs"] == len(graphs)
    assert st["k_max"] == max(t[1]["k_max"] for t in taus)
    for i, (tau_o, st_o) in enumerate(taus):
        kmax, iters, phases, passes = b.graph_info(i)
        assert kmax == st_o["k_max"], i
        assert phases == st_o["phases"], i
        # the single-graph loop ends with one pass that finds nothing left,
        # except when the k == 3 pass removed the last edge
        assert iters == st_o["iterations"] - (0 if st_o["k_last"] == 3 else 1), i
    # every count pass recounts exactly the live entries it is given
    assert st["recount_edges"] == st["live_in_passes"]


def test_batch_matches_single_graph_engine(batch_data):
    raw, graphs, taus = batch_data
    b = K.Batch(graphs)
    b.decompose()
    for i in (0, 3, 7, len(graphs) - 1):
        p = K.Plan(graphs[i])
        p.decompose()
        assert np.array_equal(p.truss_numbers(), b.truss_numbers(i)), i


@pytest.mark.parametrize("sched", [0, 1, 2, 3])
def test_batch_schedules_agree(batch_data, sched):
    raw, graphs, taus = batch_data
    b = K.Batch(graphs, K.make_variant(schedule=sched))
    b.decompose()
    _check_all(b, graphs, taus)


@pytest.mark.parametrize("tuning", [
    dict(block_size=256, chunk=64, cap_lds=64, blocks_per_cu=1, mark_chunk=64, steps_per_replay=1),
    dict(block_size=1024, chunk=512, cap_lds=4096, blocks_per_cu=8, mark_chunk=4096, steps_per_replay=7),
    dict(block_size=512, chunk=128, cap_lds=256, blocks_per_cu=3, mark_chunk=256, steps_per_replay=64),
], ids=["small", "large", "mixed"])
def test_batch_tunings_agree(batch_data, tuning):
    raw, graphs, taus = batch_data
    b = K.Batch(graphs)
    b.set_tuning(**tuning)
    st = b.decompose()
    _check_all(b, graphs, taus)
    # a second run on the same batch starts again from the pristine layout
    st2 = b.decompose()
    _check_all(b, graphs, taus)
    for key in ("steps", "iterations", "phases", "recount_edges", "merge_elems", "k_max"):
        assert st[key] == st2[key], key


def test_batch_by_index_orientation(batch_data):
    raw, graphs, taus = batch_data
    b = K.Batch(graphs, K.make_variant("directed-by-index"))
    b.decompose()
    _check_all(b, graphs, taus)


def test_batch_autotune_keeps_results(batch_data):
    raw, graphs, taus = batch_data
    b = K.Batch(graphs)
    best = b.autotune(trial_steps=3, repeats=1)
    log = b.tune_log()
    assert len(log) >= 2
    assert min(ms for _, ms in log) > 0
    assert best == b.tuning()
    b.decompose()
    _check_all(b, graphs, taus)


def test_batch_single_graph_and_rejects_other_variants():
    n, s, d = K.generate_rmat(9, 8, seed=11)
    g = K.Graph(n, s, d, pairs=True)
    b = K.Batch([g])
    b.decompose()
    p = K.Plan(g)
    p.decompose()
    assert np.array_equal(b.truss_numbers(0), p.truss_numbers())
    for v in (K.make_variant("undirected"), K.make_variant(fmt="tiled-coocsr"),
              K.make_variant(recount="affected")):
        with pytest.raises(K.KtError) as ei:
            K.Batch([g], v)
        assert ei.value.status == K.ERR_ARG
    with pytest.raises(K.KtError):
        K.Batch([], None)


def test_batch_timing_mode_same_results(batch_data):
    raw, graphs, taus = batch_data
    b = K.Batch(graphs)
    st_graph = b.decompose()
    b.set_timing(True)
    st = b.decompose()
    _check_all(b, graphs, taus)
    assert st["ms_count"] > 0
    assert st["count_launches"] >= 1
    for key in ("steps", "iterations", "phases", "recount_edges", "merge_elems"):
        assert st[key] == st_graph[key], key


@pytest.mark.parametrize("d", ["directed-by-degree", "directed-by-index"])
def test_batch_delta_matches_oracle(batch_data, d):
    raw, graphs, taus = batch_data
    b = K.Batch(graphs, K.make_variant(d, count_kernel="delta"))
    st = b.decompose()
    _check_all(b, graphs, taus)
    # every triangle of every graph is removed exactly once
    assert st["delta_triangles"] == sum(t[1]["triangles"] for t in taus)
    for i, (tau_o, st_o) in enumerate(taus):
        kmax, iters, phases, passes = b.graph_info(i)
        assert kmax == st_o["k_max"] and phases == st_o["phases"], i
        assert passes == 1, i  # the first full count only; DELTA keeps the rest exact
    st2 = b.decompose()  # a second run restores the adjacency flags
    _check_all(b, graphs, taus)
    assert st2["delta_triangles"] == st["delta_triangles"]


def test_batch_delta_tunings_and_timing(batch_data):
    raw, graphs, taus = batch_data
    b = K.Batch(graphs, K.make_variant(count_kernel="delta"))
    b.set_tuning(block_size=256, chunk=64, cap_lds=64, blocks_per_cu=1, mark_chunk=64,
                 steps_per_replay=3)
    b.decompose()
    _check_all(b, graphs, taus)
    b.set_timing(True)
    st = b.decompose()
    _check_all(b, graphs, taus)
    assert st["ms_count"] > 0
    best = b.autotune(trial_steps=3, repeats=1)
    assert best == b.tuning()
    b.set_timing(False)
    b.decompose()
    _check_all(b, graphs, taus)
