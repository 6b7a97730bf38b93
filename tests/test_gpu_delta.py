"""GPU parity of DELTA support maintenance.

DELTA replaces recounts by removing the triangles destroyed by each mark
pass from the surviving supports.  The supports it leaves are those of a
full recount, so the per-iteration deletion trace, every truss number and
every k-truss must equal the CPU oracle's; and since every triangle is
destroyed exactly once over a full decomposition, the triangles DELTA
removes add up to the graph's triangle count."""
import numpy as np
import pytest

import ktruss_amd as K
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    K.init(0)


def delta_variant(d="directed-by-degree", **kw):
    return K.make_variant(d, "coocsr", "all", count_kernel="delta", **kw)


@pytest.mark.parametrize("d", ["directed-by-degree", "directed-by-index"])
def test_delta_decompose_matches_oracle(small_graphs, d):
    for name, (n, s, dd) in small_graphs.items():
        tau_o, st_o, tr_o = O.decompose(n, s, dd, O.make_variant(d, "coocsr", "all"),
                                        O.SCHED_JUMP | O.SCHED_REUSE)
        p = K.Plan(K.Graph(n, s, dd), delta_variant(d))
        st = p.decompose()
        assert np.array_equal(p.truss_numbers(), tau_o), name
        assert st["k_max"] == st_o["k_max"], name
        assert [t[:3] for t in p.trace()] == [tuple(x[:3]) for x in tr_o], name
        assert st["delta_triangles"] == st_o["triangles"], name
        assert st["delta_passes"] > 0 or st_o["triangles"] == 0, name


@pytest.mark.parametrize("sched", [0, 1, 2, 3])
def test_delta_schedules(small_graphs, sched):
    n, s, d = small_graphs["plc300"]
    tau_o, st_o, tr_o = O.decompose(n, s, d, O.make_variant("directed-by-degree", "coocsr", "all"),
                                    sched)
    p = K.Plan(K.Graph(n, s, d), delta_variant(schedule=sched))
    p.decompose()
    assert np.array_equal(p.truss_numbers(), tau_o)
    assert [t[:3] for t in p.trace()] == [tuple(x[:3]) for x in tr_o]


def test_delta_ktruss_matches_oracle(small_graphs):
    for name, (n, s, d) in small_graphs.items():
        p = K.Plan(K.Graph(n, s, d), delta_variant())
        for k in (3, 4, 5, 7):
            os_, od_, st_o, tr_o = O.ktruss(n, s, d, k, O.make_variant())
            gs, gd, st = p.ktruss(k)
            assert set(zip(gs.tolist(), gd.tolist())) == set(zip(os_.tolist(), od_.tolist())), (name, k)
            assert len(gs) == len(os_), (name, k)


def test_delta_rmat12_and_repeat():
    n, s, d = K.generate_rmat(12, 16, seed=3)
    ss, dd = K.symmetrize(s, d)
    tau_o, st_o, tr_o = O.decompose(n, ss, dd)
    g = K.Graph(n, s, d, pairs=True)
    p = K.Plan(g, delta_variant())
    for _ in range(2):  # a second run restores the adjacency flags
        st = p.decompose()
        assert np.array_equal(p.truss_numbers(), tau_o[: len(s)])
        assert st["delta_triangles"] == st_o["triangles"]
        assert [t[:3] for t in p.trace()] == [tuple(x[:3]) for x in tr_o]
        assert st["ms_delta"] > 0 and st["balg_delta_bytes"] >= 4 * st["delta_elems"]
        assert st["count_launches"] == 1  # the first full count only


def test_delta_sweep_and_model_pick(small_graphs):
    n, s, d = small_graphs["plc300"]
    tau_o, _, _ = O.decompose(n, s, d)
    g = K.Graph(n, s, d)
    cands = [K.count_config("row", 512, 0), K.count_config("delta", 256, 0),
             K.count_config("row", 256, 256), K.count_config("delta", 512, 128)]
    p = K.Plan(g)
    p.set_candidates(cands)
    p.set_pick(K.PICK_SWEEP)
    p.decompose()
    assert np.array_equal(p.truss_numbers(), tau_o)
    ph = p.phase_log()
    assert (ph[:, 20:24] >= 0).all()
    m = K.Mlp(len(cands), seed=5)
    p.set_pick(K.PICK_MODEL, m)
    p.decompose()
    assert np.array_equal(p.truss_numbers(), tau_o)
    picks = set(int(x) for x in p.phase_log()[:, 2])
    assert picks <= set(range(len(cands)))


def test_delta_rejected_where_unsupported(small_graphs):
    n, s, d = small_graphs["gnp80"]
    g = K.Graph(n, s, d)
    for v in (K.make_variant("undirected", count_kernel="delta"),
              K.make_variant(fmt="tiled-coocsr", count_kernel="delta"),
              K.make_variant(recount="affected", count_kernel="delta")):
        with pytest.raises(K.KtError) as ei:
            K.Plan(g, v)
        assert ei.value.status == K.ERR_ARG


def test_delta_timing_off_same_results():
    """Without per-pass events (the host polls the deletion counts of the
    iterations it runs ahead) the peel is the same."""
    n, s, d = K.generate_rmat(13, 16, seed=7)
    g = K.Graph(n, s, d, pairs=True)
    p = K.Plan(g, delta_variant())
    st_on = p.decompose()
    tau_on, tr_on = p.truss_numbers(), [t[:3] for t in p.trace()]
    p.set_timing(False)
    for _ in range(2):
        st_off = p.decompose()
        assert np.array_equal(p.truss_numbers(), tau_on)
        assert [t[:3] for t in p.trace()] == tr_on
        assert st_off["delta_triangles"] == st_on["delta_triangles"]
        assert st_off["delta_passes"] == st_on["delta_passes"]
        assert st_off["ms_delta"] <= st_on["ms_delta"]
    p.set_timing(True)
    assert p.decompose()["ms_delta"] > 0
