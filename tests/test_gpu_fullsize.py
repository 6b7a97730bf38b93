"""The full-size configurations on the GPU, through properties that do not
need the CPU oracle to finish: R-MAT scale 22 (the graph the learned pick is
evaluated on) and the skewed Chung-Lu 2^22 graph with 64 planted cliques at
4096-vertex tiles.  A k-truss is unique, so every variant's k-truss at k must
be the set of edges whose truss number is at least k; the decompositions of
different count strategies must give the same truss numbers."""
import numpy as np
import pytest

import ktruss_amd as K
import learn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    K.init(0)


def keys(src, dst):
    a = np.minimum(src, dst).astype(np.uint64)
    b = np.maximum(src, dst).astype(np.uint64)
    return np.unique((a << np.uint64(32)) | b)


def test_rmat22_learned_delta_and_default_agree():
    """R-MAT 22, edge factor 16, seed 1: the learned per-phase pick and fixed
    DELTA give the same truss numbers and k_max; the default variant's
    k-truss (the reference's method, one k from scratch) at k = 3, 40 and
    k_max is exactly the edges of truss number >= k."""
    n, s, d = K.generate_rmat(22, 16, seed=1)
    g = K.Graph(n, s, d, pairs=True)
    print(f"R-MAT 22: n={n} m={g.m}", flush=True)
    pd = K.Plan(g, K.make_variant(count_kernel="delta"))
    pd.set_timing(False)
    sd = pd.decompose()
    tau = pd.truss_numbers()
    del pd
    model, configs = learn.load_model()
    pl = K.Plan(g, K.default_variant())
    pl.set_candidates(configs)
    pl.set_pick(K.PICK_MODEL, model)
    pl.set_timing(False)
    sl = pl.decompose()
    assert np.array_equal(pl.truss_numbers(), tau)
    assert sl["k_max"] == sd["k_max"] == int(tau.max())
    assert sl["iterations"] == sd["iterations"]
    del pl
    # DELTA whose first count (row-staged at this size) has 64-entry slices:
    # the rows longer than 64 entries are searched in HBM, not in LDS
    pc = K.Plan(g, K.make_variant(count_kernel="delta", block_size=256, cap_lds=64))
    pc.set_timing(False)
    pc.decompose()
    assert np.array_equal(pc.truss_numbers(), tau)
    del pc
    p = K.Plan(g, K.default_variant())
    for k in (3, 40, int(tau.max())):
        es, ed, st = p.ktruss(k)
        print(f"default variant k={k}: {len(es) // 2} edges, {st['iterations']} iterations", flush=True)
        assert np.array_equal(keys(es, ed), keys(s[tau >= k], d[tau >= k])), k


def test_skewed_chunglu22_all_variants_4096_tiles():
    """Chung-Lu 2^22, exponent 2.1, mean degree 16, 64 planted cliques of
    20-200 vertices: all twelve variants at 4096-vertex tiles (1024 tiles per
    dimension, 8 classes) find the same 8-truss, equal to the edges of truss
    number >= 8 from a DELTA decomposition."""
    n, s, d = K.generate_chung_lu(1 << 22, 2.1, 16.0, max_weight=float(1 << 15), num_cliques=64,
                                  clique_min=20, clique_max=200, seed=1)
    g = K.Graph(n, s, d, pairs=True)
    pd = K.Plan(g, K.make_variant(count_kernel="delta"))
    pd.decompose()
    tau = pd.truss_numbers()
    del pd
    want = keys(s[tau >= 8], d[tau >= 8])
    assert len(want) > 0
    for v in K.all_variants():
        v.tile_size = 4096
        p = K.Plan(g, v)
        es, ed, st = p.ktruss(8)
        print(f"{v.describe()}: {len(keys(es, ed))} edges, {st['iterations']} iterations", flush=True)
        assert np.array_equal(keys(es, ed), want), v.describe()
        del p
