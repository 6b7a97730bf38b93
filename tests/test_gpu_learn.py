"""The learned count-kernel picker on the GPU: MLP kernels against the CPU
oracle (float64), training, and bit-exact parity of SWEEP / MODEL runs."""
import numpy as np
import pytest

import ktruss_amd as K
import mlp_ref as R

pytestmark = pytest.mark.gpu

RTOL, ATOL = 2e-4, 2e-5  # FP32 kernels vs float64 oracle


@pytest.fixture(scope="module", autouse=True)
def _device():
    K.init(0)


def _data(C, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 16), dtype=np.float32)
    t = rng.normal(size=(n, C)).astype(np.float32)
    m = (rng.random((n, C)) > 0.2).astype(np.float32)
    w = rng.uniform(0.2, 2.0, n).astype(np.float32)
    return x, t, m, w


@pytest.mark.parametrize("C", [1, 5, 8, 16])
def test_forward_matches_cpu(C):
    m = K.Mlp(C, seed=3)
    x, _, _, _ = _data(C, 77, 1)
    y = m.forward(x)
    ref = R.forward(m.get_params(), C, x)
    np.testing.assert_allclose(y, ref, rtol=RTOL, atol=ATOL)


@pytest.mark.parametrize("C,n", [(8, 1), (8, 300), (3, 1000)])
def test_gradient_matches_cpu(C, n):
    m = K.Mlp(C, seed=5)
    x, t, mask, w = _data(C, n, 2)
    loss, g = m.loss_grad(x, t, mask, w)
    lref, gref = R.loss_grad(m.get_params(), C, x, t, mask, w)
    assert abs(loss - lref) <= 1e-4 * max(1.0, abs(lref))
    np.testing.assert_allclose(g, gref, rtol=1e-3, atol=1e-5)
    # deterministic reduction: the same call gives the same bits
    loss2, g2 = m.loss_grad(x, t, mask, w)
    assert loss2 == loss and np.array_equal(g, g2)


def test_training_fits_a_cost_table():
    C = 6
    rng = np.random.default_rng(4)
    x = rng.random((2000, 16), dtype=np.float32)
    W = rng.normal(size=(16, C))
    t = np.tanh(x @ W).astype(np.float32)
    m = K.Mlp(C, seed=9)
    before = m.loss_grad(x, t)[0]
    after = m.train(x, t, epochs=60, batch=128, lr=3e-3, seed=2)
    assert after < 0.2 * before
    pick = m.forward(x).argmin(1)
    assert np.mean(pick == t.argmin(1)) > 0.6


def _graph(scale=11, seed=3):
    n, s, d = K.generate_rmat(scale, 16, seed=seed)
    return K.Graph(n, s, d, pairs=True)


def test_sweep_and_model_runs_are_bit_exact():
    g = _graph()
    p = K.Plan(g, K.default_variant())
    p.decompose()
    ref = p.truss_numbers()
    cands = [K.count_config("row", 512), K.count_config("row", 256, 512),
             K.count_config("search", 256, 1024), K.count_config("merge", 256, 1024)]
    p.set_candidates(cands)
    p.set_pick(K.PICK_SWEEP)
    p.decompose()
    assert np.array_equal(p.truss_numbers(), ref)
    ph = p.phase_log()
    assert ph.shape[1] == K.PHASE_REC and len(ph) > 3
    assert np.all(ph[:, 20:24] > 0) and np.all(ph[:, 24:] == -1)
    assert np.all((ph[:, 2] >= 0) & (ph[:, 2] < 4))
    assert np.all(ph[:, 4 + 15] >= 0)
    model = K.Mlp(4, seed=11)
    p.set_pick(K.PICK_MODEL, model)
    p.decompose()
    assert np.array_equal(p.truss_numbers(), ref)
    ph = p.phase_log()
    # the device pick equals the host forward pass on the logged features
    host = model.forward(ph[:, 4:20]).argmin(1)
    assert np.array_equal(ph[:, 2].astype(int), host)


@pytest.mark.parametrize("every,offset", [(2, 0), (3, 0), (3, 2)])
def test_sampled_sweep_is_bit_exact(every, offset):
    """Sampled SWEEP: every candidate is timed only in the phases of index
    offset modulo every; the others run the fill candidate (DELTA) alone.
    Truss numbers equal the default variant's either way."""
    g = _graph()
    p = K.Plan(g, K.default_variant())
    p.decompose()
    ref = p.truss_numbers()
    cands = [K.count_config("row", 512), K.count_config("search", 256, 1024), K.count_config("delta", 512)]
    p.set_candidates(cands)
    p.set_pick(K.PICK_SWEEP)
    p.set_sweep_sampling(every, 2, offset)
    p.decompose()
    assert np.array_equal(p.truss_numbers(), ref)
    ph = p.phase_log()
    assert len(ph) > 2 * every
    swept = np.arange(len(ph)) % every == offset
    assert np.all(ph[swept][:, 20:23] >= 0) and np.any(np.all(ph[swept][:, 20:23] > 0, axis=1))
    assert np.all(ph[~swept][:, 20:23] == -1) and np.all(ph[~swept][:, 2] == -1)


def _cliques(sizes):
    """Disjoint cliques: every edge of K_s has truss number s."""
    src, dst, tau = [], [], []
    base = 0
    for s in sizes:
        i, j = np.triu_indices(s, 1)
        src.append(i + base)
        dst.append(j + base)
        tau.append(np.full(len(i), s, np.uint32))
        base += s
    return base, np.concatenate(src).astype(np.uint32), np.concatenate(dst).astype(np.uint32), np.concatenate(tau)


@pytest.mark.parametrize("every,offset", [(1 << 30, 0), (258, 257)])
def test_sampled_sweep_beyond_256_phases(every, offset):
    """A sampled SWEEP over more phases than the phase-feature buffer's first
    size (256 rows): phases without a pick do not grow it, so the run's end
    copies only the rows it holds, and a swept phase beyond it (phase 257,
    the only one swept) grows it past its row.  Disjoint cliques of 3 .. 262
    vertices peel in one phase per clique size (JUMP), and every edge of K_s
    has truss number s."""
    n, s, d, tau = _cliques(range(3, 263))
    g = K.Graph(n, s, d, pairs=True)
    p = K.Plan(g, K.default_variant())
    p.set_candidates([K.count_config("row", 512), K.count_config("delta", 512)])
    p.set_pick(K.PICK_SWEEP)
    p.set_sweep_sampling(every, 1, offset)
    p.decompose()
    assert np.array_equal(p.truss_numbers(), tau)
    ph = p.phase_log()
    assert len(ph) > 258
    swept = np.arange(len(ph)) % every == offset
    assert swept.sum() == 1
    assert np.all(ph[swept][:, 20:22] > 0) and np.all(ph[swept][:, 4 + 15] >= 0)
    assert np.all(ph[~swept][:, 2] == -1) and np.all(ph[~swept][:, 4:22] == -1)


def test_candidate_validation():
    p = K.Plan(_graph(8), K.default_variant())
    with pytest.raises(K.KtError):
        p.set_candidates([K.count_config("row", 100)])
    with pytest.raises(K.KtError):
        p.set_pick(K.PICK_SWEEP)  # no candidates yet
    p.set_candidates([K.count_config("row", 256)])
    with pytest.raises(K.KtError):
        p.set_pick(K.PICK_MODEL, K.Mlp(3))  # one output per candidate


def test_input_range_clamps_inference():
    m = K.Mlp(3, seed=4)
    rng = np.random.default_rng(2)
    x = rng.uniform(-2, 2, size=(32, K.MLP_IN)).astype(np.float32)
    lo = np.full(K.MLP_IN, -0.5, np.float32)
    hi = np.full(K.MLP_IN, 0.75, np.float32)
    y_clamped_input = m.forward(np.clip(x, lo, hi))
    m.set_input_range(lo, hi)
    got_lo, got_hi = m.get_input_range()
    assert np.array_equal(got_lo, lo) and np.array_equal(got_hi, hi)
    assert np.array_equal(m.forward(x), y_clamped_input)
    # the reference network sees the clamped input too
    import mlp_ref as R
    ref = R.forward(m.get_params(), 3, np.clip(x, lo, hi))
    assert np.allclose(m.forward(x), ref, rtol=1e-4, atol=1e-4)
