"""kt_graph_create on the device: inputs of 2^16 entries or more are checked
and numbered by kernels once the device is initialised.  Same errors as the
host checks (graph.cu:54 and the symmetric-pair rule of the input format),
reporting the lowest offending entry; valid inputs give the edges, entry ids
and degrees the oracle's decomposition agrees with."""
import re

import numpy as np
import pytest

import ktruss_amd as K
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    K.init(0)


def random_pairs(n, m, seed):
    """m distinct undirected pairs (a < b) over n vertices, shuffled."""
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n, 3 * m, dtype=np.int64)
    b = rng.integers(0, n, 3 * m, dtype=np.int64)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = np.unique(lo[lo != hi] * n + hi[lo != hi])
    key = rng.permutation(key)[:m]
    return (key // n).astype(np.uint32), (key % n).astype(np.uint32)


def symmetric(a, b, seed):
    s = np.concatenate([a, b])
    d = np.concatenate([b, a])
    p = np.random.default_rng(seed).permutation(len(s))
    return s[p].copy(), d[p].copy()


def reported_entry(err):
    m = re.search(r"entry (\d+) = ", str(err))
    assert m, str(err)
    return int(m.group(1))


def test_device_built_graph_matches_oracle():
    n = 1 << 14
    a, b = random_pairs(n, 1 << 16, 3)
    s, d = symmetric(a, b, 4)
    g = K.Graph(n, s, d)
    deg = np.bincount(np.concatenate([a, b]), minlength=n)
    assert (g.n, g.m, g.m_input, g.max_degree) == (n, len(a), len(s), int(deg.max()))
    p = K.Plan(g, K.default_variant())
    p.decompose()
    tau_o, _, _ = O.decompose(n, s, d, O.make_variant("directed-by-degree", "coocsr", "all", 0, 8),
                              O.SCHED_JUMP | O.SCHED_REUSE)
    assert np.array_equal(p.truss_numbers(), tau_o)
    # one entry per pair: the same edges
    gp = K.Graph(n, b, a, pairs=True)
    assert (gp.m, gp.m_input, gp.max_degree) == (len(a), len(a), int(deg.max()))
    pp = K.Plan(gp, K.default_variant())
    pp.decompose()
    tau_p, _, _ = O.decompose(n, np.concatenate([b, a]), np.concatenate([a, b]),
                              O.make_variant("directed-by-degree", "coocsr", "all", 0, 8),
                              O.SCHED_JUMP | O.SCHED_REUSE)
    assert np.array_equal(pp.truss_numbers(), tau_p[:len(a)])


def test_device_validation_reports_lowest_bad_entry():
    n = 1 << 14
    a, b = random_pairs(n, 1 << 16, 5)
    s, d = symmetric(a, b, 6)
    m = len(s)
    # range before a self-loop, and a self-loop before a range error
    for first, second, code in ((70000, 90000, K.ERR_RANGE), (50000, 80000, K.ERR_SELF_LOOP)):
        s2, d2 = s.copy(), d.copy()
        if code == K.ERR_RANGE:
            d2[first] = n
            s2[second] = d2[second]
        else:
            s2[first] = d2[first]
            s2[second] = n + 7
        with pytest.raises(K.KtError) as e:
            K.Graph(n, s2, d2)
        assert e.value.status == code
        assert reported_entry(e.value) == first
    # an entry without its reverse: the lone entry is reported
    keep = np.ones(m, bool)
    i = 40000
    rev = np.flatnonzero((s == d[i]) & (d == s[i]))
    assert len(rev) == 1
    keep[rev[0]] = False
    with pytest.raises(K.KtError) as e:
        K.Graph(n, s[keep], d[keep])
    assert e.value.status == K.ERR_ASYMMETRIC
    lone = i - (1 if rev[0] < i else 0)
    assert reported_entry(e.value) == lone
    # a pair given twice: the lowest entry of that edge's (min, max) copies
    j = 12345
    s3 = np.concatenate([s, [s[j], d[j]]]).astype(np.uint32)
    d3 = np.concatenate([d, [d[j], s[j]]]).astype(np.uint32)
    with pytest.raises(K.KtError) as e:
        K.Graph(n, s3, d3)
    assert e.value.status == K.ERR_DUPLICATE
    lo, hi = min(s[j], d[j]), max(s[j], d[j])
    expect = min(np.flatnonzero((s3 == lo) & (d3 == hi)))
    assert reported_entry(e.value) == expect
    # one entry per pair: both directions of one edge are a duplicate
    sp = np.concatenate([a, [b[100]]]).astype(np.uint32)
    dp = np.concatenate([b, [a[100]]]).astype(np.uint32)
    with pytest.raises(K.KtError) as e:
        K.Graph(n, sp, dp, pairs=True)
    assert e.value.status == K.ERR_DUPLICATE
    assert reported_entry(e.value) == 100
    # and the valid input still builds after the failures
    assert K.Graph(n, s, d).m == len(a)
