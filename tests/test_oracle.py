"""Pins the CPU oracle before it is trusted as the checker.

Each test is chosen so a plausible slip in the restatement (a dropped credit,
a wrong affected state, an off-by-one in k-2, a tiled segment skipped) fails:
closed forms, brute force, networkx k_truss, the reference's recorded
known-answer outputs, invariants and exact work counters on K_n."""
import json
import os

import networkx as nx
import numpy as np
import pytest
from hypothesis import given, settings, strategies as hst

import oracle as O
from conftest import GOLDEN, clique, sym

ALL12 = [(d, f, r) for d in ("undirected", "directed-by-index", "directed-by-degree")
         for f in ("coocsr", "tiled-coocsr") for r in ("all", "affected")]


def V(d, f, r, e=0, t=3):
    return O.make_variant(d, f, r, e, t)


def per_edge(tau):
    return tau[::2]


def edge_set_1based(src, dst):
    return {(int(a) + 1, int(b) + 1) for a, b in zip(src, dst)}


@pytest.mark.parametrize("v", ALL12, ids=["/".join(x) for x in ALL12])
def test_known_answer_recorded_from_reference(v):
    with open(os.path.join(GOLDEN, "known_answer.json")) as f:
        ka = json.load(f)
    edges = [(a - 1, b - 1) for a, b in ka["graph_1based"]]
    s, d = sym(edges)
    for k, expect in ka["truss_1based"].items():
        os_, od_, _, _ = O.ktruss(7, s, d, int(k), V(*v, t=2))
        want = set()
        for a, b in expect:
            want |= {(a, b), (b, a)}
        assert edge_set_1based(os_, od_) == want, k


@pytest.mark.parametrize("n", [3, 4, 5, 8, 12])
def test_clique_closed_form(n):
    s, d = sym(clique(range(n)))
    for v in ALL12:
        tau, st, _ = O.decompose(n, s, d, V(*v))
        assert set(tau.tolist()) == {n}
        assert st["k_max"] == n


@pytest.mark.parametrize("n", [5, 7, 10])
def test_clique_minus_edge_and_wheel(n):
    e = [x for x in clique(range(n)) if x != (0, 1)]
    s, d = sym(e)
    tau, _, _ = O.decompose(n, s, d)
    assert set(tau.tolist()) == {n - 1}
    rim = [(i, (i + 1) % n) for i in range(n)]
    spokes = [(n, i) for i in range(n)]
    s, d = sym(rim + spokes)
    tau, _, _ = O.decompose(n + 1, s, d)
    assert set(tau.tolist()) == {3}


def test_cycles_trees_and_unions():
    cyc = [(i, (i + 1) % 9) for i in range(9)]
    tree = [(9 + i, 9 + (i - 1) // 2) for i in range(1, 12)]
    k5 = [(21 + a, 21 + b) for a, b in clique(range(5))]
    k7 = [(26 + a, 26 + b) for a, b in clique(range(7))]
    shared = [(32 + a, 32 + b) for a, b in clique(range(4))]  # shares vertex 32 with k7
    edges = cyc + tree + k5 + k7 + shared
    s, d = sym(edges)
    tau = per_edge(O.decompose(36, s, d)[0])
    want = [2] * (len(cyc) + len(tree)) + [5] * len(k5) + [7] * len(k7) + [4] * len(shared)
    assert tau.tolist() == want


def test_planted_clique_lower_bound():
    g = nx.gnp_random_graph(200, 0.03, seed=11)
    members = list(range(0, 200, 10))[:14]
    for a, b in clique(members):
        g.add_edge(a, b)
    edges = list(g.edges())
    s, d = sym(edges)
    tau = per_edge(O.decompose(200, s, d)[0])
    for (a, b), t in zip(edges, tau):
        if a in members and b in members:
            assert t >= len(members)


def brute(n, edges):
    s, d = sym(edges)
    return per_edge(O.bruteforce_tau(n, s, d))


@settings(max_examples=40, deadline=None)
@given(hst.integers(4, 26), hst.floats(0.1, 0.7), hst.integers(0, 10_000))
def test_bruteforce_agrees(n, p, seed):
    g = nx.gnp_random_graph(n, p, seed=seed)
    edges = list(g.edges())
    if not edges:
        return
    s, d = sym(edges)
    want = brute(n, edges)
    for v in [ALL12[0], ALL12[5], ALL12[9], ALL12[10]]:
        for sched in (0, 3):
            tau = per_edge(O.decompose(n, s, d, V(*v), sched)[0])
            assert np.array_equal(tau, want), (v, sched)


@pytest.mark.parametrize("name,graph", [
    ("gnp90", lambda: nx.gnp_random_graph(90, 0.12, seed=5)),
    ("plc250", lambda: nx.powerlaw_cluster_graph(250, 5, 0.7, seed=2)),
    ("caveman", lambda: nx.relaxed_caveman_graph(8, 9, 0.15, seed=4)),
])
def test_networkx_k_truss(name, graph):
    g = graph()
    g.remove_edges_from(list(nx.selfloop_edges(g)))
    n = g.number_of_nodes()
    edges = list(g.edges())
    s, d = sym(edges)
    tau, st, _ = O.decompose(n, s, d)
    tau = per_edge(tau)
    for k in range(3, st["k_max"] + 2):
        ref = {frozenset(e) for e in nx.k_truss(g, k).edges()}
        mine = {frozenset(e) for e, t in zip(edges, tau) if t >= k}
        assert ref == mine, k
    for v in ALL12:
        for k in (3, 4, 6):
            os_, od_, _, _ = O.ktruss(n, s, d, k, V(*v))
            mine = {frozenset((int(a), int(b))) for a, b in zip(os_, od_)}
            ref = {frozenset(e) for e in nx.k_truss(g, k).edges()}
            assert mine == ref, (v, k)


def test_truss_invariants():
    g = nx.powerlaw_cluster_graph(400, 6, 0.5, seed=9)
    n = g.number_of_nodes()
    edges = list(g.edges())
    s, d = sym(edges)
    tau, st, _ = O.decompose(n, s, d)
    tau = per_edge(tau)
    assert st["triangles"] == sum(nx.triangles(g).values()) // 3
    A = np.zeros((n, n), np.int64)
    for k in range(3, st["k_max"] + 1):
        A[:] = 0
        for (a, b), t in zip(edges, tau):
            if t >= k:
                A[a, b] = A[b, a] = 1
        common = A @ A
        for (a, b), t in zip(edges, tau):
            if t >= k:
                assert common[a, b] >= k - 2  # every edge of T_k has k-2 triangles in T_k


def test_traces_identical_across_variants():
    g = nx.powerlaw_cluster_graph(300, 6, 0.6, seed=3)
    s, d = sym(list(g.edges()))
    ref = None
    for v in ALL12:
        for e in (0, 2):
            _, _, tr = O.decompose(300, s, d, V(*v, e=e), 3)
            dels = [(k, it, dl) for k, it, dl, _ in tr]
            ref = ref or dels
            assert dels == ref, (v, e)


@pytest.mark.parametrize("n", [4, 6, 9])
def test_work_counters_closed_form_on_clique(n):
    """K_n, directed by degree (equal degrees -> identity relabel), row u has
    n-1-u entries.  Plain schedule: phases k = 3..n each count once and delete
    nothing; phase n+1 deletes everything, counts once more (no live entry),
    then compacts.  Sum over oriented pairs of L_u + L_v = n (n-1)^2 / 2."""
    s, d = sym(clique(range(n)))
    m = n * (n - 1) // 2
    per_pass = n * (n - 1) ** 2 // 2
    _, st, _ = O.decompose(n, s, d, V("directed-by-degree", "coocsr", "all"), 0)
    assert st["count_passes"] == n
    assert st["phases"] == n - 1
    assert st["recount_edges"] == (n - 1) * m
    assert st["merge_elems"] == (n - 1) * per_pass
    assert st["live_in_passes"] == (n - 1) * m
    assert st["compactions"] == 1 and st["compact_elems"] == m
    # JUMP: one exact count, jump to k = n+1, delete all, one empty pass
    _, st, _ = O.decompose(n, s, d, V("directed-by-degree", "coocsr", "all"), 1)
    assert st["count_passes"] == 2 and st["phases"] == 1
    assert st["merge_elems"] == per_pass
    assert O.balg_bytes(st) == 4 * per_pass + 24 * m + 12 * m + 16 * m


def test_undirected_work_counters_closed_form():
    """K_n undirected: 2m = n(n-1) entries, every row n-1 long.  Capped counts
    give no jump, so phases k = 3..n each count once (L_u+L_v = 2(n-1) per
    entry) and phase n+1 deletes everything then counts nothing."""
    n = 6
    s, d = sym(clique(range(n)))
    _, st, tr = O.decompose(n, s, d, V("undirected", "coocsr", "all"), 3)
    assert st["count_passes"] == n
    assert st["recount_edges"] == (n - 1) * n * (n - 1)
    assert st["merge_elems"] == 2 * n * (n - 1) ** 3
    assert [x[2] for x in tr if x[2]] == [n * (n - 1) // 2]


def test_tsv_roundtrip_and_verify(tmp_path):
    edges = [(0, 1), (1, 2), (0, 2), (2, 3)]
    s, d = sym(edges)
    p = str(tmp_path / "g.tsv")
    O.write_tsv(p, s, d)
    lines = open(p).read().splitlines()
    assert lines[0] == f"{d[0] + 1}\t{s[0] + 1}\t1"  # "dst src 1" (graph.cu:99-101)
    n, s2, d2 = O.read_tsv(p)
    assert n == 4 and np.array_equal(s, s2) and np.array_equal(d, d2)
    assert O.verify(4, s, d, s2[::-1], d2[::-1])
    assert not O.verify(4, s, d, s2[:-1], d2[:-1])


def test_rejects_bad_input(tmp_path):
    with pytest.raises(O.OracleError) as e:
        O.decompose(3, np.array([0, 1]), np.array([0, 1]))
    assert e.value.code == -2
    with pytest.raises(O.OracleError) as e:
        O.decompose(3, np.array([0]), np.array([1]))
    assert e.value.code == -3
    with pytest.raises(O.OracleError) as e:
        O.decompose(3, np.array([0, 1, 0, 1]), np.array([1, 0, 1, 0]))
    assert e.value.code == -4
    p = str(tmp_path / "loop.tsv")
    with open(p, "w") as f:
        f.write("2\t1\t1\n3\t3\t1\n")
    with pytest.raises(O.OracleError) as e:
        O.read_tsv(p)
    assert e.value.code == -2


def test_golden_rmat14_fixture():
    path = os.path.join(GOLDEN, "rmat14_tau.json")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    with open(path) as f:
        gold = json.load(f)
    hist = {int(k): v for k, v in gold["tau_histogram"].items()}
    assert sum(hist.values()) == gold["m"]
    assert max(hist) == gold["k_max"]
