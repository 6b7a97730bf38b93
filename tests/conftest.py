import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "ktruss-explorer_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import ktruss_amd as K
        K.init(0)
        return True
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    # Tests marked gpu run only where a gfx950 device answers; elsewhere they
    # are deselected by -m "not gpu" and would fail loudly if forced.
    pass


def sym(edges):
    """Symmetric COO (both directions) from an edge list."""
    s, d = [], []
    for a, b in edges:
        s += [a, b]
        d += [b, a]
    return np.array(s, np.uint32), np.array(d, np.uint32)


def clique(vertices):
    vs = list(vertices)
    return [(vs[i], vs[j]) for i in range(len(vs)) for j in range(i + 1, len(vs))]


def known_answer_graph():
    """K4 on {0..3}, triangle {3,4,5}, pendant 5-6 (the 1-based graph
    K4{1..4} + triangle{4,5,6} + 6-7 on which the reference's CPU path gives
    3-truss = all but 6-7 and 4-truss = K4)."""
    return 7, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (3, 4), (3, 5), (4, 5), (5, 6)]


@pytest.fixture(scope="session")
def small_graphs():
    """Named small graphs as (n, src, dst) symmetric COO."""
    import networkx as nx
    out = {}
    n, e = known_answer_graph()
    out["known_answer"] = (n, *sym(e))
    out["K7"] = (7, *sym(clique(range(7))))
    g = nx.gnp_random_graph(80, 0.15, seed=7)
    out["gnp80"] = (80, *sym(list(g.edges())))
    g = nx.powerlaw_cluster_graph(300, 6, 0.6, seed=3)
    out["plc300"] = (300, *sym(list(g.edges())))
    # two cliques sharing a vertex plus a wheel and a path: several k levels
    e = clique(range(0, 9)) + clique(range(8, 14))
    e += [(14 + i, 14 + (i + 1) % 10) for i in range(10)] + [(24, 14 + i) for i in range(10)]
    e += [(25, 26), (26, 27)]
    out["mixed"] = (28, *sym(e))
    return out
