"""C ABI and host-side behaviour that needs no GPU: the library loads, exports
every symbol include/ktruss.h declares, validates input the way the
reference asserts it, reads and writes the reference's TSV format, and the
generators are deterministic.  The CLI's argument handling is checked too."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import ktruss_amd as K
import oracle as O
from conftest import ROOT, clique, sym

HEADER = os.path.join(ROOT, "include", "ktruss.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kt_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(K.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(K.EXPORTS) == names


def test_abi_version_and_status_strings():
    assert K.lib().kt_abi_version() == 1
    assert K.lib().kt_status_string(K.ERR_SELF_LOOP) == b"self-loop"
    assert K.lib().kt_status_string(0) == b"ok"


def test_default_variant_matches_reference_defaults():
    v = K.default_variant()
    assert v.describe() == "directed-by-degree/coocsr/all"  # config.h:192-195
    assert (v.num_edge_remove_iter, v.block_size, v.tiles_per_dim, v.num_parallel_tiles) == (0, 512, 8, 8)


@pytest.mark.parametrize("edges,n,code", [
    ([(0, 0)], 3, K.ERR_SELF_LOOP),
    ([(0, 5)], 3, K.ERR_RANGE),
])
def test_graph_rejects_bad_pairs(edges, n, code):
    s = np.array([e[0] for e in edges], np.uint32)
    d = np.array([e[1] for e in edges], np.uint32)
    with pytest.raises(K.KtError) as e:
        K.Graph(n, s, d, pairs=True)
    assert e.value.status == code


def test_graph_rejects_asymmetric_and_duplicates():
    with pytest.raises(K.KtError) as e:
        K.Graph(3, np.array([0, 1, 1]), np.array([1, 0, 2]))
    assert e.value.status == K.ERR_ASYMMETRIC
    with pytest.raises(K.KtError) as e:
        K.Graph(3, np.array([0, 1, 0, 1]), np.array([1, 0, 1, 0]))
    assert e.value.status == K.ERR_DUPLICATE
    with pytest.raises(K.KtError) as e:
        K.Graph(3, np.array([0, 1]), np.array([1, 0]), pairs=True)
    assert e.value.status == K.ERR_DUPLICATE
    with pytest.raises(K.KtError) as e:
        K.Graph(1 << 31, np.array([0]), np.array([1]), pairs=True)
    assert e.value.status == K.ERR_RANGE
    assert "below 2^31" in K.lib().kt_last_error().decode()


def test_graph_numbers_edges():
    s, d = sym(clique(range(5)))
    g = K.Graph(5, s, d)
    assert (g.n, g.m, g.m_input, g.max_degree) == (5, 10, 20, 4)
    g2 = K.Graph(5, d[::2], s[::2], pairs=True)
    assert (g2.m, g2.m_input) == (10, 10)


def test_tsv_matches_reference_format(tmp_path):
    s, d = sym([(0, 1), (1, 2), (0, 2), (2, 3)])
    p = str(tmp_path / "g.tsv")
    K.write_tsv(p, s, d)
    q = str(tmp_path / "o.tsv")
    O.write_tsv(q, s, d)
    assert open(p).read() == open(q).read()
    n, s2, d2 = K.read_tsv(p)
    assert n == 4 and np.array_equal(s2, s) and np.array_equal(d2, d)
    assert K.verify_edges(s, d, s2[::-1], d2[::-1])
    assert not K.verify_edges(s, d, s2, d2[::-1])


def test_tsv_errors(tmp_path):
    p = str(tmp_path / "loop.tsv")
    with open(p, "w") as f:
        f.write("2 1 1\n3 3 1\n")
    with pytest.raises(K.KtError) as e:
        K.read_tsv(p)
    assert e.value.status == K.ERR_SELF_LOOP
    with pytest.raises(K.KtError) as e:
        K.read_tsv(str(tmp_path / "missing.tsv"))
    assert e.value.status == K.ERR_IO
    p = str(tmp_path / "trunc.tsv")
    with open(p, "w") as f:
        f.write("2 1 1\n3 1\n")
    with pytest.raises(K.KtError) as e:
        K.read_tsv(p)
    assert e.value.status == K.ERR_IO


def test_rmat_generator_deterministic_and_clean():
    n, s, d = K.generate_rmat(12, 16, seed=5)
    n2, s2, d2 = K.generate_rmat(12, 16, seed=5)
    assert n == 4096 and np.array_equal(s, s2) and np.array_equal(d, d2)
    assert np.all(s < d)
    key = s.astype(np.uint64) << 32 | d
    assert np.all(np.diff(key.astype(np.int64)) > 0)  # sorted, unique
    assert 40000 < len(s) < 16 * 4096
    _, s3, _ = K.generate_rmat(12, 16, seed=6)
    assert len(s3) != len(s) or not np.array_equal(s3, s)
    deg = np.bincount(np.concatenate([s, d]), minlength=n)
    assert deg.max() > 20 * deg.mean()  # skewed (a = 0.57)


def test_rmat_thread_count_independent():
    env = dict(os.environ, KT_HOST_THREADS="1")
    code = ("import sys; sys.path.insert(0, %r); import ktruss_amd as K, hashlib; "
            "n, s, d = K.generate_rmat(13, 8, seed=2); "
            "print(hashlib.sha256(s.tobytes() + d.tobytes()).hexdigest())") % os.path.join(ROOT, "ktruss-explorer_amd")
    one = subprocess.run(["python", "-c", code], env=env, capture_output=True, text=True).stdout
    env["KT_HOST_THREADS"] = "7"
    seven = subprocess.run(["python", "-c", code], env=env, capture_output=True, text=True).stdout
    assert one and one == seven


def test_erdos_renyi_degree():
    n, s, d = K.generate_erdos_renyi(20000, 12.0, seed=3)
    assert abs(2 * len(s) / n - 12.0) < 0.2
    assert np.all(s < d)


def test_chung_lu_weights_and_cliques():
    n, s, d = K.generate_chung_lu(1 << 14, 2.1, 16.0, max_weight=1024.0, num_cliques=4,
                                  clique_min=30, clique_max=40, seed=9)
    deg = np.bincount(np.concatenate([s, d]), minlength=n)
    assert 10 < 2 * len(s) / n < 17
    assert 400 < deg.max() < 1400
    assert np.all(s < d)


def test_plan_without_device_fails_loudly():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    s, d = sym(clique(range(4)))
    g = K.Graph(4, s, d)
    with pytest.raises(K.KtError) as e:
        K.Plan(g)
    assert e.value.status == K.ERR_NO_DEVICE


def run_cli(*args):
    return subprocess.run([K.CLI_PATH, *args], capture_output=True, text=True, timeout=60)


def test_cli_help_and_bad_options():
    r = run_cli("-h")
    assert r.returncode == 0 and "-d <directedness>" in r.stderr
    r = run_cli("-d", "sideways")
    assert r.returncode == 2 and "Unrecognized -d option: sideways" in r.stderr
    r = run_cli("-f", "csr")
    assert r.returncode == 2 and "Unrecognized -f option" in r.stderr
    r = run_cli("-r", "some")
    assert r.returncode == 2
    r = run_cli("-p", "cpu")
    assert r.returncode == 2
    r = run_cli("-Q")
    assert r.returncode == 2 and "Unrecognized option" in r.stderr


def test_cli_missing_graph_is_an_error(tmp_path):
    r = run_cli("-g", str(tmp_path / "nothing.tsv"))
    assert r.returncode == 1 and "reading the graph" in r.stderr
