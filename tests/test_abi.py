"""C ABI and host-side behaviour that needs no GPU: the library loads, exports
every symbol include/ktruss.h declares, validates input the way the
reference asserts it, reads and writes the reference's TSV format, and the
generators are deterministic.  The CLI's argument handling is checked too."""
import ctypes
import os
import re
import struct
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


def big_symmetric(n=1 << 16, m=1 << 17, seed=7):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n, m).astype(np.uint32)
    b = rng.integers(0, n, m).astype(np.uint32)
    b[a == b] = (b[a == b] + 1) % n
    return np.concatenate([a, b]), np.concatenate([b, a])


def test_tsv_large_file_matches_oracle_io(tmp_path):
    """Files of several MiB are parsed and formatted in parallel chunks; the
    bytes and the parsed entries equal the oracle's serial reader/writer."""
    s, d = big_symmetric()
    p, q = str(tmp_path / "k.tsv"), str(tmp_path / "o.tsv")
    K.write_tsv(p, s, d)
    O.write_tsv(q, s, d)
    assert os.path.getsize(p) > (1 << 21)
    assert open(p, "rb").read() == open(q, "rb").read()
    n, s2, d2 = K.read_tsv(p)
    no, so, do = O.read_tsv(p)
    assert n == no == int(max(s.max(), d.max())) + 1
    assert np.array_equal(s2, so) and np.array_equal(d2, do)
    assert np.array_equal(s2, s) and np.array_equal(d2, d)
    # mixed whitespace and no final newline parse the same
    text = open(p).read().replace("\t", "  ", 1000).rstrip("\n")
    open(p, "w").write(text)
    n3, s3, d3 = K.read_tsv(p)
    assert n3 == n and np.array_equal(s3, s) and np.array_equal(d3, d)


def test_tsv_large_file_errors_are_the_first_in_the_file(tmp_path):
    s, d = big_symmetric()
    lines = [f"{b + 1}\t{a + 1}\t1\n" for a, b in zip(s.tolist(), d.tolist())]
    cases = []
    loop = list(lines)
    loop[150000] = "5\t5\t1\n"
    loop[200000] = "0\t3\t1\n"
    cases.append((loop, K.ERR_SELF_LOOP, "itself"))
    zero = list(lines)
    zero[100000] = "0\t3\t1\n"
    zero[200000] = "5\t5\t1\n"
    cases.append((zero, K.ERR_RANGE, "1-based"))
    word = list(lines)
    word[-3] = "7\tx\t1\n"
    cases.append((word, K.ERR_IO, "non-numeric"))
    big = list(lines)
    big[20000] = "4294967296\t1\t1\n"
    cases.append((big, K.ERR_IO, "out of range"))
    cases.append((lines + ["1\t2\n"], K.ERR_IO, "truncated"))
    for i, (ls, code, text) in enumerate(cases):
        p = str(tmp_path / f"bad{i}.tsv")
        open(p, "w").write("".join(ls))
        with pytest.raises(K.KtError) as e:
            K.read_tsv(p)
        assert e.value.status == code and text in str(e.value), (i, str(e.value))


def test_truss_tsv_round_trip_and_format(tmp_path):
    s = np.array([0, 1, 1, 2, 9], np.uint32)
    d = np.array([1, 0, 2, 1, 4294967294], np.uint32)
    t = np.array([2, 2, 3, 3, 4294967295], np.uint32)
    p = str(tmp_path / "t.tsv")
    K.write_truss_tsv(p, s, d, t)
    assert open(p).read() == "".join(f"{b + 1}\t{a + 1}\t{w}\n" for a, b, w in zip(s.tolist(), d.tolist(), t.tolist()))
    n, s2, d2, t2 = K.read_truss_tsv(p)
    assert n == 4294967295 and np.array_equal(s2, s) and np.array_equal(d2, d) and np.array_equal(t2, t)
    # the graph reader takes it too (the third column is the weight)
    _, s3, d3 = K.read_tsv(p)
    assert np.array_equal(s3, s) and np.array_equal(d3, d)
    with pytest.raises(ValueError):
        K.write_truss_tsv(p, s, d, t[:2])


def test_binary_edge_cache_format_and_errors(tmp_path):
    s, d = big_symmetric(n=1000, m=5000, seed=3)
    p = str(tmp_path / "g.ktb")
    K.write_edges_bin(p, 1000, s, d)
    raw = open(p, "rb").read()
    assert len(raw) == 32 + 8 * len(s)
    magic, n, flags, m, _ = struct.unpack("<8sIIQQ", raw[:32])
    assert (magic, n, flags, m) == (b"KTEDGES1", 1000, K.INPUT_SYMMETRIC, len(s))
    assert np.array_equal(np.frombuffer(raw[32:32 + 4 * m], "<u4"), s)
    assert np.array_equal(np.frombuffer(raw[32 + 4 * m:], "<u4"), d)
    assert K.is_edges_bin(p)
    n2, s2, d2, pairs = K.read_edges_bin(p)
    assert n2 == 1000 and not pairs and np.array_equal(s2, s) and np.array_equal(d2, d)
    K.write_edges_bin(p, 1000, s[:10], d[:10], pairs=True)
    assert K.read_edges_bin(p)[3] is True
    # empty
    K.write_edges_bin(p, 0, [], [])
    n0, s0, _, _ = K.read_edges_bin(p)
    assert n0 == 0 and len(s0) == 0
    bad = bytearray(raw)
    for where in (40, 32 + 4 * m + 8):  # one src, one dst value
        b2 = bytearray(raw)
        b2[where] ^= 1
        open(p, "wb").write(bytes(b2))
        with pytest.raises(K.KtError) as e:
            K.read_edges_bin(p)
        assert e.value.status == K.ERR_IO and "checksum" in str(e.value)
    # two entries swapped
    b2 = bytearray(raw)
    b2[32:36], b2[36:40] = raw[36:40], raw[32:36]
    if raw[32:36] != raw[36:40]:
        open(p, "wb").write(bytes(b2))
        with pytest.raises(K.KtError):
            K.read_edges_bin(p)
    open(p, "wb").write(bytes(bad[:-4]))
    with pytest.raises(K.KtError) as e:
        K.read_edges_bin(p)
    assert "size" in str(e.value)
    open(p, "wb").write(b"KTEDGES2" + bytes(bad[8:]))
    with pytest.raises(K.KtError) as e:
        K.read_edges_bin(p)
    assert "not a binary edge file" in str(e.value)
    assert not K.is_edges_bin(str(tmp_path / "missing.ktb"))


def test_cost_model_export_for_the_command_line(tmp_path):
    """learn.export_model_bin writes the picker's model in the layout the
    ktruss command line reads (-M); the committed file is the current model."""
    import learn
    p = str(tmp_path / "m.ktm")
    learn.export_model_bin(learn.MODEL_PATH, p)
    raw = open(p, "rb").read()
    assert raw == open(learn.MODEL_BIN_PATH, "rb").read()
    with np.load(learn.MODEL_PATH, allow_pickle=False) as z:
        params, cands, lo, hi = z["params"], z["candidates"], z["lo"], z["hi"]
    C, P = struct.unpack("<II", raw[8:16])
    assert raw[:8] == b"KTMODEL1" and (C, P) == (len(cands), len(params))
    o = 16
    c4 = np.frombuffer(raw[o:o + 16 * C], "<u4").reshape(C, 4)
    o += 16 * C
    assert np.array_equal(c4[:, :3], cands) and not c4[:, 3].any()
    assert np.array_equal(np.frombuffer(raw[o:o + 4 * P], "<f4"), params)
    o += 4 * P
    assert np.array_equal(np.frombuffer(raw[o:o + 64], "<f4"), lo)
    assert np.array_equal(np.frombuffer(raw[o + 64:o + 128], "<f4"), hi)
    assert len(raw) == o + 128


def test_host_validation_reports_lowest_bad_entry():
    """Inputs checked on the host (under 2^16 entries, or before the device is
    initialised) name the lowest offending entry, as the device checks do,
    whichever thread finds which error first."""
    n = 60000
    rng = np.random.default_rng(11)
    key = np.unique(rng.integers(0, n, 40000) * n + rng.integers(0, n, 40000))
    a, b = (key // n).astype(np.uint32), (key % n).astype(np.uint32)
    keep = a != b
    a, b = np.minimum(a, b)[keep][:15000], np.maximum(a, b)[keep][:15000]
    s = np.concatenate([a, b])
    d = np.concatenate([b, a])
    # an entry without its reverse late in the input on a low vertex, a
    # repeated pair early in the input on a high vertex
    lo_v = int(np.argmin(a))
    hi_v = int(np.argmax(a))
    s2 = np.concatenate([[a[hi_v], b[hi_v]], s, [a[lo_v] if a[lo_v] != 0 else 1], ]).astype(np.uint32)
    d2 = np.concatenate([[b[hi_v], a[hi_v]], d, [n - 1]]).astype(np.uint32)
    if (s2[-1], d2[-1]) in set(zip(s.tolist(), d.tolist())) or s2[-1] == d2[-1]:
        pytest.skip("planted entry collides")
    with pytest.raises(K.KtError) as e:
        K.Graph(n, s2, d2)
    assert e.value.status == K.ERR_DUPLICATE
    assert re.search(r"entry (\d+) = ", str(e.value)).group(1) == "0"
    # a range error before a self-loop: the range error is reported
    s3, d3 = s.copy(), d.copy()
    s3[100], d3[100] = 7, 7
    s3[50] = n
    with pytest.raises(K.KtError) as e:
        K.Graph(n, s3, d3)
    assert e.value.status == K.ERR_RANGE and "entry 50 " in str(e.value)
