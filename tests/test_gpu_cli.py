"""The ./ktruss command line end to end on the GPU: the reference's flags,
its TSV input (graph.cu:37-79), the k-truss it writes (-o) and the
verification against a compare file (-c, verify.cu:12-40; the default
compare file is the reference's, so runs without one pass -c none), checked
against the CPU oracle."""
import subprocess

import numpy as np
import pytest

import ktruss_amd as K
import oracle as O

pytestmark = pytest.mark.gpu


def run_cli(*args):
    return subprocess.run([K.CLI_PATH, *args], capture_output=True, text=True, timeout=120)


def undirected(src, dst):
    return {(min(a, b), max(a, b)) for a, b in zip(np.asarray(src).tolist(), np.asarray(dst).tolist())}


@pytest.mark.parametrize("d", ["directed-by-degree", "undirected"])
@pytest.mark.parametrize("k", [3, 4, 6])
def test_cli_ktruss_matches_oracle(small_graphs, tmp_path, d, k):
    for name, (n, s, dd) in small_graphs.items():
        graph = tmp_path / f"{name}.tsv"
        K.write_tsv(str(graph), s, dd)
        out = tmp_path / f"{name}_k{k}.tsv"
        r = run_cli("-g", str(graph), "-k", str(k), "-p", "gpu", "-d", d, "-o", str(out), "-c", "none",
                    "-v", "0")
        assert r.returncode == 0, (name, r.stderr)
        _, gs, gd = K.read_tsv(str(out)) if out.stat().st_size else (0, [], [])
        os_, od_, _, _ = O.ktruss(n, s, dd, k, O.make_variant(d))
        assert undirected(gs, gd) == undirected(os_, od_), (name, k, d)


def test_cli_compare_file_and_decompose(small_graphs, tmp_path):
    n, s, d = small_graphs["plc300"]
    graph = tmp_path / "g.tsv"
    K.write_tsv(str(graph), s, d)
    tau_o, _, _ = O.decompose(n, s, d)
    k = int(np.median(tau_o)) + 1
    # the oracle's k-truss as the compare file: verification passes
    keep = tau_o >= k
    ref = tmp_path / "ref.tsv"
    K.write_tsv(str(ref), s[keep], d[keep])
    r = run_cli("-g", str(graph), "-k", str(k), "-p", "gpu", "-c", str(ref), "-v", "0")
    assert r.returncode == 0, r.stdout + r.stderr
    # a compare file missing an edge (both directions) fails the verification
    ks, kd = s[keep], d[keep]
    a, b = ks[0], kd[0]
    drop = ((ks == a) & (kd == b)) | ((ks == b) & (kd == a))
    bad = tmp_path / "bad.tsv"
    K.write_tsv(str(bad), ks[~drop], kd[~drop])
    r = run_cli("-g", str(graph), "-k", str(k), "-p", "gpu", "-c", str(bad), "-v", "0")
    assert r.returncode != 0
    # -D: one line per input entry, "dst src tau" 1-based
    out = tmp_path / "tau.tsv"
    r = run_cli("-g", str(graph), "-D", "-p", "gpu", "-o", str(out), "-c", "none", "-v", "0")
    assert r.returncode == 0, r.stderr
    rows = np.loadtxt(str(out), dtype=np.int64, ndmin=2)
    got = {(int(a) - 1, int(b) - 1): int(t) for a, b, t in rows}
    exp = {(int(b), int(a)): int(t) for a, b, t in zip(s, d, tau_o)}
    assert got == exp


def test_cli_binary_edge_cache(small_graphs, tmp_path):
    """-B writes the input as a binary edge cache; -g reads it back by its
    magic and gives the same truss file as the text input."""
    n, s, d = small_graphs["plc300"]
    graph = tmp_path / "g.tsv"
    K.write_tsv(str(graph), s, d)
    cache = tmp_path / "g.ktb"
    out_txt = tmp_path / "tau_txt.tsv"
    r = run_cli("-g", str(graph), "-B", str(cache), "-D", "-p", "gpu", "-o", str(out_txt), "-c", "none", "-v", "0")
    assert r.returncode == 0, r.stderr
    assert K.is_edges_bin(str(cache)) and not K.is_edges_bin(str(graph))
    n2, s2, d2, pairs = K.read_edges_bin(str(cache))
    assert (n2, pairs) == (n, False) and np.array_equal(s2, s) and np.array_equal(d2, d)
    out_bin = tmp_path / "tau_bin.tsv"
    r = run_cli("-g", str(cache), "-D", "-p", "gpu", "-o", str(out_bin), "-c", "none", "-v", "0")
    assert r.returncode == 0, r.stderr
    assert out_bin.read_bytes() == out_txt.read_bytes()
    tau_o, _, _ = O.decompose(n, s, d)
    _, ts, td, tt = K.read_truss_tsv(str(out_bin))
    assert np.array_equal(ts, s) and np.array_equal(td, d) and np.array_equal(tt, tau_o)


@pytest.mark.parametrize("kernel", ["delta", "search", "merge"])
def test_cli_count_kernel_option(small_graphs, tmp_path, kernel):
    """-K picks the support-pass kernel: the truss file equals the default
    kernel's byte for byte; an unknown name exits with status 2."""
    n, s, d = small_graphs["plc300"]
    graph = tmp_path / "g.tsv"
    K.write_tsv(str(graph), s, d)
    ref, out = tmp_path / "ref.tsv", tmp_path / "out.tsv"
    r = run_cli("-g", str(graph), "-D", "-p", "gpu", "-o", str(ref), "-c", "none", "-v", "0")
    assert r.returncode == 0, r.stderr
    r = run_cli("-g", str(graph), "-D", "-K", kernel, "-p", "gpu", "-o", str(out), "-c", "none", "-v", "1")
    assert r.returncode == 0, r.stderr
    assert f"Count kernel: {kernel}" in r.stdout
    assert out.read_bytes() == ref.read_bytes()
    assert run_cli("-K", "bogus").returncode == 2


def test_cli_cost_model_pick(small_graphs, tmp_path):
    """-M: the learned cost model picks each phase's support-pass kernel; the
    truss file equals the default kernel's; a file that is not a model fails."""
    import learn
    n, s, d = small_graphs["plc300"]
    graph = tmp_path / "g.tsv"
    K.write_tsv(str(graph), s, d)
    ref, out = tmp_path / "ref.tsv", tmp_path / "out.tsv"
    r = run_cli("-g", str(graph), "-D", "-p", "gpu", "-o", str(ref), "-c", "none", "-v", "0")
    assert r.returncode == 0, r.stderr
    r = run_cli("-g", str(graph), "-D", "-M", learn.MODEL_BIN_PATH, "-p", "gpu", "-o", str(out), "-c", "none",
                "-v", "1")
    assert r.returncode == 0, r.stderr
    assert "cost model" in r.stdout
    assert out.read_bytes() == ref.read_bytes()
    r = run_cli("-g", str(graph), "-D", "-M", str(graph), "-p", "gpu", "-c", "none", "-v", "0")
    assert r.returncode != 0 and "not a cost model" in r.stderr
