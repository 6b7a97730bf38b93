"""The ./ktruss command line end to end on the GPU: the reference's flags,
its TSV input (graph.cu:37-79), the k-truss it writes (-o) and the
verification against a compare file (-c, verify.cu:12-40; the default
compare file is the reference's, so runs without one pass -c none), checked
against the CPU oracle."""
import re
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


ANSI = re.compile(r"\x1b\[[0-9;]*m")
MS = r"\d+\.\d{6} ms"


def lines_of(out):
    return [ANSI.sub("", x) for x in out.splitlines() if ANSI.sub("", x) != ""]


def assert_in_order(lines, patterns):
    """Every pattern matches a line, in this order (other lines may sit between)."""
    i = 0
    for p in patterns:
        rx = re.compile(p + "$")
        while i < len(lines) and not rx.match(lines[i]):
            i += 1
        assert i < len(lines), f"no line matching {p!r} after the previous ones in:\n" + "\n".join(lines)
        i += 1


@pytest.mark.parametrize("d,f", [("directed-by-degree", "coocsr"), ("undirected", "tiled-coocsr")])
def test_cli_csv_row_and_passed_line(small_graphs, tmp_path, d, f):
    """-v 0 prints one CSV row per timed run (printConfigAsCSV, config.h:247-272,
    then "%f ms") and nothing else; -v 1 ends the verification with "    PASSED"."""
    n, s, dd = small_graphs["plc300"]
    graph = tmp_path / "g.tsv"
    K.write_tsv(str(graph), s, dd)
    tau_o, _, _ = O.decompose(n, s, dd)
    k = int(np.median(tau_o)) + 1
    ref = tmp_path / "ref.tsv"
    K.write_tsv(str(ref), s[tau_o >= k], dd[tau_o >= k])
    r = run_cli("-g", str(graph), "-k", str(k), "-d", d, "-f", f, "-w", "1", "-x", "2", "-c", str(ref), "-v", "0")
    assert r.returncode == 0, r.stderr
    tiles = "8, 8, " if f == "tiled-coocsr" else "N/A, N/A, "
    row = (re.escape(f"{graph}, {n}, {len(s)}, {k}, gpu, {d}, {f}, 0, all, 512, {tiles}") + MS)
    got = lines_of(r.stdout)
    assert len(got) == 2 and all(re.fullmatch(row, x) for x in got), r.stdout
    r = run_cli("-g", str(graph), "-k", str(k), "-d", d, "-f", f, "-c", str(ref), "-v", "1")
    assert r.returncode == 0, r.stderr
    assert_in_order(lines_of(r.stdout), ["Verifying result", "    PASSED", "    Verification time: " + MS,
                                         "Cleaning up", "    Clean up time: " + MS])


def test_cli_v1_phase_labels(small_graphs, tmp_path):
    """-v 1: the reference's stage labels in its order (main.cu:13-53,
    kernel.cu:19, 36, kernel_d.cu:20-84, kernel_ddc.cu:223-254 and the tiled /
    undirected wrappers' labels)."""
    n, s, dd = small_graphs["plc300"]
    graph = tmp_path / "g.tsv"
    K.write_tsv(str(graph), s, dd)
    r = run_cli("-g", str(graph), "-k", "5", "-c", "none", "-v", "1")
    assert r.returncode == 0, r.stderr
    assert_in_order(lines_of(r.stdout), [
        "Setting up", "    Set up time: " + MS, re.escape(f"    Graph: {graph}"), "        # vertices = 300",
        re.escape(f"        # edges = {len(s)}"), "    k: 5", "    Processor: gpu", "        Block size = 512",
        "    Directedness: directed-by-degree", "    Input format: coocsr",
        "    Number of initial iterations to remove edges: 0", "    Recount strategy: all",
        "Computing", re.escape("    Relabel vertices (not optimized): ") + MS,
        "    Copy COO graph to GPU: " + MS, "    Allocate intermediate data structures: " + MS,
        "    Convert undirected COO to directed COO: " + MS,
        re.escape("    Convert directed COO to COOCSR (not optimized): ") + MS,
        "    Performing K-truss", "        Total K-truss time: " + MS,
        "    Convert COO/CSR to COO: " + MS, "    Deallocate COOCSR: " + MS,
        "    Convert directed COO to undirected COO: " + MS, "    Deallocate directed COO: " + MS,
        "    Deallocate intermediate data structures: " + MS, "    Copy COO truss from GPU: " + MS,
        "    Deallocate COO graph and truss from GPU: " + MS,
        re.escape("    Restore original vertex labels (not optimized): ") + MS,
        "    Total time: " + MS, "Cleaning up", "    Clean up time: " + MS])
    r = run_cli("-g", str(graph), "-k", "5", "-d", "undirected", "-f", "tiled-coocsr", "-c", "none", "-v", "1")
    assert r.returncode == 0, r.stderr
    got = lines_of(r.stdout)
    assert_in_order(got, [
        "    Input format: tiled-coocsr", "        Number of tiles = 8",
        "        Number of tiles processed in parallel = 8", "Computing",
        "    Copy COO graph to GPU: " + MS, "    Allocate intermediate data structures: " + MS,
        re.escape("    Convert undirected COO to tiled COOCSR (not optimized): ") + MS,
        "    Performing K-truss", "        Total K-truss time: " + MS,
        "    Convert tiled COOCSR to COO: " + MS, "    Deallocate tiled COOCSR: " + MS,
        "    Deallocate intermediate data structures: " + MS, "    Copy COO truss from GPU: " + MS])
    assert not any(x.startswith(("    Relabel vertices", "    Convert undirected COO to directed COO",
                                 "    Convert directed COO to undirected COO", "    Deallocate directed COO"))
                   for x in got)


@pytest.mark.parametrize("d,f", [("directed-by-degree", "coocsr"), ("directed-by-index", "tiled-coocsr"),
                                 ("undirected", "coocsr")])
def test_cli_v2_iteration_lines(small_graphs, tmp_path, d, f):
    """-v 2 with -e 1 and -r affected: per iteration "Iteration i", the count
    and mark times, in iteration 0 the removal with "# edges remaining" (not
    printed by the undirected COO+CSR variant, whose mark line ends in
    "time"), the affected marking, and the removal after the loop
    (kernel_ddc.cu:126-198, 207-212; kernel_duc.cu:105-188); as many
    iterations as the oracle's peel, and the oracle's remaining edges."""
    n, s, dd = small_graphs["plc300"]
    graph = tmp_path / "g.tsv"
    K.write_tsv(str(graph), s, dd)
    tau_o, _, _ = O.decompose(n, s, dd)
    k = int(np.median(tau_o)) + 1
    _, _, _, trace = O.ktruss(n, s, dd, k, O.make_variant(d, f, "affected", 1))
    assert len(trace) >= 3
    r = run_cli("-g", str(graph), "-k", str(k), "-d", d, "-f", f, "-r", "affected", "-e", "1", "-c", "none",
                "-v", "2")
    assert r.returncode == 0, r.stderr
    got = lines_of(r.stdout)
    duc = d == "undirected" and f == "coocsr"
    mark = "            Mark deleted edges" + (" time" if duc else "") + ": " + MS
    m_und = len(s) // 2
    pats = ["    Performing K-truss"]
    for it, (_, i, deleted, _) in enumerate(trace):
        assert i == it
        pats += [f"        Iteration {it}", "            Triangle counting time: " + MS, mark]
        if deleted and it == 0:
            pats.append("            Remove deleted edges: " + MS)
            if not duc:
                stored = m_und if d != "undirected" else 2 * m_und
                per = 1 if d != "undirected" else 2
                pats.append(re.escape(f"                # edges remaining = {stored - per * deleted}"))
        if deleted and k != 3:
            pats.append("            Mark affected edges: " + MS)
    if any(t[2] for t in trace[1:]):
        pats.append("        Remove deleted edges: " + MS)
    pats.append("        Total K-truss time: " + MS)
    assert_in_order(got, pats)
    assert sum(1 for x in got if re.match(r"        Iteration \d+$", x)) == len(trace)
    assert not any("# edges remaining" in x for x in got) or not duc
