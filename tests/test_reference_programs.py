"""The reference's own programs, unchanged, on the MI355X.

oracle/ref_programs.py compiles the reference's C++ tests (tests/*/*.cpp) and
example programs (examples/*/*.cpp) against this repository's include/ and
libsygraph_hip into oracle/_ref/.  Here the tests must exit 0 with their
assert()s enabled, and the examples must validate (BFS, SSSP: the examples'
own CPU checks) or agree with the oracle on what they print (distances,
triangle count).  The programs are skipped where they were not built (no
reference tree at build time).
"""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")

REF_TESTS = [
    "frontier_bitmap_frontier", "frontier_mlb_frontier", "formats_csr", "formats_coo2csr",
    "formats_coo2csr_unweighted", "formats_properties", "graph_graph_teardown", "operators_advance",
    "operators_advance_graph", "operators_advance_pull", "operators_filter_compute", "operators_intersection",
    "algorithms_bfs", "algorithms_sssp", "algorithms_cc", "algorithms_tc", "algorithms_bc",
]


def _exe(kind, name):
    p = os.path.join(REF, kind, name)
    if not os.path.exists(p):
        pytest.skip(f"{kind}/{name} not built (reference tree absent at build time)")
    return p


def _run(args, timeout=120):
    print("running:", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    return r


@pytest.mark.parametrize("name", REF_TESTS)
def test_reference_cpp_test(device, name):
    r = _run([_exe("tests", name)])
    assert r.returncode == 0, f"{name} rc={r.returncode}\nstdout:\n{r.stdout[-2000:]}\nstderr:\n{r.stderr[-2000:]}"


def _write_coo(path, g, weights=None):
    """COO text (io/read_coo.hpp): header "n n m", then "u v [w]" per line;
    each undirected edge once (u < v), read back with -u."""
    src = np.repeat(np.arange(g.n), np.diff(g.offsets))
    keep = src < g.cols
    u, v = src[keep], g.cols[keep]
    with open(path, "w") as f:
        f.write(f"% generated RMAT graph\n{g.n} {g.n} {len(u)}\n")
        if weights is None:
            np.savetxt(f, np.stack([u, v], 1), fmt="%d")
        else:
            np.savetxt(f, np.stack([u, v, weights[keep]], 1), fmt="%d")


@pytest.fixture(scope="module")
def rmat_file(tmp_path_factory, oracle_mod):
    g = oracle_mod.rmat(11, 8, seed=21)
    w = oracle_mod.fill_weights(g, 3)
    d = tmp_path_factory.mktemp("graphs")
    plain, weighted = str(d / "rmat11.coo"), str(d / "rmat11_w.coo")
    _write_coo(plain, g)
    _write_coo(weighted, g, w)
    return g, plain, weighted


def _printed_table(stdout):
    rows = {}
    for line in stdout.splitlines():
        m = re.match(r"^(\d+)\s+(\d+)(?:\s+\S+)?\s*$", line)
        if m:
            rows[int(m.group(1))] = int(m.group(2))
    return rows


@pytest.mark.parametrize("mode", ["push", "pull", "hybrid"])
def test_example_bfs(device, oracle_mod, rmat_file, mode):
    g, plain, _ = rmat_file
    src = int(np.argmax(np.diff(g.offsets)))
    r = _run([_exe("examples", "bfs"), "-u", "-c", "-p", "-s", str(src), "--advance", mode, plain])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Validation: [Success]" in r.stdout
    ref, levels = oracle_mod.bfs(g, src)
    got = _printed_table(r.stdout)
    expect = {i: int(d) for i, d in enumerate(ref) if d != g.n + 1}
    assert got == expect
    assert f"Iterations: {levels}" in r.stderr


def test_example_bfs_shared_prefetch(device, oracle_mod, rmat_file):
    """examples/bfs built with GRAPH_LOCATION=2 (graph in managed memory) and
    ENABLE_PREFETCH: the reference's own validation and the oracle's depths."""
    g, plain, _ = rmat_file
    src = int(np.argmax(np.diff(g.offsets)))
    r = _run([_exe("examples", "bfs_shared"), "-u", "-c", "-p", "-s", str(src), "--advance", "hybrid", plain])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Validation: [Success]" in r.stdout
    ref, _ = oracle_mod.bfs(g, src)
    assert _printed_table(r.stdout) == {i: int(d) for i, d in enumerate(ref) if d != g.n + 1}


def test_example_sssp(device, oracle_mod, rmat_file):
    g, _, weighted = rmat_file
    oracle_mod.fill_weights(g, 3)
    src = 5
    r = _run([_exe("examples", "sssp"), "-u", "-c", "-p", "-s", str(src), weighted])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Validation: [Success]" in r.stdout
    ref = oracle_mod.sssp(g, src)
    got = _printed_table(r.stdout)
    assert got == {i: int(d) for i, d in enumerate(ref) if d != g.n + 1}


def test_example_tc(device, oracle_mod, rmat_file):
    g, plain, _ = rmat_file
    r = _run([_exe("examples", "tc"), "-u", "-p", plain])
    assert r.returncode == 0, r.stderr[-2000:]
    m = re.search(r"Total num triangles: (\d+)", r.stdout)
    assert m and int(m.group(1)) == oracle_mod.tc(g)


@pytest.mark.parametrize("prog", ["cc", "bc"])
def test_example_runs(device, rmat_file, prog):
    """The cc and bc examples print no result (their print branch is a TODO
    and their validate() returns false, examples/cc/cc.cpp and
    examples/bc/bc.cpp), so what can be compared is their own output: the run
    line and, with -c, the reference's validation stub reporting Failed."""
    _, plain, _ = rmat_file
    r = _run([_exe("examples", prog), "-u", "-c", "-s", "3", plain])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "[!] Done" in r.stderr
    assert f"[*] Running {prog.upper()} on source 3" in r.stdout
    assert "Validation: [Failed] | Validation Time: " in r.stdout


def _run_lengths(dirs):
    """examples/bfs/bfs.cpp:74-97 printAdvanceDetails: the per-level directions
    as runs, "(push x2, pull x3, push x1)"."""
    runs = []
    for d in dirs:
        name = "push" if d == 0 else "pull"
        if runs and runs[-1][0] == name:
            runs[-1][1] += 1
        else:
            runs.append([name, 1])
    return "(" + ", ".join(f"{k} x{c}" for k, c in runs) + ")"


def test_example_bfs_hybrid_run_details(device, oracle_mod, rmat_file):
    """The hybrid run details the bfs example prints (Iterations, Push steps,
    Pull steps and the direction runs, examples/bfs/bfs.cpp:74-97 and
    :145-151) against the reference rule (alpha 15, beta 24) restated in the
    oracle's OpenMP port of the plain method."""
    g, plain, _ = rmat_file
    deg = np.diff(g.offsets)
    for src in (int(np.argmax(deg)), int(np.nonzero(deg == 1)[0][0])):
        r = _run([_exe("examples", "bfs"), "-u", "-s", str(src), "--advance", "hybrid", plain])
        assert r.returncode == 0, r.stderr[-2000:]
        _, levels, dirs = oracle_mod.bfs_plain_omp(g, src, oracle_mod.BFS_HYBRID, 15.0, 24.0, with_dirs=True)
        assert f"Iterations: {levels}" in r.stderr
        assert f"Push steps: {int((dirs == 0).sum())}" in r.stderr
        assert f"Pull steps: {int((dirs == 1).sum())}" in r.stderr
        assert _run_lengths(dirs) in r.stdout.splitlines()
        assert f"[*] Running BFS (hybrid advance, alpha=15, beta=24) from source vertex {src}" in r.stdout


def _profile_lines(stdout):
    kernels = re.findall(r"^ Kernel \[(\S+) x (\d+)\] (.*)Time: ([0-9.eE+-]+) ms$", stdout, re.M)
    total = re.search(r"^Total GPU Time: ([0-9.eE+-]+) ms$", stdout, re.M)
    mteps = re.search(r"^Total Edge-Througput \(MTEPS\): ([0-9.eE+-]+) MTEPS$", stdout, re.M)
    return kernels, total, mteps


@pytest.mark.parametrize("detailed", [False, True])
def test_example_bfs_profiler_layout(device, oracle_mod, rmat_file, detailed):
    """Profiler::print (utils/profiler.hpp:47-68): one " Kernel [tag x n] "
    line per tag (with -V each event's "(i) t ms " before "| "), then "Total
    GPU Time: t ms" = the sum of the tags' times and "Total Edge-Througput
    (MTEPS): x MTEPS" = visited edges / 1e6 / (t / 1e3), the visited edges
    being the degrees of the reached vertices (algorithms/bfs.hpp:88-99),
    here from the oracle's depths."""
    g, plain, _ = rmat_file
    src = int(np.argmax(np.diff(g.offsets)))
    r = _run([_exe("examples", "bfs_profiling"), "-u", "-V" if detailed else "-v", "-s", str(src), plain])
    assert r.returncode == 0, r.stderr[-2000:]
    kernels, total, mteps = _profile_lines(r.stdout)
    assert kernels and total and mteps, r.stdout[-2000:]
    t_sum = sum(float(k[3]) for k in kernels)
    t_tot = float(total.group(1))
    assert abs(t_sum - t_tot) <= 1e-3 * max(t_tot, 1e-3) + 1e-4
    for tag, count, detail, t in kernels:
        if detailed:
            per = [float(m[1]) for m in re.findall(r"\((\d+)\) ([0-9.eE+-]+) ms ", detail)]
            assert len(per) == int(count) and detail.endswith("| ")
            assert abs(sum(per) - float(t)) <= 1e-3 * max(float(t), 1e-3) + 1e-4 * int(count)
        else:
            assert detail == ""
    depth, _ = oracle_mod.bfs(g, src)
    visited = int(np.diff(g.offsets)[depth != g.n + 1].sum())
    want = visited / 1e6 / (t_tot / 1e3)
    assert abs(float(mteps.group(1)) - want) <= 2e-3 * want + 1e-3


@pytest.mark.parametrize("prog", ["bfs", "sssp", "cc", "tc", "bc"])
def test_example_profiling_build(device, rmat_file, prog):
    _, plain, weighted = rmat_file
    r = _run([_exe("examples", prog + "_profiling"), "-u", "-v", "-s", "3", weighted if prog == "sssp" else plain])
    assert r.returncode == 0, r.stderr[-2000:]
