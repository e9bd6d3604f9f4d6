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
    _, plain, _ = rmat_file
    r = _run([_exe("examples", prog), "-u", "-s", "3", plain])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "[!] Done" in r.stderr


@pytest.mark.parametrize("prog", ["bfs", "sssp", "cc", "tc", "bc"])
def test_example_profiling_build(device, rmat_file, prog):
    _, plain, weighted = rmat_file
    r = _run([_exe("examples", prog + "_profiling"), "-u", "-v", "-s", "3", weighted if prog == "sssp" else plain])
    assert r.returncode == 0, r.stderr[-2000:]
