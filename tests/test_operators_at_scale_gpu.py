"""The drop-in operator API at scale, as a user of the reference drives it
(tests/cpp/operators_at_scale.cpp, built by tests/cpp/build.py): BFS through
advance::frontier with push, pull and pull_all and each of the four load
balancers, SSSP through the push advance and filter::external with each load
balancer, filter::inplace and compute::reduce with plus, multiplies, minimum
and maximum -- on RMAT scale 14 (the CPU-path configuration: edge factor 16,
a/b/c/d = .57/.19/.19/.05, seed 1, source the highest-degree vertex) and on a
graph of two hubs of 3,000 and 2,000 arcs (the workgroup bin of the
bucketing balancer, and the wavefront-wide walk of long in-lists in pull),
against the oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_bin", "operators_at_scale")
RUNS = [f"bfs {d} {lb}" for d in ("push", "pull", "pull_all")
        for lb in ("workitem_mapped", "subgroup_mapped", "workgroup_mapped", "bucketing")]
RUNS += [f"sssp push {lb}" for lb in ("workitem_mapped", "subgroup_mapped", "workgroup_mapped", "bucketing")]


def _two_hubs(oracle_mod):
    rng = np.random.default_rng(5)
    n = 6000
    u = [np.zeros(3000, dtype=np.int64), np.full(2000, 3001, dtype=np.int64)]
    v = [np.arange(1, 3001), np.arange(3002, 5002)]
    u.append(np.arange(5001, 5999))  # a path from the second hub's last leaf
    v.append(np.arange(5002, 6000))
    u.append(rng.integers(1, 5001, 1500))  # chords between leaves
    v.append(rng.integers(1, 5001, 1500))
    u.append(np.array([2999]))  # the two stars joined by one leaf-to-hub arc
    v.append(np.array([3001]))
    return oracle_mod.csr_from_pairs(n, np.concatenate(u), np.concatenate(v))


@pytest.mark.parametrize("graph", ["rmat14", "two_hubs"])
def test_operator_api_at_scale(oracle_mod, graph, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("tests/cpp/_bin/operators_at_scale not built (tests/cpp/build.py)")
    g = oracle_mod.rmat(14, 16, seed=1) if graph == "rmat14" else _two_hubs(oracle_mod)
    w = oracle_mod.fill_weights(g, 3)
    src = int(np.argmax(g.degrees()))
    path = tmp_path / "g.bin"
    with open(path, "wb") as f:
        np.array([g.n, len(g.cols)], dtype="<i8").tofile(f)
        g.offsets.astype("<i8").tofile(f)
        g.cols.astype("<i4").tofile(f)
        np.asarray(w, dtype="<u4").tofile(f)
    out = tmp_path / "out.bin"
    print(f"running {EXE} on {graph} (n={g.n}, m={len(g.cols)}, source {src})", flush=True)
    r = subprocess.run([EXE, str(path), str(src), str(out)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    raw = out.read_bytes()
    n = g.n
    arr = np.frombuffer(raw, dtype="<u4", count=len(RUNS) * n).reshape(len(RUNS), n)
    member = np.frombuffer(raw, dtype=np.uint8, count=n, offset=4 * len(RUNS) * n)
    red = np.frombuffer(raw, dtype="<u8", count=5, offset=4 * len(RUNS) * n + n)
    depth, _ = oracle_mod.bfs(g, src)
    dist = oracle_mod.sssp(g, src, w)
    for i, name in enumerate(RUNS):
        want = depth if name.startswith("bfs") else dist
        assert np.array_equal(arr[i], want.astype(np.uint32)), name
    reached = np.nonzero(depth != n + 1)[0]
    d = depth.astype(np.int64)
    assert np.array_equal(member.astype(bool), np.isin(np.arange(n), reached) & (d % 2 == 0))
    lvl1 = np.nonzero(depth == 1)[0]
    small = lvl1[lvl1 < 64]
    want = [int(d[reached].sum()), int(d[reached].max()), int((d[reached] + reached % 7).min()),
            int(np.prod(1 + small % 2, dtype=np.uint64)), 5 + len(lvl1)]
    assert [int(x) for x in red] == want
