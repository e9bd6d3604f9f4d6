"""Batched BC and multi-source BFS at full size on both batched-BC graphs,
256 sources in four 64-lane batches, against committed oracle fixtures
(tests/golden/c4_bc.json, written by tests/golden/make_golden.py --c4):

  rmat -- RMAT scale 20, edge factor 56, a/b/c/d = .55/.20/.20/.05, seed 6:
          1,048,576 vertices, 106,398,472 arcs, average degree 101.5;
  grid -- the 4096 x 4096 grid with 10% of its edges deleted (seed 4):
          16,777,216 vertices, ~8,000 BFS levels from a corner.

BC (FP32 path counts with a per-(vertex, source) exponent) must be within
1e-5 relative of the FP64 oracle at 2,048 sampled vertices and at the 64
largest scores, and its total within 1e-5 of the oracle's, which itself
equals the path-length identity sum_s sum_t (d(s,t) - 1) (checked when the
fixture was written).  Multi-source BFS depths are compared per source with
the oracle's depth-array digests, bit-exact.
"""
import json
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
RTOL = 1e-5


def _fixture(key):
    with open(os.path.join(GOLDEN, "c4_bc.json")) as f:
        return json.load(f)[key]


def _graph(sg, device, rec):
    p = rec["params"]
    if p["kind"] == "rmat":
        G = sg.Graph.rmat(device, p["scale"], p["edge_factor"], p["a"], p["b"], p["c"], seed=p["seed"])
    else:
        G = sg.Graph.grid(device, p["rows"], p["cols"], p["p_delete"], seed=p["seed"])
    assert (G.n, G.m) == (rec["n"], rec["m"])
    assert [str(x) for x in G.digest()] == rec["digest"]
    return G


def _rel_err(got, ref):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)


@pytest.mark.parametrize("key", ["rmat", "grid"])
def test_c4_bc_batched_against_fixture(sg, device, key):
    rec = _fixture(key)
    assert abs(rec["bc_sum"] - rec["identity_sum"]) <= 1e-9 * rec["identity_sum"]
    G = _graph(sg, device, rec)
    b = sg.BC(G)
    b.init(rec["sources"])
    t0 = time.perf_counter()
    b.run(batched=True)
    print(f"\nC4 {key}: batched BC over {len(rec['sources'])} sources in {time.perf_counter() - t0:.3f} s", flush=True)
    got = b.get_values().astype(np.float64)
    sample = np.asarray(rec["sample"], dtype=np.int64)
    ref = np.asarray(rec["sample_bc"])
    err = _rel_err(got[sample], ref)
    ok = np.abs(got[sample] - ref) <= RTOL * np.abs(ref) + 1e-6
    assert ok.all(), f"sampled vertices: max rel err {err.max():.3e} at {sample[np.argmax(err)]}"
    top = np.asarray(rec["top"], dtype=np.int64)
    terr = _rel_err(got[top], rec["top_bc"])
    assert terr.max() <= RTOL, f"largest scores: max rel err {terr.max():.3e}"
    total = float(got.sum())
    assert abs(total - rec["bc_sum"]) <= RTOL * rec["bc_sum"], (total, rec["bc_sum"])


@pytest.mark.parametrize("key", ["rmat", "grid"])
def test_c4_msbfs_against_fixture(sg, device, key):
    rec = _fixture(key)
    G = _graph(sg, device, rec)
    t0 = time.perf_counter()
    out = sg.bfs_multi(G, rec["sources"], on_device=True)
    print(f"\nC4 {key}: multi-source BFS over {len(rec['sources'])} sources in {time.perf_counter() - t0:.3f} s",
          flush=True)
    bad = [i for i in range(len(rec["sources"])) if str(out.digest(G.n, 4, i * G.n)) != rec["depth_digest"][i]]
    assert not bad, f"{len(bad)} sources differ, first rows {bad[:5]}"
