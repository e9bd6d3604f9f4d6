"""Writes the golden fixtures under tests/golden/ (run from the repo root).

    python tests/golden/make_golden.py            # small fixtures (seconds)
    python tests/golden/make_golden.py --rmat24   # + the RMAT scale-24 summary (minutes)
    python tests/golden/make_golden.py --c3       # + grid / RMAT-20 CC and TC summaries
    python tests/golden/make_golden.py --c4       # + batched BC / MS-BFS fixtures (minutes)

reference_cases.json -- the adjacency-matrix fixtures of the reference's own
  tests (tests/include/test_utils.hpp:21-45, include/sygraph/io/matrices.hpp:21-41)
  and the answers those tests assert, each with its file:line; plus answers that
  follow from closed forms (CC max-label, BC on a path).  These are data typed
  in from the reference tests, cross-checked here against networkx.
rmat14_seed1.npz -- the CPU-device configuration (RMAT scale 14, edge factor 16,
  a/b/c/d = .57/.19/.19/.05, seed 1): digests, source (highest degree, lowest id
  on ties), oracle BFS depths, SSSP distances (weights seed 3), CC labels, TC
  count and FP64 BC for three sources; each cross-checked against scipy/networkx.
c3_cc_tc.json -- CC labels digest / component count on the 4096x4096 grid with 10%
  of edges deleted (seed 4) and on RMAT scale 20 edge factor 32 (seed 5); exact
  triangle count of that RMAT graph.
c4_bc.json -- batched BC / multi-source BFS configurations, 256 sources each:
  (a) RMAT scale 20, edge factor 56, a/b/c/d = .55/.20/.20/.05, seed 6 (1,048,576
  vertices, average degree 101.5) and (b) the 4096x4096 grid with 10% of edges
  deleted (seed 4).  Per graph: CSR digests, the sources (uniform over vertices
  with arcs, seed 6), the FP64 oracle's BC sum, the path-length identity
  sum_s sum_t (d(s,t) - 1), the BC values at 2,048 sampled vertices and at the
  64 largest, and per source the oracle BFS depth digest.
rmat24_seed2.json -- RMAT scale 24 (seed 2): vertex/arc counts, CSR digests, 16
  sources drawn uniformly (seed 2) from the largest component, and per source
  the BFS level count, reached vertices/edges and the depth-array digest; SSSP
  (weights seed 3) distance digests for the first two sources.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

MATRICES = {
    # tests/include/test_utils.hpp:21-26
    "line_5": "5\n0 1 0 0 0\n1 0 1 0 0\n0 1 0 1 0\n0 0 1 0 1\n0 0 0 1 0",
    # tests/include/test_utils.hpp:28-33
    "star_5": "5\n0 1 1 1 1\n1 0 0 0 0\n1 0 0 0 0\n1 0 0 0 0\n1 0 0 0 0",
    # tests/include/test_utils.hpp:35-38
    "triangle_3": "3\n0 1 1\n1 0 1\n1 1 0",
    # tests/include/test_utils.hpp:40-45 (directed, weighted)
    "weighted_directed_5": "5\n0 1 4 0 0\n0 0 2 6 0\n0 0 0 1 5\n0 0 0 0 1\n0 0 0 0 0",
    # include/sygraph/io/matrices.hpp:21-28
    "symmetric_6nodes": "6\n0 1 1 0 0 0\n1 0 1 0 0 0\n1 1 0 1 1 0\n0 0 1 0 0 0\n0 0 1 0 0 1\n0 0 0 0 1 0",
    # include/sygraph/io/matrices.hpp:33-41
    "two_cc": "6\n0 1 1 0 0 0\n1 0 1 0 0 0\n1 1 0 1 1 0\n0 0 1 0 0 0\n0 0 1 0 0 0\n0 0 0 0 0 0",
}

EXPECT = {
    # tests/algorithms/bfs.cpp:11,19,27 (push, pull, hybrid) and :28 iterations == 5
    "bfs_line_5": [0, 1, 2, 3, 4],
    "bfs_line_5_iterations": 5,
    # tests/operators/advance.cpp:41
    "advance_bfs_symmetric_6nodes": [0, 1, 1, 2, 2, 3],
    # tests/operators/advance_pull.cpp:58-64: in = {0,1}; out = {2}; vertex 2 visited once (short circuit)
    "advance_pull_in": [0, 1],
    "advance_pull_out": [2],
    "advance_pull_visits_2": 1,
    # tests/operators/advance_graph.cpp:22: functor degree(u) != 0 over the graph view
    "advance_graph_two_cc": [True, True, True, True, True, False],
    # tests/operators/filter_compute.cpp:15-38 on star_5 with input {0..4}
    "filter_external_even": [0, 2, 4],
    "filter_inplace_lt2": [2, 3, 4],
    "compute_execute_marks": [1, 0, 1, 0, 1],
    "compute_reduce_sum": 6,
    # tests/operators/intersection.cpp:14-20 on line_5
    "intersection_lhs": [0, 1, 3],
    "intersection_rhs": [1, 2, 3],
    "intersection_out": [1, 3],
    # tests/frontier/mlb_frontier.cpp:10-24 on star_5
    "mlb_insert": [1, 3, 4],
    "mlb_after_remove_3": [1, 4],
    # tests/algorithms/sssp.cpp:18 (directed weighted fixture, source 0)
    "sssp_weighted_directed_5": [0, 1, 3, 4, 5],
    # tests/algorithms/tc.cpp:11
    "tc_triangle_3": 1,
    # max vertex id per component of two_cc (fixpoint of algorithms/cc.hpp:103-169)
    "cc_two_cc": [4, 4, 4, 4, 4, 5],
    # path from end 0: delta(v) = n-1-v (closed form; tests/algorithms/bc.cpp only runs)
    "bc_line_5_source0": [0.0, 3.0, 2.0, 1.0, 0.0],
}


def _nx(g):
    import networkx as nx
    import scipy.sparse as sp

    return nx.from_scipy_sparse_array(sp.csr_matrix((np.ones(g.m), g.cols, g.offsets), shape=(g.n, g.n)))


def write_reference_cases():
    import networkx as nx

    # cross-check the typed-in answers against networkx / the oracle
    g = oracle.from_dense(oracle.parse_matrix_text(MATRICES["line_5"]))
    assert [nx.shortest_path_length(_nx(g), 0, v) for v in range(5)] == EXPECT["bfs_line_5"]
    g = oracle.from_dense(oracle.parse_matrix_text(MATRICES["symmetric_6nodes"]))
    assert [nx.shortest_path_length(_nx(g), 0, v) for v in range(6)] == EXPECT["advance_bfs_symmetric_6nodes"]
    g = oracle.from_dense(oracle.parse_matrix_text(MATRICES["weighted_directed_5"]))
    assert oracle.sssp(g, 0).tolist() == EXPECT["sssp_weighted_directed_5"]
    with open(os.path.join(OUT, "reference_cases.json"), "w") as f:
        json.dump({"matrices": MATRICES, "expect": EXPECT}, f, indent=1)


def write_rmat14():
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    g = oracle.rmat(14, 16, 0.57, 0.19, 0.19, seed=1)
    src = int(np.argmax(g.degrees()))
    d, lv = oracle.bfs(g, src)
    A = sp.csr_matrix((np.ones(g.m), g.cols, g.offsets), shape=(g.n, g.n))
    ref = cg.shortest_path(A, unweighted=True, indices=src)
    assert np.array_equal(np.where(np.isinf(ref), g.n + 1, ref).astype(np.int32), d)
    w = oracle.fill_weights(g, 3)
    ds = oracle.sssp(g, src)
    ref = cg.dijkstra(sp.csr_matrix((w.astype(float), g.cols, g.offsets), shape=(g.n, g.n)), indices=src)
    assert np.array_equal(np.where(np.isinf(ref), g.n + 1, ref).astype(np.uint32), ds)
    lab = oracle.cc(g)
    tcount = oracle.tc(g)
    bc_src = np.array([src, 7, 12345], dtype=np.int32)
    b = oracle.bc(g, bc_src)
    np.savez_compressed(os.path.join(OUT, "rmat14_seed1.npz"), n=g.n, m=g.m, digest=np.array(g.digest(), np.uint64),
                        source=src, depth=d, levels=lv, sssp=ds, cc=lab, tc=tcount, bc_sources=bc_src, bc=b)


def write_c3():
    """CC on the 4096x4096 grid (10% of edges deleted, seed 4) and on RMAT scale 20
    edge factor 32 (seed 5); TC on the same RMAT graph."""
    t0 = time.time()
    g = oracle.grid(4096, 4096, 0.1, seed=4)
    lab = oracle.cc(g)
    out = {"grid": {"n": g.n, "m": g.m, "digest": [str(x) for x in g.digest()],
                    "components": int(len(np.unique(lab))), "labels_digest": str(oracle.digest_i32(lab.view(np.int32)))}}
    print(f"grid done {time.time() - t0:.1f}s", flush=True)
    g = oracle.rmat(20, 32, 0.57, 0.19, 0.19, seed=5)
    lab = oracle.cc(g)
    t1 = time.time()
    tri = oracle.tc(g)
    print(f"rmat20 tc {tri} in {time.time() - t1:.1f}s", flush=True)
    out["rmat20"] = {"n": g.n, "m": g.m, "digest": [str(x) for x in g.digest()],
                     "components": int(len(np.unique(lab))), "labels_digest": str(oracle.digest_i32(lab.view(np.int32))),
                     "triangles": str(tri)}
    with open(os.path.join(OUT, "c3_cc_tc.json"), "w") as f:
        json.dump(out, f, indent=1)


def write_rmat24():
    t0 = time.time()
    g = oracle.rmat(24, 16, 0.57, 0.19, 0.19, seed=2)
    print(f"rmat24 built in {time.time() - t0:.1f}s: n={g.n} m={g.m}", flush=True)
    lab = oracle.cc(g)
    vals, cnts = np.unique(lab, return_counts=True)
    big = vals[np.argmax(cnts)]
    members = np.nonzero(lab == big)[0]
    pick = [int(members[int(oracle.lib().or_rand64(2, k) % len(members))]) for k in range(16)]
    deg = g.degrees()
    out = {"n": g.n, "m": g.m, "digest": [str(x) for x in g.digest()], "lcc_size": int(cnts.max()),
           "isolated": int((deg == 0).sum()), "max_degree": int(deg.max()), "sources": pick, "bfs": []}
    for s in pick:
        d, lv = oracle.bfs(g, s)
        reached = d != g.n + 1
        out["bfs"].append({"source": s, "levels": lv, "reached_vertices": int(reached.sum()),
                           "reached_edges": int(deg[reached].sum()), "depth_digest": str(oracle.digest_i32(d))})
        print(f"  bfs {s}: levels {lv}", flush=True)
    oracle.fill_weights(g, 3)
    out["sssp"] = []
    for s in pick[:2]:
        ds = oracle.sssp(g, s)
        out["sssp"].append({"source": s, "dist_digest": str(oracle.digest_i32(ds.view(np.int32))),
                            "max_dist": int(ds[ds != g.n + 1].max())})
    with open(os.path.join(OUT, "rmat24_seed2.json"), "w") as f:
        json.dump(out, f, indent=1)


C4 = {
    "rmat": dict(kind="rmat", scale=20, edge_factor=56, a=0.55, b=0.20, c=0.20, seed=6),
    "grid": dict(kind="grid", rows=4096, cols=4096, p_delete=0.1, seed=4),
}


def c4_graph(key: str):
    p = C4[key]
    if p["kind"] == "rmat":
        return oracle.rmat(p["scale"], p["edge_factor"], p["a"], p["b"], p["c"], seed=p["seed"])
    return oracle.grid(p["rows"], p["cols"], p["p_delete"], seed=p["seed"])


def c4_sources(g, count: int = 256, seed: int = 6) -> list[int]:
    cand = np.flatnonzero(g.degrees() > 0)
    rng = np.random.default_rng(seed)
    return [int(x) for x in cand[rng.integers(0, len(cand), size=count)]]


def write_c4():
    """Batched BC / MS-BFS fixtures (256 sources) for both C4 graphs."""
    from concurrent.futures import ThreadPoolExecutor
    out = {}
    for key in ("rmat", "grid"):
        t0 = time.time()
        g = c4_graph(key)
        srcs = c4_sources(g)
        with ThreadPoolExecutor(8) as ex:
            bfs = list(ex.map(lambda s: oracle.bfs(g, s), srcs))
        ident = 0
        for d, _ in bfs:
            r = d[d != g.n + 1].astype(np.int64)
            ident += int(np.maximum(r - 1, 0).sum())
        b = oracle.bc(g, srcs)
        rng = np.random.default_rng(7)
        sample = np.unique(rng.integers(0, g.n, size=2048))
        top = np.argsort(-b, kind="stable")[:64]
        out[key] = {"params": C4[key], "n": g.n, "m": g.m, "avg_degree": g.m / g.n,
                    "digest": [str(x) for x in g.digest()], "sources": srcs,
                    "bc_sum": float(b.sum()), "identity_sum": ident,
                    "sample": [int(x) for x in sample], "sample_bc": [float(x) for x in b[sample]],
                    "top": [int(x) for x in top], "top_bc": [float(x) for x in b[top]],
                    "levels": [int(lv) for _, lv in bfs],
                    "depth_digest": [str(oracle.digest_i32(d)) for d, _ in bfs]}
        print(f"c4 {key}: n={g.n} m={g.m} bc sum {b.sum():.6e} identity {ident} in {time.time() - t0:.1f}s",
              flush=True)
    with open(os.path.join(OUT, "c4_bc.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--rmat24", action="store_true")
    ap.add_argument("--c3", action="store_true")
    ap.add_argument("--c4", action="store_true")
    a = ap.parse_args()
    write_reference_cases()
    write_rmat14()
    if a.rmat24:
        write_rmat24()
    if a.c3:
        write_c3()
    if a.c4:
        write_c4()
