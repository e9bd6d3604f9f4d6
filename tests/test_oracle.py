"""Pins the CPU oracle (oracle/) before anything is compared against it.

Each oracle function is checked against (1) the known answers held by the
reference's own tests, (2) textbook library routines on the same graph
(scipy.sparse.csgraph, networkx), (3) closed forms and (4) brute force, chosen
so that a plausible slip (an off-by-one level, a missing symmetrisation, a
transposed operand, a wrong sign in the dependency update) fails at least one.
"""
import itertools
import json
import os

import networkx as nx
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _cases():
    with open(os.path.join(GOLDEN, "reference_cases.json")) as f:
        return json.load(f)


def _graph(oracle_mod, name):
    c = _cases()
    return oracle_mod.from_dense(oracle_mod.parse_matrix_text(c["matrices"][name]))


def _scipy(g, weights=None):
    data = np.ones(g.m) if weights is None else weights.astype(np.float64)
    return sp.csr_matrix((data, g.cols, g.offsets), shape=(g.n, g.n))


# ----------------------------------------------------- reference known answers
def test_bfs_line5_known_answer(oracle_mod):
    # tests/algorithms/bfs.cpp:11 expects {0,1,2,3,4}; iterations == 5 (:28)
    g = _graph(oracle_mod, "line_5")
    d, lv = oracle_mod.bfs(g, 0)
    assert d.tolist() == _cases()["expect"]["bfs_line_5"]
    assert lv == 5


def test_bfs_symmetric6_known_answer(oracle_mod):
    # tests/operators/advance.cpp:41 expects {0,1,1,2,2,3}
    g = _graph(oracle_mod, "symmetric_6nodes")
    d, _ = oracle_mod.bfs(g, 0)
    assert d.tolist() == _cases()["expect"]["advance_bfs_symmetric_6nodes"]


def test_sssp_weighted_directed_known_answer(oracle_mod):
    # tests/algorithms/sssp.cpp:18 expects {0,1,3,4,5} on the directed weighted fixture
    g = _graph(oracle_mod, "weighted_directed_5")
    d = oracle_mod.sssp(g, 0)
    assert d.tolist() == _cases()["expect"]["sssp_weighted_directed_5"]


def test_tc_triangle_known_answer(oracle_mod):
    # tests/algorithms/tc.cpp:11 expects 1
    assert oracle_mod.tc(_graph(oracle_mod, "triangle_3")) == 1


def test_cc_two_cc(oracle_mod):
    # two_cc (io/matrices.hpp:33-41): {0..4} connected, 5 isolated -> max ids
    lab = oracle_mod.cc(_graph(oracle_mod, "two_cc"))
    assert lab.tolist() == _cases()["expect"]["cc_two_cc"]


def test_bc_line5_closed_form(oracle_mod):
    # path P_n from end 0: delta_0(v) = n-1-v for 0 < v < n
    g = _graph(oracle_mod, "line_5")
    assert oracle_mod.bc(g, [0]).tolist() == _cases()["expect"]["bc_line_5_source0"]


def test_bc_star_closed_forms(oracle_mod):
    g = _graph(oracle_mod, "star_5")
    assert np.all(oracle_mod.bc(g, [0]) == 0.0)  # from the centre nothing passes through
    b = oracle_mod.bc(g, [1])  # from a leaf every other leaf is reached via the centre
    assert b.tolist() == [3.0, 0.0, 0.0, 0.0, 0.0]


# ----------------------------------------------------- generators / invariants
def _check_csr_invariants(g):
    off, cols = g.offsets, g.cols
    assert off[0] == 0 and np.all(np.diff(off) >= 0) and off[-1] == len(cols)
    rows = np.repeat(np.arange(g.n), np.diff(off))
    assert np.all(cols != rows), "self loop"
    # strictly ascending within each row (sorted, no duplicates)
    same_row = rows[1:] == rows[:-1]
    assert np.all(cols[1:][same_row] > cols[:-1][same_row])
    # symmetric
    A = sp.csr_matrix((np.ones(len(cols)), cols, off), shape=(g.n, g.n))
    assert (A != A.T).nnz == 0


def test_grid_edge_count_closed_form(oracle_mod):
    r, c = 37, 53
    g = oracle_mod.grid(r, c, 0.0)
    assert g.m == 2 * (r * (c - 1) + (r - 1) * c)
    _check_csr_invariants(g)
    deg = g.degrees().reshape(r, c)
    assert deg[0, 0] == 2 and deg[1, 1] == 4 and deg[0, 1] == 3


def test_grid_deletion_rate(oracle_mod):
    g = oracle_mod.grid(256, 256, 0.1, seed=4)
    full = 2 * (256 * 255 * 2)
    kept = g.m / full
    assert abs(kept - 0.9) < 0.01
    _check_csr_invariants(g)


def test_rmat_invariants_and_skew(oracle_mod):
    g = oracle_mod.rmat(12, 16, seed=7)
    _check_csr_invariants(g)
    deg = g.degrees()
    assert deg.max() > 20 * deg.mean()  # power-law skew
    # relabel is a bijection on 2^scale ids
    ids = np.array([oracle_mod.lib().or_rmat_relabel(i, 12, 7) for i in range(1 << 12)])
    assert len(np.unique(ids)) == 1 << 12


def test_rmat_without_relabel_concentrates_on_low_ids(oracle_mod):
    # with a > d, quadrant (0,0) dominates, so vertex 0 has the highest degree
    g = oracle_mod.rmat(10, 16, seed=3, relabel=False)
    assert int(np.argmax(g.degrees())) == 0


def test_er_invariants(oracle_mod):
    g = oracle_mod.erdos_renyi(3000, 12000, seed=5)
    _check_csr_invariants(g)
    assert abs(g.m / (2 * 12000) - 1.0) < 0.01


def test_weights_symmetric_in_range(oracle_mod):
    g = oracle_mod.rmat(10, 8, seed=2)
    w = oracle_mod.fill_weights(g, 3)
    assert w.min() >= 1 and w.max() <= 255
    A = sp.csr_matrix((w.astype(np.int64), g.cols, g.offsets), shape=(g.n, g.n))
    assert (A != A.T).nnz == 0


def test_digest_order_sensitive(oracle_mod):
    x = np.arange(100, dtype=np.int32)
    y = x.copy()
    y[[3, 4]] = y[[4, 3]]
    assert oracle_mod.digest_i32(x) != oracle_mod.digest_i32(y)


# ----------------------------------------------------- library cross-checks
@pytest.fixture(scope="module")
def rmat13(oracle_mod):
    g = oracle_mod.rmat(13, 16, seed=11)
    oracle_mod.fill_weights(g, 12)
    return g


def test_bfs_matches_scipy(oracle_mod, rmat13):
    g = rmat13
    A = _scipy(g)
    for src in [0, int(np.argmax(g.degrees())), 4095]:
        d, _ = oracle_mod.bfs(g, src)
        ref = cg.shortest_path(A, unweighted=True, indices=src)
        ref = np.where(np.isinf(ref), g.n + 1, ref).astype(np.int32)
        assert np.array_equal(d, ref)


def test_sssp_matches_scipy_dijkstra(oracle_mod, rmat13):
    g = rmat13
    A = _scipy(g, g.weights)
    for src in [1, int(np.argmax(g.degrees()))]:
        d = oracle_mod.sssp(g, src)
        ref = cg.dijkstra(A, indices=src)
        ref = np.where(np.isinf(ref), g.n + 1, ref).astype(np.uint32)
        assert np.array_equal(d, ref)


def test_cc_matches_scipy(oracle_mod, rmat13):
    g = rmat13
    lab = oracle_mod.cc(g)
    ncomp, comp = cg.connected_components(_scipy(g), directed=False)
    assert len(np.unique(lab)) == ncomp
    for c in range(ncomp):
        members = np.nonzero(comp == c)[0]
        assert np.all(lab[members] == members.max())


def test_tc_matches_networkx_and_brute_force(oracle_mod):
    g = oracle_mod.rmat(11, 8, seed=9)
    G = nx.from_scipy_sparse_array(_scipy(g))
    assert oracle_mod.tc(g) == sum(nx.triangles(G).values()) // 3
    small = oracle_mod.erdos_renyi(40, 200, seed=1)
    A = _scipy(small).toarray() > 0
    brute = sum(1 for a, b, c in itertools.combinations(range(40), 3) if A[a, b] and A[b, c] and A[a, c])
    assert oracle_mod.tc(small) == brute


def test_tc_complete_graph_closed_form(oracle_mod):
    n = 12
    adj = np.ones((n, n), dtype=np.int64) - np.eye(n, dtype=np.int64)
    assert oracle_mod.tc(oracle_mod.from_dense(adj)) == n * (n - 1) * (n - 2) // 6
    assert oracle_mod.tc(oracle_mod.grid(20, 20)) == 0


def test_bc_matches_networkx(oracle_mod):
    g = oracle_mod.rmat(9, 8, seed=4)
    G = nx.from_scipy_sparse_array(_scipy(g))
    srcs = [0, 5, int(np.argmax(g.degrees()))]
    b = oracle_mod.bc(g, srcs)
    ref = nx.betweenness_centrality_subset(G, sources=srcs, targets=list(G.nodes()), normalized=False)
    ref = np.array([ref[i] for i in range(g.n)]) * 2.0  # networkx halves undirected scores
    assert np.allclose(b, ref, rtol=1e-12, atol=1e-9)


def _random_digraph(oracle_mod, n, m, seed):
    rng = np.random.default_rng(seed)
    u, v = rng.integers(0, n, m), rng.integers(0, n, m)
    keep = u != v
    pairs = np.unique(np.stack([u[keep], v[keep]], 1), axis=0)
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(pairs[:, 0], minlength=n), out=off[1:])
    return oracle_mod.CSRGraph(n, off, np.ascontiguousarray(pairs[:, 1], dtype=np.int32)), pairs


def test_transpose_is_inverse_csr(oracle_mod):
    g, pairs = _random_digraph(oracle_mod, 200, 900, seed=2)
    t = oracle_mod.transpose(g)
    back = oracle_mod.transpose(t)
    assert np.array_equal(back.offsets, g.offsets) and np.array_equal(back.cols, g.cols)
    for v in (0, 17, 199):
        assert t.cols[t.offsets[v]:t.offsets[v + 1]].tolist() == sorted(pairs[pairs[:, 1] == v, 0].tolist())


def test_bc_directed_matches_networkx(oracle_mod):
    # directed: sigma over in-arcs, dependencies over out-arcs, no halving
    n = 300
    g, pairs = _random_digraph(oracle_mod, n, 1500, seed=5)
    srcs = [0, 7, 42]
    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    G.add_edges_from(pairs.tolist())
    ref = nx.betweenness_centrality_subset(G, sources=srcs, targets=list(G.nodes()), normalized=False)
    ref = np.array([ref[i] for i in range(n)])
    assert np.allclose(oracle_mod.bc(g, srcs, inverse=oracle_mod.transpose(g)), ref, rtol=1e-12, atol=1e-9)


def test_bc_grid_scaling_stays_finite(oracle_mod):
    # path counts on a 200x200 grid exceed 1e116 at the far corner; the scaled
    # oracle must stay finite and symmetric under the grid's reflection.
    g = oracle_mod.grid(200, 200)
    b = oracle_mod.bc(g, [0])
    assert np.all(np.isfinite(b))
    B = b.reshape(200, 200)
    assert np.allclose(B, B.T, rtol=1e-9)
    # everything but the source is passed through at most n-2 times
    assert b.max() <= g.n - 2


def test_bc_unscaled_equivalence_small_grid(oracle_mod):
    g = oracle_mod.grid(12, 9)
    G = nx.grid_2d_graph(12, 9)
    mapping = {(r, c): r * 9 + c for r, c in G.nodes()}
    G = nx.relabel_nodes(G, mapping)
    ref = nx.betweenness_centrality_subset(G, sources=[0, 50], targets=list(G.nodes()), normalized=False)
    ref = np.array([ref[i] for i in range(g.n)]) * 2.0
    assert np.allclose(oracle_mod.bc(g, [0, 50]), ref, rtol=1e-12, atol=1e-9)


def test_plain_omp_matches_oracle_all_modes(oracle_mod, rmat13):
    g = rmat13
    src = int(np.argmax(g.degrees()))
    d, lv = oracle_mod.bfs(g, src)
    for mode in (oracle_mod.BFS_PUSH, oracle_mod.BFS_PULL, oracle_mod.BFS_HYBRID):
        d2, lv2 = oracle_mod.bfs_plain_omp(g, src, mode)
        assert np.array_equal(d, d2) and lv == lv2
    ds = oracle_mod.sssp(g, src)
    ds2, _ = oracle_mod.sssp_plain_omp(g, src)
    assert np.array_equal(ds, ds2)


def test_hybrid_schedule_line5(oracle_mod):
    # tests/algorithms/bfs.cpp:25-28: hybrid with alpha=beta=1 runs 5 iterations
    g = _graph(oracle_mod, "line_5")
    d, lv, dirs = oracle_mod.bfs_plain_omp(g, 0, oracle_mod.BFS_HYBRID, 1.0, 1.0, with_dirs=True)
    assert d.tolist() == [0, 1, 2, 3, 4] and lv == 5 and len(dirs) == 5


def test_pull_hint_brute_force(oracle_mod):
    g = oracle_mod.rmat(9, 8, seed=3)
    deg = g.degrees()
    h = oracle_mod.pull_hint(g)
    for v in range(g.n):
        nb = g.cols[g.offsets[v]:g.offsets[v + 1]]
        if len(nb) == 0:
            assert h[v] == -1
            continue
        best = max(nb.tolist(), key=lambda u: (deg[u], -u))  # largest degree, lowest id on ties
        assert h[v] == best
    # directed: hints come from in-arcs and rank by out-degree
    t = oracle_mod.transpose(oracle_mod.CSRGraph(3, np.array([0, 2, 3, 3]), np.array([1, 2, 2], dtype=np.int32)))
    hd = oracle_mod.pull_hint(oracle_mod.CSRGraph(3, np.array([0, 2, 3, 3]), np.array([1, 2, 2], dtype=np.int32)), t)
    assert hd.tolist() == [-1, 0, 0]


def test_pull_scan_count_brute_force(oracle_mod):
    """With hints, a vertex whose hint misses scans its in-list, except a
    single in-arc: the list is the hint, already probed."""
    g = oracle_mod.rmat(8, 8, seed=3)
    d, lv = oracle_mod.bfs(g, int(np.argmax(g.degrees())))
    hint = oracle_mod.pull_hint(g)
    single_misses = 0
    for level in range(lv):
        for use_hint in (False, True):
            tot, cnt, miss = oracle_mod.pull_scan_count(g, d, level, hint if use_hint else None)
            t2 = c2 = m2 = 0
            for u in range(g.n):
                if d[u] <= level:
                    continue
                nb = g.cols[g.offsets[u]:g.offsets[u + 1]]
                if len(nb) == 0:
                    continue
                c2 += 1
                if use_hint:
                    t2 += 1
                    if d[hint[u]] == level:
                        continue
                    m2 += 1
                    if len(nb) == 1:
                        single_misses += 1
                        continue
                hit = np.nonzero(d[nb] == level)[0]
                t2 += (hit[0] + 1) if len(hit) else len(nb)
            assert (tot, cnt, miss) == (t2, c2, m2)
    assert single_misses > 0


# ----------------------------------------------------- committed golden vectors
def test_golden_rmat14(oracle_mod):
    z = np.load(os.path.join(GOLDEN, "rmat14_seed1.npz"))
    g = oracle_mod.rmat(14, 16, seed=1)
    assert g.n == int(z["n"]) and g.m == int(z["m"])
    assert list(g.digest()) == [int(x) for x in z["digest"]]
    src = int(z["source"])
    assert src == int(np.argmax(g.degrees()))
    d, lv = oracle_mod.bfs(g, src)
    assert np.array_equal(d, z["depth"]) and lv == int(z["levels"])
    oracle_mod.fill_weights(g, 3)
    assert np.array_equal(oracle_mod.sssp(g, src), z["sssp"])
    assert np.array_equal(oracle_mod.cc(g), z["cc"])
    assert oracle_mod.tc(g) == int(z["tc"])
    assert np.allclose(oracle_mod.bc(g, z["bc_sources"]), z["bc"], rtol=1e-12, atol=1e-12)


def test_bc_path_length_identity_on_long_grids(oracle_mod):
    """Sum over v != s of delta_s(v) is the number of interior vertices of the
    shortest paths, sum_t (d(s,t) - 1) over reached t != s (each s-t path
    passes through d(s,t) - 1 vertices, weighted 1/sigma_st per path).  On a
    1024x1024 grid from a corner path counts reach ~2^2040 and differ by more
    than 2^1074 within a level, beyond FP64 with one scale per level; the
    per-vertex exponent keeps every score finite and the identity exact to
    rounding.  A dropped term or a wrong exponent sign breaks the identity."""
    g = oracle_mod.grid(1024, 1024, 0.1, seed=4)
    srcs = [0, 123457, 777777]
    b = oracle_mod.bc(g, srcs)
    assert np.all(np.isfinite(b)) and b.min() >= 0.0
    want = 0.0
    for s in srcs:
        d, _ = oracle_mod.bfs(g, s)
        reached = d <= g.n
        want += float(np.maximum(d[reached].astype(np.int64) - 1, 0).sum())
    assert abs(b.sum() - want) <= 1e-9 * want
    # and chunking over sources (16 per chunk, summed in chunk order) does not
    # change a sum over 20 sources beyond rounding
    srcs20 = list(range(0, 20 * 997, 997))
    g2 = oracle_mod.grid(150, 150)
    b20 = oracle_mod.bc(g2, srcs20)
    ref = sum(oracle_mod.bc(g2, [s]) for s in srcs20)
    assert np.allclose(b20, ref, rtol=1e-12, atol=1e-12)
