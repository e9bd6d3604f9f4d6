"""SSSP, CC and TC on the MI355X through the C ABI, bit-exact against the oracle.

Known answers come from the reference's tests (tests/algorithms/sssp.cpp:18,
tests/algorithms/tc.cpp:11) and closed forms; larger cases compare with the
oracle on generated graphs, and the full-size configurations (RMAT scale 24
SSSP, 4096x4096 grid and RMAT scale 20 CC/TC) compare with committed digests.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _cases():
    with open(os.path.join(GOLDEN, "reference_cases.json")) as f:
        return json.load(f)


def _fixture(oracle_mod, name):
    return oracle_mod.from_dense(oracle_mod.parse_matrix_text(_cases()["matrices"][name]))


# ----------------------------------------------------------------- SSSP ---
@pytest.mark.parametrize("weight_float", [False, True])
def test_sssp_weighted_directed_known_answer(sg, device, oracle_mod, weight_float):
    g = _fixture(oracle_mod, "weighted_directed_5")
    G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=g.weights, directed=True, weight_float=weight_float)
    s = sg.SSSP(G)
    s.init(0)
    s.run()
    assert s.get_distances().tolist() == _cases()["expect"]["sssp_weighted_directed_5"]
    # unreachable from 4: everything else stays n+1 (algorithms/sssp.hpp:43)
    s.init(4)
    s.run()
    assert s.get_distances().tolist() == [6, 6, 6, 6, 0]


@pytest.mark.parametrize("record_relax", [False, True])
def test_sssp_rmat14_golden(sg, device, oracle_mod, record_relax):
    """record_relax: each relaxing launch timed by events, the host waiting on
    them; otherwise the host polls each iteration's record."""
    z = np.load(os.path.join(GOLDEN, "rmat14_seed1.npz"))
    G = sg.Graph.rmat(device, 14, 16, seed=1, weight_seed=3)
    s = sg.SSSP(G)
    s.init(int(z["source"]))
    st = s.run(record_relax=record_relax)
    assert np.array_equal(s.get_distances(), z["sssp"])
    assert st.reached_vertices == int((z["sssp"] != G.n + 1).sum())
    assert (st.relax_launches > 0 and st.relax_ms > 0) == record_relax


@pytest.mark.parametrize("lb", [0, 2, 3])
def test_sssp_random_graphs(sg, device, oracle_mod, lb):
    rng = np.random.default_rng(7)
    graphs = [oracle_mod.rmat(12, 16, seed=31), oracle_mod.grid(50, 70, 0.2, seed=32),
              oracle_mod.erdos_renyi(4000, 12000, seed=33)]
    for g in graphs:
        oracle_mod.fill_weights(g, 9)
        G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=g.weights)
        s = sg.SSSP(G)
        for src in rng.integers(0, g.n, size=3):
            s.init(int(src))
            s.run(lb)
            assert np.array_equal(s.get_distances(), oracle_mod.sssp(g, int(src)))


@pytest.mark.parametrize("delta", [float("inf"), -1.0, 0.0, 0.5, 1.0, 7.0, 40.0, 1000.0])
def test_sssp_bucket_width_does_not_change_distances(sg, device, oracle_mod, delta):
    g = oracle_mod.rmat(12, 16, seed=44)
    oracle_mod.fill_weights(g, 5)
    G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=g.weights)
    s = sg.SSSP(G)
    for src in (0, int(np.argmax(g.degrees()))):
        s.init(src)
        st = s.run(delta=delta)
        assert np.array_equal(s.get_distances(), oracle_mod.sssp(g, src))
    if delta in (float("inf"), -1.0):  # the reference's schedule: every dirty vertex each round
        assert st.relaxations >= st.reached_edges


def test_sssp_width_changes_on_one_graph(sg, device, oracle_mod):
    """The light-first adjacency is rebuilt whenever a run asks for another
    bucket width; alternating widths on one graph keep every distance exact."""
    g = oracle_mod.rmat(12, 8, seed=47)
    oracle_mod.fill_weights(g, 8)
    G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=g.weights)
    s = sg.SSSP(G)
    ref = oracle_mod.sssp(g, 5)
    for delta in (7.0, 40.0, 7.0, 0.0, -1.0, 255.0, 1.0, 7.0):
        s.init(5)
        s.run(delta=delta)
        assert np.array_equal(s.get_distances(), ref), delta


def test_sssp_buckets_relax_fewer_arcs(sg, device, oracle_mod):
    g = oracle_mod.rmat(13, 16, seed=45)
    oracle_mod.fill_weights(g, 6)
    G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=g.weights)
    s = sg.SSSP(G)
    src = int(np.argmax(g.degrees()))
    s.init(src)
    bf = s.run(delta=-1.0).relaxations
    s.init(src)
    buckets = s.run(delta=64.0).relaxations  # half the mean of weights U[1,255]
    assert np.array_equal(s.get_distances(), oracle_mod.sssp(g, src))
    assert buckets < bf, (buckets, bf)
    # the automatic schedule of a graph this small (under 2^25 arcs) is the
    # reference's (relaxation counts vary a little from run to run: a round's
    # atomics race with each other)
    s.init(src)
    auto = s.run().relaxations
    assert abs(auto - bf) < 0.05 * bf and abs(auto - bf) < abs(auto - buckets), (auto, bf, buckets)


@pytest.mark.parametrize("kind", ["grid", "rmat"])
def test_sssp_automatic_schedule_rule(sg, device, kind):
    """The automatic schedule takes buckets only on a graph of at least 2^25
    arcs and 8 arcs per vertex: RMAT scale 20 edge factor 32 (about 60 M
    arcs, ~57 per vertex) relaxes fewer arcs than the reference's schedule;
    a 3072 x 3072 grid with 5 % of its edges deleted (about 36 M arcs, 3.8 per
    vertex) relaxes as many.  Distances equal the reference schedule's."""
    if kind == "grid":
        G = sg.Graph.grid(device, 3072, 3072, 0.05, seed=61, weight_seed=62)
    else:
        G = sg.Graph.rmat(device, 20, 32, seed=63, weight_seed=64)
    assert G.m >= 1 << 25
    buckets_expected = G.m >= 8 * G.n
    assert buckets_expected == (kind == "rmat")
    off, _, _ = G.download()
    src = int(np.argmax(np.diff(off)))
    s = sg.SSSP(G)
    s.init(src)
    bf = s.run(delta=-1.0).relaxations
    ref = s.dist.digest(G.n, 4)
    s.init(src)
    auto = s.run().relaxations
    assert s.dist.digest(G.n, 4) == ref
    if buckets_expected:
        assert auto < 0.9 * bf, (auto, bf)
    else:  # a round's atomics race with each other: counts vary a little
        assert abs(auto - bf) < 0.05 * bf, (auto, bf)
    G.destroy()


@pytest.mark.parametrize("delta", [0.0, 0.3, -1.0])
def test_sssp_float_weights_buckets(sg, device, oracle_mod, delta):
    g = oracle_mod.rmat(11, 8, seed=46)
    w = oracle_mod.fill_weights(g, 5)
    G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=(w / 7.0).astype(np.float32), weight_float=True)
    s = sg.SSSP(G)
    s.init(9)
    s.run(delta=delta)
    ref = oracle_mod.sssp(g, 9).astype(np.float64)
    got = s.get_distances().astype(np.float64)
    reach = ref != g.n + 1
    # float sums differ from the integer oracle only by rounding (path sums of w/7)
    assert np.array_equal(got == g.n + 1, ~reach)
    assert np.allclose(got[reach], ref[reach] / 7.0, rtol=1e-5)


@pytest.mark.parametrize("heavy_pull", [1, -1, 0])
def test_sssp_heavy_round_pull_and_push(sg, device, oracle_mod, heavy_pull):
    """A bucket's heavy round either pulls into the vertices not yet settled
    or pushes from the bucket's settled set; the distances are the same.  The
    ER graph has isolated vertices and several components, the grid a long
    diameter (many buckets)."""
    graphs = [oracle_mod.rmat(12, 16, seed=48), oracle_mod.grid(40, 50, 0.1, seed=49),
              oracle_mod.erdos_renyi(3000, 9000, seed=50)]
    for g in graphs:
        oracle_mod.fill_weights(g, 11)
        G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=g.weights)
        s = sg.SSSP(G)
        for src, delta in [(0, 0.0), (g.n // 2, 20.0), (int(np.argmax(g.degrees())), 3.0)]:
            s.init(src)
            s.run(delta=delta, heavy_pull=heavy_pull)
            assert np.array_equal(s.get_distances(), oracle_mod.sssp(g, src)), (g.n, src, delta)


def test_sssp_heavy_pull_relaxes_fewer_arcs_on_rmat(sg, device, oracle_mod):
    # after the first bucket almost every vertex is settled: pulling into the
    # rest relaxes far fewer heavy arcs than pushing all of the settled set's
    g = oracle_mod.rmat(13, 16, seed=52)
    oracle_mod.fill_weights(g, 12)
    G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=g.weights)
    s = sg.SSSP(G)
    src = int(np.argmax(g.degrees()))
    counts = {}
    for mode in (1, -1):
        s.init(src)
        counts[mode] = s.run(delta=64.0, heavy_pull=mode).relaxations
        assert np.array_equal(s.get_distances(), oracle_mod.sssp(g, src))
    assert counts[1] < counts[-1], counts


def test_sssp_float_heavy_pull_matches_push(sg, device, oracle_mod):
    # each final distance is the minimum of fl(dist[u] + w) over final
    # dist[u], whichever direction relaxed it: pull and push agree bit for bit
    g = oracle_mod.rmat(11, 8, seed=51)
    w = oracle_mod.fill_weights(g, 5)
    G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=(w / 7.0).astype(np.float32), weight_float=True)
    s = sg.SSSP(G)
    got = {}
    for mode in (1, -1):
        s.init(9)
        s.run(delta=9.0, heavy_pull=mode)  # about half the mean of the weights w / 7
        got[mode] = s.get_distances()
    assert np.array_equal(got[1].view(np.uint32), got[-1].view(np.uint32))
    ref = oracle_mod.sssp(g, 9).astype(np.float64)
    reach = ref != g.n + 1
    assert np.allclose(got[1].astype(np.float64)[reach], ref[reach] / 7.0, rtol=1e-5)


def test_sssp_heavy_pull_directed_pushes(sg, device, oracle_mod):
    # a directed graph's heavy rounds push whatever is asked
    g = _fixture(oracle_mod, "weighted_directed_5")
    G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=g.weights, directed=True)
    s = sg.SSSP(G)
    s.init(0)
    s.run(delta=2.0, heavy_pull=1)
    assert s.get_distances().tolist() == _cases()["expect"]["sssp_weighted_directed_5"]
    with pytest.raises(sg.SygraphError):
        s.run(heavy_pull=2)


@pytest.mark.parametrize("burst_arcs", [-1, 0, 64, 8192])
def test_sssp_bursts_keep_distances(sg, device, oracle_mod, burst_arcs):
    """Small rounds of the reference's schedule run inside one workgroup
    (never, default, tiny and largest thresholds): same distances on a long
    grid, a weighted path, a small RMAT (hubs in small rounds) and an ER graph
    with several components."""
    path = oracle_mod.grid(1, 3000, 0.0, seed=1)
    graphs = [oracle_mod.grid(120, 90, 0.1, seed=53), path, oracle_mod.rmat(11, 16, seed=54),
              oracle_mod.erdos_renyi(4000, 5000, seed=55)]
    for g in graphs:
        oracle_mod.fill_weights(g, 13)
        G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=g.weights)
        s = sg.SSSP(G)
        for src in (0, g.n // 3, int(np.argmax(g.degrees()))):
            s.init(src)
            st = s.run(delta=-1.0, burst_arcs=burst_arcs)
            assert np.array_equal(s.get_distances(), oracle_mod.sssp(g, src)), (g.n, src)
            assert st.rounds >= 1


def test_sssp_float_bursts_match_launched_rounds(sg, device, oracle_mod):
    g = oracle_mod.grid(80, 80, 0.05, seed=56)
    w = oracle_mod.fill_weights(g, 14)
    G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=(w / 3.0).astype(np.float32), weight_float=True)
    s = sg.SSSP(G)
    got = {}
    for ba in (-1, 0):
        s.init(17)
        s.run(delta=-1.0, burst_arcs=ba)
        got[ba] = s.get_distances()
    assert np.array_equal(got[-1].view(np.uint32), got[0].view(np.uint32))
    with pytest.raises(sg.SygraphError):
        s.run(burst_arcs=8193)


def test_sssp_float_weights(sg, device, oracle_mod):
    g = oracle_mod.rmat(11, 8, seed=41)
    w = oracle_mod.fill_weights(g, 5)
    G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=w.astype(np.float32), weight_float=True)
    s = sg.SSSP(G)
    s.init(3)
    s.run()
    assert np.array_equal(s.get_distances(), oracle_mod.sssp(g, 3).astype(np.float32))


def test_sssp_requires_weights(sg, device, oracle_mod):
    g = oracle_mod.rmat(8, 4, seed=1)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    with pytest.raises(sg.SygraphError):
        sg.SSSP(G)


def test_sssp_rmat24_full_size(sg, device, oracle_mod):
    with open(os.path.join(GOLDEN, "rmat24_seed2.json")) as f:
        ref = json.load(f)
    G = sg.Graph.rmat(device, 24, 16, 0.57, 0.19, 0.19, seed=2, weight_seed=3)
    assert [str(x) for x in G.digest()] == ref["digest"]
    s = sg.SSSP(G)
    for rec in ref["sssp"]:
        s.init(rec["source"])
        s.run()
        d = s.get_distances()
        assert str(oracle_mod.digest_i32(d.view(np.int32))) == rec["dist_digest"]
        assert int(d[d != G.n + 1].max()) == rec["max_dist"]


# ------------------------------------------------------------------- CC ---
def test_cc_two_cc_known_answer(sg, device, oracle_mod):
    g = _fixture(oracle_mod, "two_cc")
    cc = sg.CC(sg.Graph.from_csr(device, g.offsets, g.cols))
    cc.init(0)
    cc.run()
    assert cc.get_labels().tolist() == _cases()["expect"]["cc_two_cc"]


def test_cc_matches_oracle(sg, device, oracle_mod):
    for g in [oracle_mod.rmat(14, 16, seed=1), oracle_mod.grid(300, 200, 0.45, seed=3),
              oracle_mod.erdos_renyi(20000, 15000, seed=2)]:
        cc = sg.CC(sg.Graph.from_csr(device, g.offsets, g.cols))
        cc.init()
        cc.run()
        assert np.array_equal(cc.get_labels(), oracle_mod.cc(g))


def test_cc_directed_max_reaching_label(sg, device, oracle_mod):
    # reference semantics on a directed graph: largest id that reaches v
    g = _fixture(oracle_mod, "weighted_directed_5")
    cc = sg.CC(sg.Graph.from_csr(device, g.offsets, g.cols, directed=True))
    cc.init()
    cc.run()
    expect = []
    for v in range(g.n):
        reach = [u for u in range(g.n) if oracle_mod.bfs(g, u)[0][v] != g.n + 1]
        expect.append(max(reach))
    assert cc.get_labels().tolist() == expect


def test_cc_full_size(sg, device, oracle_mod):
    with open(os.path.join(GOLDEN, "c3_cc_tc.json")) as f:
        ref = json.load(f)
    G = sg.Graph.grid(device, 4096, 4096, 0.1, seed=4)
    assert [str(x) for x in G.digest()] == ref["grid"]["digest"]
    cc = sg.CC(G)
    cc.init()
    cc.run()
    lab = cc.get_labels()
    assert str(oracle_mod.digest_i32(lab.view(np.int32))) == ref["grid"]["labels_digest"]
    G = sg.Graph.rmat(device, 20, 32, seed=5)
    cc = sg.CC(G)
    cc.init()
    cc.run()
    lab = cc.get_labels()
    assert str(oracle_mod.digest_i32(lab.view(np.int32))) == ref["rmat20"]["labels_digest"]


# ------------------------------------------------------------------- TC ---
def test_tc_known_answers(sg, device, oracle_mod):
    t = sg.TC(sg.Graph.from_csr(device, *_csr(_fixture(oracle_mod, "triangle_3"))))
    t.init()
    assert t.run() == _cases()["expect"]["tc_triangle_3"]
    n = 40
    full = oracle_mod.from_dense(np.ones((n, n), dtype=np.int64) - np.eye(n, dtype=np.int64))
    assert sg.TC(sg.Graph.from_csr(device, *_csr(full))).run() == n * (n - 1) * (n - 2) // 6
    assert sg.TC(sg.Graph.grid(device, 64, 64)).run() == 0


def _csr(g):
    return g.offsets, g.cols


def test_tc_matches_oracle(sg, device, oracle_mod):
    for g in [oracle_mod.rmat(14, 16, seed=1), oracle_mod.rmat(12, 32, a=0.45, b=0.2, c=0.2, seed=6),
              oracle_mod.erdos_renyi(3000, 40000, seed=3)]:
        t = sg.TC(sg.Graph.from_csr(device, g.offsets, g.cols))
        assert t.run(per_vertex=True) == oracle_mod.tc(g)
        assert int(t.per_vertex.download(np.uint64).sum()) == t.count


def test_tc_per_vertex_at_lowest_corner(sg, device, oracle_mod):
    """Each triangle is counted at its lowest corner in (degree, id) order:
    the count of u is row u of (D @ D) * D summed, D the adjacency matrix
    oriented from lower to higher rank (brute force, independent of the
    kernel's hashing and of which wave takes which vertex)."""
    import scipy.sparse as sp
    for g in [oracle_mod.rmat(11, 16, seed=8), oracle_mod.erdos_renyi(1500, 20000, seed=9)]:
        deg = np.diff(g.offsets).astype(np.int64)
        src = np.repeat(np.arange(g.n), deg)
        dst = g.cols.astype(np.int64)
        up = (deg[src] < deg[dst]) | ((deg[src] == deg[dst]) & (src < dst))
        D = sp.csr_matrix((np.ones(int(up.sum())), (src[up], dst[up])), shape=(g.n, g.n))
        want = np.asarray((D @ D).multiply(D).sum(axis=1)).ravel().astype(np.uint64)
        assert int(want.sum()) == oracle_mod.tc(g)
        t = sg.TC(sg.Graph.from_csr(device, g.offsets, g.cols))
        for _ in range(2):  # the second call reuses the graph's buffers and claim counter
            t.run(per_vertex=True)
            assert np.array_equal(t.per_vertex.download(np.uint64)[:g.n], want)


def test_tc_rmat20_full_size(sg, device, oracle_mod):
    with open(os.path.join(GOLDEN, "c3_cc_tc.json")) as f:
        ref = json.load(f)
    G = sg.Graph.rmat(device, 20, 32, seed=5)
    assert [str(x) for x in G.digest()] == ref["rmat20"]["digest"]
    assert str(sg.TC(G).run()) == ref["rmat20"]["triangles"]


def test_cc_sampling_shapes(sg, device, oracle_mod):
    """Undirected CC links a sample of arcs, takes the most frequent root as the
    giant component and links only the other vertices' remaining arcs: graphs
    without a giant component (a perfect matching, many small stars), with
    two large components of equal size, and with hubs, all give the oracle's
    labels (largest id per component)."""
    rng = np.random.default_rng(48)
    n = 6000
    graphs = []
    perm = rng.permutation(n)
    graphs.append(oracle_mod.csr_from_pairs(n, perm[0::2], perm[1::2]))  # perfect matching
    centres = rng.choice(n, 300, replace=False)
    leaves = np.setdiff1d(np.arange(n), centres)
    graphs.append(oracle_mod.csr_from_pairs(n, centres[rng.integers(0, 300, len(leaves))], leaves))  # stars
    half = n // 2  # two random trees of equal size, interleaved ids
    ids = rng.permutation(n)
    a, b = ids[:half], ids[half:]
    u = np.concatenate([a[1:], b[1:]])
    v = np.concatenate([a[rng.integers(0, np.arange(1, half))], b[rng.integers(0, np.arange(1, half))]])
    graphs.append(oracle_mod.csr_from_pairs(n, u, v))
    graphs.append(oracle_mod.rmat(13, 4, seed=49))  # hubs, many isolated vertices
    for g in graphs:
        G = sg.Graph.from_csr(device, g.offsets, g.cols)
        c = sg.CC(G)
        c.init()
        c.run()
        assert np.array_equal(c.get_labels(), oracle_mod.cc(g))


def test_tc_long_dag_lists_clique(sg, device, oracle_mod):
    """A 1,100-vertex clique joined to a sparse RMAT graph: in the (degree,
    id) orientation the clique's lowest vertices keep more than 1,024
    higher-ranked neighbours, the lists that are binary-searched in global
    memory instead of hashed in LDS.  The clique alone has C(1100, 3)
    triangles, and vertex u of an isolated clique (all degrees equal, so rank
    = id) is the lowest corner of C(1099 - u, 2) of them; joined to the RMAT
    graph, total and per-vertex counts are checked against the brute-force
    lowest-corner count."""
    import scipy.sparse as sp
    from math import comb
    k = 1100
    iu, iv = np.triu_indices(k, 1)
    clique = oracle_mod.csr_from_pairs(k, iu, iv)
    t = sg.TC(sg.Graph.from_csr(device, clique.offsets, clique.cols))
    assert t.run(per_vertex=True) == comb(k, 3)
    want = np.array([comb(k - 1 - u, 2) for u in range(k)], dtype=np.uint64)
    assert np.array_equal(t.per_vertex.download(np.uint64)[:k], want)
    r = oracle_mod.rmat(12, 8, seed=31)
    n = k + r.n
    src = np.repeat(np.arange(r.n), np.diff(r.offsets)) + k
    keep = src < r.cols + k
    rng = np.random.default_rng(31)
    bridge_u = rng.integers(0, k, 3000)  # clique vertices wired into the sparse part
    bridge_v = rng.integers(k, n, 3000)
    g = oracle_mod.csr_from_pairs(n, np.concatenate([iu, src[keep], bridge_u]),
                                  np.concatenate([iv, r.cols[keep].astype(np.int64) + k, bridge_v]))
    deg = np.diff(g.offsets).astype(np.int64)
    s = np.repeat(np.arange(g.n), deg)
    d = g.cols.astype(np.int64)
    up = (deg[s] < deg[d]) | ((deg[s] == deg[d]) & (s < d))
    assert np.bincount(s[up], minlength=g.n).max() > 1024
    D = sp.csr_matrix((np.ones(int(up.sum())), (s[up], d[up])), shape=(g.n, g.n))
    want = np.asarray((D @ D).multiply(D).sum(axis=1)).ravel().astype(np.uint64)
    assert int(want.sum()) == oracle_mod.tc(g)
    t = sg.TC(sg.Graph.from_csr(device, g.offsets, g.cols))
    assert t.run(per_vertex=True) == int(want.sum())
    assert np.array_equal(t.per_vertex.download(np.uint64)[:g.n], want)


def _max_reaching(n, src, dst):
    """Fixpoint of the reference's propagation (cc.hpp:103-169) by numpy
    sweeps: labels[v] = largest id with a directed path to v."""
    lab = np.arange(n, dtype=np.int64)
    while True:
        new = lab.copy()
        np.maximum.at(new, dst, lab[src])
        if np.array_equal(new, lab):
            return lab.astype(np.uint32)
        lab = new


def test_cc_asymmetric_csr_without_directed_flag(sg, device, oracle_mod):
    """A CSR that lacks some reverse arcs but is not flagged directed gets the
    reference's labels (largest id reaching each vertex along out-arcs), not
    its weak components."""
    rng = np.random.default_rng(77)
    n, m = 4000, 9000
    u, v = rng.integers(0, n, m), rng.integers(0, n, m)
    keep = u != v
    pairs = np.unique(np.stack([u[keep], v[keep]], 1), axis=0)
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(pairs[:, 0], minlength=n), out=off[1:])
    cols = np.ascontiguousarray(pairs[:, 1], dtype=np.int32)
    want = _max_reaching(n, pairs[:, 0], pairs[:, 1])
    weak = oracle_mod.cc(oracle_mod.CSRGraph(n, off, cols))
    assert not np.array_equal(want, weak)  # the two definitions differ on this graph
    G = sg.Graph.from_csr(device, off, cols)  # directed=False
    c = sg.CC(G)
    c.init()
    c.run()
    assert np.array_equal(c.get_labels(), want)
    # its symmetrised form still takes the union-find path: weak components
    g = oracle_mod.csr_from_pairs(n, pairs[:, 0], pairs[:, 1])
    c = sg.CC(sg.Graph.from_csr(device, g.offsets, g.cols))
    c.init()
    c.run()
    assert np.array_equal(c.get_labels(), oracle_mod.cc(g))
