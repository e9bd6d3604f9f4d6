"""BFS on the MI355X through the C ABI, bit-exact against the oracle.

Covers the reference's BFS known answers (tests/algorithms/bfs.cpp,
tests/operators/advance.cpp), every direction policy, directed graphs (pull
over the inverse CSR), isolated sources, high-diameter grids, the hybrid
direction schedule (algorithms/bfs.hpp:325-350) and, at full size, RMAT
scale 24 through committed digests.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
POLICIES = ["push", "pull", "hybrid"]


def _cases():
    with open(os.path.join(GOLDEN, "reference_cases.json")) as f:
        return json.load(f)


def _upload(sg, device, g, directed=False):
    return sg.Graph.from_csr(device, g.offsets, g.cols, directed=directed)


def _run(sg, G, src, policy, alpha=15.0, beta=24.0):
    b = sg.BFS(G)
    b.init(src)
    pol = {"push": sg.POLICY_PUSH, "pull": sg.POLICY_PULL, "hybrid": sg.POLICY_HYBRID}[policy]
    st = b.run(pol, alpha, beta)
    return b.get_distances(), st, b.directions()


@pytest.mark.parametrize("policy", POLICIES)
def test_line5_known_answer(sg, device, oracle_mod, policy):
    c = _cases()
    g = oracle_mod.from_dense(oracle_mod.parse_matrix_text(c["matrices"]["line_5"]))
    G = _upload(sg, device, g)
    d, st, dirs = _run(sg, G, 0, policy, 1.0, 1.0)
    assert d.tolist() == c["expect"]["bfs_line_5"]
    assert st.levels == c["expect"]["bfs_line_5_iterations"]
    if policy == "push":
        assert set(dirs) == {"push"}
    if policy == "pull":
        assert set(dirs) == {"pull"}


@pytest.mark.parametrize("policy", POLICIES)
def test_symmetric6_known_answer(sg, device, oracle_mod, policy):
    c = _cases()
    g = oracle_mod.from_dense(oracle_mod.parse_matrix_text(c["matrices"]["symmetric_6nodes"]))
    d, _, _ = _run(sg, _upload(sg, device, g), 0, policy)
    assert d.tolist() == c["expect"]["advance_bfs_symmetric_6nodes"]


@pytest.mark.parametrize("policy", POLICIES)
def test_isolated_source(sg, device, oracle_mod, policy):
    c = _cases()
    g = oracle_mod.from_dense(oracle_mod.parse_matrix_text(c["matrices"]["two_cc"]))
    d, st, _ = _run(sg, _upload(sg, device, g), 5, policy)
    assert d.tolist() == [7, 7, 7, 7, 7, 0] and st.levels == 1


@pytest.mark.parametrize("policy", POLICIES)
def test_directed_graph(sg, device, oracle_mod, policy):
    c = _cases()
    g = oracle_mod.from_dense(oracle_mod.parse_matrix_text(c["matrices"]["weighted_directed_5"]))
    G = _upload(sg, device, g, directed=True)
    for src in range(5):
        ref, lv = oracle_mod.bfs(g, src)
        d, st, _ = _run(sg, G, src, policy)
        assert np.array_equal(d, ref) and st.levels == lv


@pytest.fixture(scope="module")
def rmat14(sg, device, oracle_mod):
    z = np.load(os.path.join(GOLDEN, "rmat14_seed1.npz"))
    G = sg.Graph.rmat(device, 14, 16, 0.57, 0.19, 0.19, seed=1)
    return G, z


def test_rmat14_generator_matches_oracle(sg, rmat14):
    G, z = rmat14
    assert G.n == int(z["n"]) and G.m == int(z["m"])
    assert list(G.digest()) == [int(x) for x in z["digest"]]


@pytest.mark.parametrize("policy", POLICIES)
def test_rmat14_golden_depths(sg, oracle_mod, rmat14, policy):
    G, z = rmat14
    d, st, _ = _run(sg, G, int(z["source"]), policy)
    assert np.array_equal(d, z["depth"])
    assert st.levels == int(z["levels"])


def test_rmat14_hybrid_schedule_matches_reference_rule(sg, oracle_mod, rmat14):
    G, z = rmat14
    off, cols, _ = G.download()
    g = oracle_mod.CSRGraph(G.n, off, cols)
    for src in [int(z["source"]), 1, 777, 9000]:
        for alpha, beta in [(15.0, 24.0), (1.0, 1.0), (4.0, 100.0)]:
            ref, lv, ref_dirs = oracle_mod.bfs_plain_omp(g, src, oracle_mod.BFS_HYBRID, alpha, beta, with_dirs=True)
            d, st, dirs = _run(sg, G, src, "hybrid", alpha, beta)
            assert np.array_equal(d, ref)
            assert dirs == ["push" if x == 0 else "pull" for x in ref_dirs]


def test_rmat14_level_counters_match_oracle(sg, oracle_mod, rmat14):
    """Per-level discovered counts and the pull scan cost sum k_u (the byte model)."""
    G, z = rmat14
    off, cols, _ = G.download()
    g = oracle_mod.CSRGraph(G.n, off, cols)
    src = int(z["source"])
    ref = z["depth"]
    hint = oracle_mod.pull_hint(g)
    for policy in ["push", "pull"]:
        _, st, _ = _run(sg, G, src, policy)
        for L in range(st.levels):
            ls = st.level[L]
            assert ls.discovered == int((ref == L + 1).sum())
            if policy == "pull":
                tot, cnt, miss = oracle_mod.pull_scan_count(g, ref, L, hint)
                assert (ls.edges_inspected, ls.vertices_scanned, ls.hint_misses) == (tot, cnt, miss)
            else:
                fr = ref == L
                assert ls.frontier_vertices == int(fr.sum())
                assert ls.edges_inspected == int(g.degrees()[fr].sum())


@pytest.mark.parametrize("policy", POLICIES)
def test_random_sources_and_shapes(sg, device, oracle_mod, policy):
    graphs = [
        oracle_mod.rmat(12, 8, seed=21),
        oracle_mod.rmat(11, 32, a=0.45, b=0.22, c=0.22, seed=22),
        oracle_mod.grid(64, 80, 0.15, seed=23),
        oracle_mod.erdos_renyi(5000, 9000, seed=24),
    ]
    rng = np.random.default_rng(5)
    for g in graphs:
        G = _upload(sg, device, g)
        for src in rng.integers(0, g.n, size=4):
            ref, lv = oracle_mod.bfs(g, int(src))
            d, st, _ = _run(sg, G, int(src), policy)
            assert np.array_equal(d, ref), (g.n, int(src))
            assert st.levels == lv


def test_generators_match_oracle(sg, device, oracle_mod):
    cases = [
        (lambda: sg.Graph.grid(device, 100, 73, 0.1, seed=4), lambda: oracle_mod.grid(100, 73, 0.1, seed=4)),
        (lambda: sg.Graph.erdos_renyi(device, 7000, 30000, seed=9), lambda: oracle_mod.erdos_renyi(7000, 30000, 9)),
        (lambda: sg.Graph.rmat(device, 13, 8, 0.45, 0.25, 0.2, seed=3, relabel=False),
         lambda: oracle_mod.rmat(13, 8, 0.45, 0.25, 0.2, seed=3, relabel=False)),
    ]
    for mk_gpu, mk_cpu in cases:
        G, g = mk_gpu(), mk_cpu()
        assert (G.n, G.m) == (g.n, g.m)
        assert list(G.digest()) == list(g.digest())


def test_weights_match_oracle(sg, device, oracle_mod):
    G = sg.Graph.rmat(device, 12, 8, seed=5, weight_seed=3)
    off, cols, w = G.download(weights=True)
    g = oracle_mod.CSRGraph(G.n, off, cols)
    assert np.array_equal(w, oracle_mod.fill_weights(g, 3))


def test_repeatable_and_reusable(sg, oracle_mod, rmat14):
    G, z = rmat14
    b = sg.BFS(G)
    for _ in range(3):
        b.init(int(z["source"]))
        b.run(sg.POLICY_HYBRID, 15.0, 24.0)
        assert np.array_equal(b.get_distances(), z["depth"])


def test_invalid_source_raises(sg, rmat14):
    G, _ = rmat14
    b = sg.BFS(G)
    b.init(G.n + 5)
    with pytest.raises(sg.SygraphError):
        b.run()


def test_rmat24_full_size(sg, device, oracle_mod):
    """RMAT scale 24 at full size: generator digest, then BFS from the 16 committed
    sources under every policy against the committed per-source depth digests."""
    with open(os.path.join(GOLDEN, "rmat24_seed2.json")) as f:
        ref = json.load(f)
    G = sg.Graph.rmat(device, 24, 16, 0.57, 0.19, 0.19, seed=2)
    assert (G.n, G.m) == (ref["n"], ref["m"])
    assert [str(x) for x in G.digest()] == ref["digest"]
    b = sg.BFS(G)
    for rec in ref["bfs"]:
        for policy in POLICIES:
            b.init(rec["source"])
            pol = {"push": sg.POLICY_PUSH, "pull": sg.POLICY_PULL, "hybrid": sg.POLICY_HYBRID}[policy]
            st = b.run(pol, 15.0, 24.0)
            d = b.get_distances()
            assert str(oracle_mod.digest_i32(d)) == rec["depth_digest"], (rec["source"], policy)
            assert st.levels == rec["levels"]
            e, v = G.reached_edges(b.depth)
            assert (e, v) == (rec["reached_edges"], rec["reached_vertices"])


@pytest.mark.parametrize("policy", POLICIES)
def test_lookahead_matches_level_by_level(sg, device, oracle_mod, policy):
    """Levels enqueued ahead of the host (device-side decisions) give the same
    depths, schedule and per-level counters as waiting for every level."""
    g = oracle_mod.rmat(13, 16, seed=17)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    pol = {"push": sg.POLICY_PUSH, "pull": sg.POLICY_PULL, "hybrid": sg.POLICY_HYBRID}[policy]
    src = int(np.argmax(g.degrees()))
    runs = []
    for sync in (False, True):
        b = sg.BFS(G)
        b.init(src)
        st = b.run(pol, 15.0, 24.0, sync_per_level=sync)
        lv = [(st.level[i].direction, st.level[i].discovered, st.level[i].edges_inspected,
               st.level[i].vertices_scanned, st.level[i].unexplored_edges) for i in range(st.levels)]
        runs.append((b.get_distances(), st.levels, st.reached_edges, lv))
    assert np.array_equal(runs[0][0], runs[1][0]) and runs[0][1:] == runs[1][1:]
    assert np.array_equal(runs[0][0], oracle_mod.bfs(g, src)[0])


def _directed_rmat(oracle_mod, scale, ef, seed):
    """RMAT arcs kept one-way (no symmetrization): sorted, deduplicated, no loops."""
    n = 1 << scale
    u, v = oracle_mod.rmat_pairs(scale, ef, 0.57, 0.19, 0.19, seed)
    key = np.unique(u.astype(np.int64)[u != v] * n + v.astype(np.int64)[u != v])
    off = np.zeros(n + 1, dtype=np.int64)
    np.add.at(off, key // n + 1, 1)
    return oracle_mod.CSRGraph(n, np.cumsum(off), (key % n).astype(np.int32))


@pytest.mark.parametrize("graph", ["grid", "path", "rmat", "rmat_directed"])
@pytest.mark.parametrize("policy", ["push", "hybrid"])
@pytest.mark.parametrize("timed", [True, False])
def test_bursts_leave_depths_and_levels_unchanged(sg, device, oracle_mod, graph, policy, timed):
    """Small push levels run back to back in one single-workgroup launch; every
    threshold (off, tiny, default, maximal) gives the oracle's depths and the
    same per-level records as running every level on its own."""
    g = {"grid": lambda: oracle_mod.grid(120, 90, 0.15, seed=8),
         "path": lambda: oracle_mod.grid(1, 3000),
         "rmat": lambda: oracle_mod.rmat(13, 8, seed=19),
         "rmat_directed": lambda: _directed_rmat(oracle_mod, 12, 8, 21)}[graph]()
    directed = graph == "rmat_directed"
    G = sg.Graph.from_csr(device, g.offsets, g.cols, directed=directed)
    pol = sg.POLICY_PUSH if policy == "push" else sg.POLICY_HYBRID
    srcs = [int(np.argmax(g.degrees())), int(np.nonzero(g.degrees() > 0)[0][-1])]
    for src in srcs:
        ref, levels = oracle_mod.bfs(g, src)
        runs = {}
        for burst in (-1, 64, 0, 65536):
            b = sg.BFS(G)
            b.init(src)
            st = b.run(pol, 15.0, 24.0, tuning={"burst_arcs": burst, "grid_burst_arcs": -1}, record_levels=timed)
            assert np.array_equal(b.get_distances(), ref), (src, burst)
            assert st.levels == levels
            lv = [(st.level[i].direction, st.level[i].frontier_vertices, st.level[i].frontier_edges,
                   st.level[i].discovered, st.level[i].unexplored_edges, st.level[i].max_degree,
                   st.level[i].action) for i in range(st.levels)]
            runs[burst] = ((st.n_push, st.n_pull, st.reached_vertices, st.reached_edges), lv,
                           sum(st.level[i].burst for i in range(st.levels)))
        assert runs[-1][2] == 0
        for burst in (64, 0, 65536):
            assert runs[burst][:2] == runs[-1][:2], (src, burst)
        if graph in ("grid", "path"):
            assert runs[0][2] >= levels // 2  # most levels of a high-diameter graph run inside bursts
    with pytest.raises(sg.SygraphError):
        b.run(pol, 15.0, 24.0, tuning={"burst_arcs": 65537})


@pytest.mark.parametrize("graph", ["grid", "path", "rmat", "rmat_directed", "star_chain"])
@pytest.mark.parametrize("policy", ["push", "hybrid"])
@pytest.mark.parametrize("timed", [True, False])
def test_grid_bursts_leave_depths_and_levels_unchanged(sg, device, oracle_mod, graph, policy, timed):
    """Push levels between the single-workgroup burst's limit and
    grid_burst_arcs run back to back in one cooperative launch with a grid
    barrier per level.  Every combination of thresholds gives the oracle's
    depths and the per-level records of running each level on its own; on the
    high-diameter graphs most levels run inside grid bursts when the
    single-workgroup bursts are off.  star_chain: hubs of 3,000 arcs on a long
    path, so grid bursts cut long lists into pieces.  timed: each step timed
    by events (the host waits on them) or not (the host polls each step's
    record, published once by the step's last writer)."""
    def star_chain():
        hubs, leaves = 40, 3000
        u, v = [], []
        for h in range(hubs):
            c = h * (leaves + 1)
            if h:
                u.append(c - leaves - 1)
                v.append(c)
            u.extend([c] * leaves)
            v.extend(range(c + 1, c + 1 + leaves))
        return oracle_mod.csr_from_pairs(hubs * (leaves + 1), np.array(u), np.array(v))
    g = {"grid": lambda: oracle_mod.grid(120, 90, 0.15, seed=8),
         "path": lambda: oracle_mod.grid(1, 3000),
         "rmat": lambda: oracle_mod.rmat(13, 8, seed=19),
         "rmat_directed": lambda: _directed_rmat(oracle_mod, 12, 8, 21),
         "star_chain": star_chain}[graph]()
    directed = graph == "rmat_directed"
    G = sg.Graph.from_csr(device, g.offsets, g.cols, directed=directed)
    pol = sg.POLICY_PUSH if policy == "push" else sg.POLICY_HYBRID
    srcs = [int(np.argmax(g.degrees())), int(np.nonzero(g.degrees() > 0)[0][-1])]
    configs = [(-1, -1), (-1, 0), (64, 0), (-1, 128), (0, 0), (-1, 1 << 24)]
    for src in srcs:
        ref, levels = oracle_mod.bfs(g, src)
        runs = {}
        for burst, grid in configs:
            b = sg.BFS(G)
            b.init(src)
            st = b.run(pol, 15.0, 24.0, tuning={"burst_arcs": burst, "grid_burst_arcs": grid}, record_levels=timed)
            assert np.array_equal(b.get_distances(), ref), (src, burst, grid)
            assert st.levels == levels
            lv = [(st.level[i].direction, st.level[i].frontier_vertices, st.level[i].frontier_edges,
                   st.level[i].discovered, st.level[i].unexplored_edges, st.level[i].max_degree,
                   st.level[i].action) for i in range(min(st.levels, sg.MAX_LEVELS_RECORDED))]
            runs[(burst, grid)] = ((st.n_push, st.n_pull, st.reached_vertices, st.reached_edges), lv,
                                   sum(1 for i in range(min(st.levels, sg.MAX_LEVELS_RECORDED))
                                       if st.level[i].burst == 2))
        assert runs[(-1, -1)][2] == 0
        for c in configs[1:]:
            assert runs[c][:2] == runs[(-1, -1)][:2], (src, c)
        if graph in ("grid", "path") and policy == "push":
            assert runs[(-1, 0)][2] >= min(levels, sg.MAX_LEVELS_RECORDED) // 2
    with pytest.raises(sg.SygraphError):
        b.run(pol, 15.0, 24.0, tuning={"grid_burst_arcs": (1 << 24) + 1})


def test_bursts_with_fine_degree_classes(sg, device, oracle_mod):
    """Bursts hand their last discovered set back as queues: with small class
    thresholds and chunks, short and long lists both become chunked (class 2)
    entries, written per thread (short lists) and per wave (long lists)."""
    g = oracle_mod.rmat(13, 16, seed=23)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    for src in [int(np.argmax(g.degrees())), 17, 4000]:
        ref, levels = oracle_mod.bfs(g, src)
        for tuning in ({"t_medium": 8, "t_large": 16, "chunk": 64, "burst_arcs": 65536},
                       {"t_medium": 8, "t_large": 16, "chunk": 64, "burst_arcs": 512}):
            b = sg.BFS(G)
            b.init(src)
            st = b.run(sg.POLICY_HYBRID, 15.0, 24.0, tuning=tuning)
            assert np.array_equal(b.get_distances(), ref), (src, tuning)
            assert st.levels == levels


@pytest.mark.parametrize("policy", POLICIES)
def test_bfs_many_rows_match_oracle(sg, device, oracle_mod, policy):
    """sg_bfs_many: the sources' traversals back to back without a host
    synchronisation between them; row i equals the oracle's depths from
    sources[i], the totals are the sum of the rows' reached vertices and
    their arcs, and a bad source is refused before anything runs."""
    for g in (oracle_mod.rmat(12, 16, seed=41), oracle_mod.grid(60, 200, 0.1, seed=42)):
        G = sg.Graph.from_csr(device, g.offsets, g.cols)
        deg = g.degrees()
        srcs = [int(x) for x in np.nonzero(deg > 0)[0][[0, 5, -1]]] + [int(np.argmax(deg))]
        pol = {"push": sg.POLICY_PUSH, "pull": sg.POLICY_PULL, "hybrid": sg.POLICY_HYBRID}[policy]
        st = sg.bfs_many(G, srcs, policy=pol)
        rows = st.depth.download(np.int32, len(srcs) * G.n).reshape(len(srcs), G.n)
        rv = re = 0
        for i, s in enumerate(srcs):
            ref, _ = oracle_mod.bfs(g, s)
            assert np.array_equal(rows[i], ref), (i, s)
            hit = ref != G.n + 1
            rv += int(hit.sum())
            re += int(deg[hit].sum())
        assert (st.reached_vertices, st.reached_edges) == (rv, re)
        assert st.kernel_ms > 0
        empty = sg.bfs_many(G, [], policy=pol)
        assert empty.reached_vertices == 0
        with pytest.raises(sg.SygraphError):
            sg.bfs_many(G, [0, G.n], policy=pol)
        G.destroy()
