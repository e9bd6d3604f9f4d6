"""Multi-source BFS (64 sources per bit-lane batch) against one oracle BFS per
source: depths must be bit-exact for every (source, vertex) pair."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(sg, G, g, oracle_mod, srcs, inverse=None):
    d = sg.bfs_multi(G, srcs)
    assert d.shape == (len(srcs), g.n)
    for i, s in enumerate(srcs):
        assert np.array_equal(d[i], oracle_mod.bfs(g, s)[0]), f"source {s} (row {i})"


def test_msbfs_full_batch_rmat(sg, device, oracle_mod):
    g = oracle_mod.rmat(12, 16, seed=21)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    srcs = np.random.default_rng(1).integers(0, g.n, size=64).tolist()
    _check(sg, G, g, oracle_mod, srcs)


def test_msbfs_several_batches_and_duplicates(sg, device, oracle_mod):
    g = oracle_mod.rmat(11, 8, seed=22)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    rng = np.random.default_rng(2)
    srcs = rng.integers(0, g.n, size=150).tolist()
    srcs[5] = srcs[6] = srcs[70]  # the same source in several lanes and batches
    srcs[9] = int(np.flatnonzero(np.diff(g.offsets) == 0)[0])  # an isolated vertex
    _check(sg, G, g, oracle_mod, srcs)


def test_msbfs_long_diameter_grid(sg, device, oracle_mod):
    g = oracle_mod.grid(70, 90, 0.1, seed=5)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    _check(sg, G, g, oracle_mod, [0, 1, 6299, 3150, 77, 5000, 123])


def test_msbfs_directed(sg, device, oracle_mod):
    rng = np.random.default_rng(9)
    n, m = 3000, 12000
    u, v = rng.integers(0, n, m), rng.integers(0, n, m)
    keep = u != v
    pairs = np.unique(np.stack([u[keep], v[keep]], 1), axis=0)
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(pairs[:, 0], minlength=n), out=off[1:])
    g = oracle_mod.CSRGraph(n, off, np.ascontiguousarray(pairs[:, 1], dtype=np.int32))
    G = sg.Graph.from_csr(device, g.offsets, g.cols, directed=True)
    _check(sg, G, g, oracle_mod, rng.integers(0, n, size=40).tolist())


def test_msbfs_invalid_source(sg, device, oracle_mod):
    g = oracle_mod.rmat(8, 4, seed=1)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    with pytest.raises(sg.SygraphError):
        sg.bfs_multi(G, [0, g.n])


@pytest.mark.parametrize("graph", ["rmat", "grid", "directed"])
def test_msbfs_direction_rules_same_depths(sg, device, oracle_mod, graph):
    """Every per-level direction rule of a batch -- always push, always pull,
    the reference's alpha / beta rule and the learned policy on the batch's
    totals, the default n/20 rule -- gives the oracle's depths."""
    import os
    if graph == "rmat":
        g = oracle_mod.rmat(13, 16, seed=24)
        directed = False
    elif graph == "grid":
        g = oracle_mod.grid(80, 70, 0.1, seed=25)
        directed = False
    else:
        rng = np.random.default_rng(26)
        n, m = 4000, 24000
        u, v = rng.integers(0, n, m), rng.integers(0, n, m)
        keep = u != v
        pairs = np.unique(np.stack([u[keep], v[keep]], 1), axis=0)
        off = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(np.bincount(pairs[:, 0], minlength=n), out=off[1:])
        g = oracle_mod.CSRGraph(n, off, np.ascontiguousarray(pairs[:, 1], dtype=np.int32))
        directed = True
    G = sg.Graph.from_csr(device, g.offsets, g.cols, directed=directed)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pol = sg.Policy(device, np.fromfile(os.path.join(root, "sygraph_amd", "data", "bfs_policy.f32"),
                                        dtype=np.float32))
    srcs = np.random.default_rng(27).choice(np.nonzero(g.degrees() > 0)[0], 40, replace=False).tolist()
    ref = [oracle_mod.bfs(g, s)[0] for s in srcs]
    for policy in (sg.MSBFS_AUTO, sg.POLICY_PUSH, sg.POLICY_PULL, sg.POLICY_HYBRID, sg.POLICY_LEARNED):
        d = sg.bfs_multi(G, srcs, policy=policy, learned=pol if policy == sg.POLICY_LEARNED else None)
        for i in range(len(srcs)):
            assert np.array_equal(d[i], ref[i]), (policy, srcs[i])
    with pytest.raises(sg.SygraphError):
        sg.bfs_multi(G, srcs, policy=sg.POLICY_LEARNED)  # no trained policy given
