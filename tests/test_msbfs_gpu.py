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
