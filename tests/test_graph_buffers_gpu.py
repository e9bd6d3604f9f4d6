"""Device buffers kept with a graph across calls (CC, TC, MS-BFS, batched BC).

These algorithms keep their working buffers with the graph, grown when a call
needs more and reused otherwise.  Each call below is compared with the oracle,
so stale contents left by an earlier call on the same graph (a larger source
batch, a deeper BFS, another algorithm sharing a buffer name) would show up as
a mismatch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _bc(sg, G, sources):
    b = sg.BC(G)
    b.init(sources)
    b.run(batched=True)
    return b.get_values().astype(np.float64)


def _close(got, ref):
    return np.all(np.abs(got - ref) <= RTOL * np.abs(ref) + 1e-6)


def test_repeated_calls_on_one_graph(sg, device, oracle_mod):
    g = oracle_mod.rmat(12, 16, seed=21)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    rng = np.random.default_rng(5)
    cc_ref = oracle_mod.cc(g)
    tc_ref = oracle_mod.tc(g)
    for sources in [rng.integers(0, g.n, size=3).tolist(),
                    rng.integers(0, g.n, size=130).tolist(),   # three 64-wide batches
                    [int(np.argmax(g.degrees()))]]:
        assert _close(_bc(sg, G, sources), oracle_mod.bc(g, sources))
        d = sg.bfs_multi(G, sources[:64])
        for i, s in enumerate(sources[:64]):
            assert np.array_equal(d[i], oracle_mod.bfs(g, s)[0])
        cc = sg.CC(G)
        cc.init()
        cc.run()
        assert np.array_equal(cc.get_labels(), cc_ref)
        assert sg.TC(G).run() == tc_ref


def _union(oracle_mod, a, b):
    """Disjoint union: b's vertices follow a's."""
    off = np.concatenate([a.offsets, a.offsets[-1] + b.offsets[1:]])
    cols = np.concatenate([a.cols, b.cols + np.int32(a.n)]).astype(np.int32)
    return oracle_mod.CSRGraph(a.n + b.n, off, cols)


def test_buffers_grow_with_depth(sg, device, oracle_mod):
    # one graph with a shallow RMAT part and a long-diameter grid part:
    # sources in the RMAT part first, then in the grid (the batched BC's
    # per-level masks and scale table grow), then the RMAT part again, which
    # reuses the larger buffers
    a = oracle_mod.rmat(11, 8, seed=3)
    g = _union(oracle_mod, a, oracle_mod.grid(40, 60, 0.0, seed=4))
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    for srcs in [[0, 5, 9], [a.n, a.n + 1234, g.n - 1], [0, 5, 9], [1, a.n + 7]]:
        assert _close(_bc(sg, G, srcs), oracle_mod.bc(g, srcs))
        d = sg.bfs_multi(G, srcs)
        for i, s in enumerate(srcs):
            assert np.array_equal(d[i], oracle_mod.bfs(g, s)[0])
