"""Betweenness centrality and multi-source BFS on the MI355X through the C ABI.

BC results are compared with the FP64 oracle (itself pinned against networkx
for undirected and directed graphs) at relative tolerance 1e-5 of the FP64
value, the bound stated for BC; the float output rounds the FP64 total once,
so the observed error is ~6e-8 relative.  Known answers: the reference's
line_5 fixture (tests/algorithms/bc.cpp:9-15 runs BC from source 0 on it) and
star closed forms.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
RTOL = 1e-5


def _cases():
    with open(os.path.join(GOLDEN, "reference_cases.json")) as f:
        return json.load(f)


def _fixture(oracle_mod, name):
    return oracle_mod.from_dense(oracle_mod.parse_matrix_text(_cases()["matrices"][name]))


def _bc(sg, G, sources, batched=False):
    b = sg.BC(G)
    b.init(sources)
    b.run(batched=batched)
    return b.get_values()


def _close(got, ref):
    got = got.astype(np.float64)
    return np.all(np.abs(got - ref) <= RTOL * np.abs(ref) + 1e-6)


@pytest.mark.parametrize("batched", [False, True])
def test_bc_line5_known_answer(sg, device, oracle_mod, batched):
    g = _fixture(oracle_mod, "line_5")
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    assert _bc(sg, G, [0], batched).tolist() == _cases()["expect"]["bc_line_5_source0"]


@pytest.mark.parametrize("batched", [False, True])
def test_bc_star_closed_forms(sg, device, oracle_mod, batched):
    n = 9
    adj = np.zeros((n, n), dtype=np.int64)
    adj[0, 1:] = adj[1:, 0] = 1
    G = sg.Graph.from_csr(device, *_csr(oracle_mod.from_dense(adj)))
    assert np.all(_bc(sg, G, [0], batched) == 0.0)       # from the centre nothing passes through
    b = _bc(sg, G, [1], batched)                         # from a leaf the centre carries n-2 targets
    assert b[0] == n - 2 and np.all(b[1:] == 0.0)


def _csr(g):
    return g.offsets, g.cols


@pytest.mark.parametrize("batched", [False, True])
def test_bc_rmat14_golden(sg, device, oracle_mod, batched):
    z = np.load(os.path.join(GOLDEN, "rmat14_seed1.npz"))
    G = sg.Graph.rmat(device, 14, 16, seed=1)
    assert _close(_bc(sg, G, z["bc_sources"], batched), z["bc"])


@pytest.mark.parametrize("batched", [False, True])
def test_bc_matches_oracle_shapes(sg, device, oracle_mod, batched):
    rng = np.random.default_rng(11)
    for g in [oracle_mod.rmat(12, 16, seed=9), oracle_mod.grid(60, 45, 0.15, seed=8),
              oracle_mod.erdos_renyi(3000, 6000, seed=7)]:
        srcs = rng.integers(0, g.n, size=5).tolist()
        G = sg.Graph.from_csr(device, g.offsets, g.cols)
        assert _close(_bc(sg, G, srcs, batched), oracle_mod.bc(g, srcs))


@pytest.mark.parametrize("batched", [False, True])
def test_bc_long_diameter_grid_stays_finite(sg, device, oracle_mod, batched):
    # ~1e116 shortest paths to the far corner: sigma needs the per-level scaling
    g = oracle_mod.grid(200, 200)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    b = _bc(sg, G, [0], batched)
    assert np.all(np.isfinite(b))
    assert _close(b, oracle_mod.bc(g, [0]))


@pytest.mark.parametrize("batched", [False, True])
def test_bc_directed(sg, device, oracle_mod, batched):
    rng = np.random.default_rng(5)
    n, m = 2000, 9000
    u, v = rng.integers(0, n, m), rng.integers(0, n, m)
    keep = u != v
    pairs = np.unique(np.stack([u[keep], v[keep]], 1), axis=0)
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(pairs[:, 0], minlength=n), out=off[1:])
    g = oracle_mod.CSRGraph(n, off, np.ascontiguousarray(pairs[:, 1], dtype=np.int32))
    G = sg.Graph.from_csr(device, g.offsets, g.cols, directed=True)
    srcs = [0, 3, 99, 1234]
    assert _close(_bc(sg, G, srcs, batched), oracle_mod.bc(g, srcs, inverse=oracle_mod.transpose(g)))


@pytest.mark.parametrize("batched", [False, True])
def test_bc_accumulates_and_isolated_source(sg, device, oracle_mod, batched):
    g = oracle_mod.grid(30, 30, 0.3, seed=2)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    iso = int(np.flatnonzero(np.diff(g.offsets) == 0)[0])
    b = sg.BC(G)
    b.init([5, iso])
    b.run(batched=batched)
    b.run(batched=batched)  # a second run adds the same dependencies again
    assert _close(b.get_values(), 2.0 * oracle_mod.bc(g, [5, iso]))


@pytest.mark.parametrize("batched", [False, True])
def test_bc_invalid_source(sg, device, oracle_mod, batched):
    g = oracle_mod.rmat(8, 4, seed=1)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    b = sg.BC(G)
    b.init([g.n])
    with pytest.raises(sg.SygraphError):
        b.run(batched=batched)


def test_bc_batched_many_sources_and_hubs(sg, device, oracle_mod):
    """Two lane batches (100 sources, repeats included), vertices whose lists
    span several work chunks (degrees above 256) and a star centre of degree
    599: the batched sums equal the per-source ones up to FP64 rounding."""
    rng = np.random.default_rng(23)
    g = oracle_mod.rmat(12, 16, seed=31)
    assert g.degrees().max() > 3 * 256
    srcs = rng.integers(0, g.n, size=98).tolist() + [int(np.argmax(g.degrees()))] * 2
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    ref = oracle_mod.bc(g, srcs)
    batched, single = _bc(sg, G, srcs, True), _bc(sg, G, srcs, False)
    assert _close(batched, ref) and _close(single, ref)
    n = 600
    adj = np.zeros((n, n), dtype=np.int64)
    adj[0, 1:] = adj[1:, 0] = 1
    G = sg.Graph.from_csr(device, *_csr(oracle_mod.from_dense(adj)))
    b = _bc(sg, G, list(range(1, 71)), True)
    assert b[0] == 70 * (n - 2) and np.all(b[1:] == 0.0)


def test_bfs_multi_matches_oracle(sg, device, oracle_mod):
    g = oracle_mod.rmat(12, 8, seed=13)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    srcs = [0, 1, 77, int(np.argmax(g.degrees()))]
    d = sg.bfs_multi(G, srcs)
    for i, s in enumerate(srcs):
        assert np.array_equal(d[i], oracle_mod.bfs(g, s)[0])
