"""sg_cc_noalloc (CC with every buffer beyond the labels in the library's
static scratch array) against sg_cc, through the C ABI: identical labels on
two calls in a row, the second on a smaller graph, for an undirected RMAT
graph (union-find path), a directed graph and an undirected-flagged CSR
missing reverse arcs (label propagation path).  Collected last: the array
keeps the last call's contents, and the calls must not depend on them."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _labels(sg, G, fn):
    buf = G.dev.alloc(4 * max(G.n, 1))
    r = C.c_int32()
    sg.check(fn(G.h, C.c_void_p(buf.ptr), C.byref(r)))
    out = buf.download(np.uint32, G.n)
    buf.free()
    return out, r.value


def _directed(oracle_mod, n, m, seed):
    rng = np.random.default_rng(seed)
    u, v = rng.integers(0, n, m), rng.integers(0, n, m)
    keep = u != v
    pairs = np.unique(np.stack([u[keep], v[keep]], 1), axis=0)
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(pairs[:, 0], minlength=n), out=off[1:])
    return oracle_mod.CSRGraph(n, off, np.ascontiguousarray(pairs[:, 1], dtype=np.int32))


def test_cc_noalloc_matches_cc_twice_undirected(sg, device, oracle_mod):
    lib = sg.lib()
    for scale, ef, seed in ((16, 8, 41), (12, 4, 42)):  # the second call on a smaller graph
        G = sg.Graph.rmat(device, scale, ef, seed=seed)
        ref, _ = _labels(sg, G, lib.sg_cc)
        got, rounds = _labels(sg, G, lib.sg_cc_noalloc)
        assert np.array_equal(got, ref), f"RMAT s{scale}: labels differ"
        assert rounds == 1
        G.destroy()
    g = oracle_mod.grid(60, 70, 0.3, seed=7)  # many components, none dominant
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    got, _ = _labels(sg, G, lib.sg_cc_noalloc)
    assert np.array_equal(got, oracle_mod.cc(g)), "grid: labels differ from the oracle"


def test_cc_noalloc_matches_cc_twice_directed(sg, device, oracle_mod):
    lib = sg.lib()
    for n, m, seed in ((20000, 50000, 3), (3000, 5000, 4)):
        g = _directed(oracle_mod, n, m, seed)
        G = sg.Graph.from_csr(device, g.offsets, g.cols, directed=True)
        ref, _ = _labels(sg, G, lib.sg_cc)
        got, rounds = _labels(sg, G, lib.sg_cc_noalloc)
        assert np.array_equal(got, ref), f"directed n={n}: labels differ"
        assert rounds >= 1
        G.destroy()


def test_cc_noalloc_asymmetric_csr_without_directed_flag(sg, device, oracle_mod):
    """An undirected-flagged CSR missing some reverse arcs: the no-alloc path
    takes the label propagation too, as sg_cc does."""
    rng = np.random.default_rng(77)
    n, m = 4000, 9000
    u, v = rng.integers(0, n, m), rng.integers(0, n, m)
    keep = u != v
    pairs = np.unique(np.stack([u[keep], v[keep]], 1), axis=0)
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(pairs[:, 0], minlength=n), out=off[1:])
    G = sg.Graph.from_csr(device, off, np.ascontiguousarray(pairs[:, 1], dtype=np.int32))
    ref, _ = _labels(sg, G, sg.lib().sg_cc)
    got, _ = _labels(sg, G, sg.lib().sg_cc_noalloc)
    assert np.array_equal(got, ref)
