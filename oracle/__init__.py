"""CPU oracle for the SYgraph traversal core -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or the timed
CPU baseline.  The product (``sygraph_amd``, ``include/``) never uses it.

Thin numpy/ctypes front-end over ``liboracle.so`` (plain C, built by
``oracle/Makefile``).  Each wrapper names the reference file:line that the C
function restates; see ``oracle.c`` for the parity pins.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_mix64.restype = C.c_uint64
        L.or_mix64.argtypes = [C.c_uint64]
        L.or_rand64.restype = C.c_uint64
        L.or_rand64.argtypes = [C.c_uint64, C.c_uint64]
        L.or_rmat_pairs.restype = None
        L.or_rmat_pairs.argtypes = [C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int,
                                    _u32p, _u32p]
        L.or_rmat_relabel.restype = C.c_uint32
        L.or_rmat_relabel.argtypes = [C.c_uint32, C.c_int, C.c_uint64]
        L.or_grid_pairs.restype = C.c_int64
        L.or_grid_pairs.argtypes = [C.c_int, C.c_int, C.c_double, C.c_uint64, _u32p, _u32p]
        L.or_er_pairs.restype = None
        L.or_er_pairs.argtypes = [C.c_int64, C.c_int64, C.c_uint64, _u32p, _u32p]
        L.or_build_csr.restype = C.c_int64
        L.or_build_csr.argtypes = [C.c_int64, C.c_int64, _u32p, _u32p, _i64p, _i32p]
        L.or_edge_weight.restype = C.c_uint32
        L.or_edge_weight.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64]
        L.or_fill_weights.restype = None
        L.or_fill_weights.argtypes = [C.c_int64, _i64p, _i32p, C.c_uint64, _u32p]
        L.or_digest_i64.restype = C.c_uint64
        L.or_digest_i64.argtypes = [_i64p, C.c_int64]
        L.or_digest_i32.restype = C.c_uint64
        L.or_digest_i32.argtypes = [_i32p, C.c_int64]
        L.or_bfs.restype = C.c_int32
        L.or_bfs.argtypes = [C.c_int64, _i64p, _i32p, C.c_int32, _i32p]
        L.or_sssp.restype = None
        L.or_sssp.argtypes = [C.c_int64, _i64p, _i32p, _u32p, C.c_int32, _u32p]
        L.or_cc.restype = None
        L.or_cc.argtypes = [C.c_int64, _i64p, _i32p, _u32p]
        L.or_tc.restype = C.c_uint64
        L.or_tc.argtypes = [C.c_int64, _i64p, _i32p]
        L.or_bc.restype = None
        L.or_bc.argtypes = [C.c_int64, _i64p, _i32p, C.c_void_p, C.c_void_p, C.c_int32, _i32p, _f64p]
        L.or_pull_scan_count.restype = C.c_int64
        L.or_pull_scan_count.argtypes = [C.c_int64, _i64p, _i32p, _i32p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_pull_hint.restype = None
        L.or_pull_hint.argtypes = [C.c_int64, _i64p, _i32p, _i64p, _i32p]
        L.or_bfs_plain_omp.restype = C.c_int32
        L.or_bfs_plain_omp.argtypes = [C.c_int64, _i64p, _i32p, C.c_int32, C.c_int, C.c_float, C.c_float, C.c_int,
                                       _i32p, _i32p, C.c_int32]
        L.or_sssp_plain_omp.restype = C.c_int32
        L.or_sssp_plain_omp.argtypes = [C.c_int64, _i64p, _i32p, _u32p, C.c_int32, C.c_int, _u32p]
        L.or_omp_max_threads.restype = C.c_int
        L.or_omp_max_threads.argtypes = []
        _lib = L
    return _lib


# ---------------------------------------------------------------- graphs --
@dataclass
class CSRGraph:
    """Symmetric CSR graph: int64 row offsets, int32 ascending column ids."""

    n: int
    offsets: np.ndarray
    cols: np.ndarray
    weights: np.ndarray | None = None

    @property
    def m(self) -> int:
        return int(self.offsets[-1])

    def degrees(self) -> np.ndarray:
        return np.diff(self.offsets)

    def digest(self) -> tuple[int, int]:
        return digest_i64(self.offsets), digest_i32(self.cols)


def transpose(g: CSRGraph) -> CSRGraph:
    """In-arc CSR of a directed graph: row v lists every u with an arc u -> v,
    ascending (the inverse CSR of graph/impls/graph_csr.hpp:423-425)."""
    src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.offsets))
    order = np.lexsort((src, g.cols.astype(np.int64)))
    off = np.zeros(g.n + 1, dtype=np.int64)
    np.cumsum(np.bincount(g.cols, minlength=g.n), out=off[1:])
    w = g.weights[order] if g.weights is not None else None
    return CSRGraph(g.n, off, np.ascontiguousarray(src[order], dtype=np.int32), w)


def csr_from_pairs(n: int, u: np.ndarray, v: np.ndarray) -> CSRGraph:
    u = np.ascontiguousarray(u, dtype=np.uint32)
    v = np.ascontiguousarray(v, dtype=np.uint32)
    off = np.zeros(n + 1, dtype=np.int64)
    cols = np.zeros(max(2 * len(u), 1), dtype=np.int32)
    m = lib().or_build_csr(n, len(u), u, v, off, cols)
    return CSRGraph(n, off, cols[:m].copy())


def rmat_pairs(scale: int, edge_factor: int, a: float, b: float, c: float, seed: int, relabel: bool = True):
    ne = edge_factor << scale
    u = np.empty(ne, dtype=np.uint32)
    v = np.empty(ne, dtype=np.uint32)
    lib().or_rmat_pairs(scale, ne, a, b, c, seed, 1 if relabel else 0, u, v)
    return u, v


def rmat(scale: int, edge_factor: int, a: float = 0.57, b: float = 0.19, c: float = 0.19, seed: int = 1,
         relabel: bool = True) -> CSRGraph:
    """Graph500-style RMAT, symmetrised, self loops and duplicates removed."""
    u, v = rmat_pairs(scale, edge_factor, a, b, c, seed, relabel)
    return csr_from_pairs(1 << scale, u, v)


def grid(rows: int, cols: int, p_delete: float = 0.0, seed: int = 4) -> CSRGraph:
    """rows x cols 4-neighbour grid, each undirected edge deleted with prob p_delete."""
    cap = max(2 * rows * cols, 1)
    u = np.empty(cap, dtype=np.uint32)
    v = np.empty(cap, dtype=np.uint32)
    k = lib().or_grid_pairs(rows, cols, p_delete, seed, u, v)
    return csr_from_pairs(rows * cols, u[:k], v[:k])


def erdos_renyi(n: int, m: int, seed: int) -> CSRGraph:
    u = np.empty(m, dtype=np.uint32)
    v = np.empty(m, dtype=np.uint32)
    lib().or_er_pairs(n, m, seed, u, v)
    return csr_from_pairs(n, u, v)


def from_dense(adj) -> CSRGraph:
    """CSR of a dense adjacency matrix, row-major scan (io/read_csr.hpp:40-67)."""
    a = np.asarray(adj)
    n = a.shape[0]
    off = np.zeros(n + 1, dtype=np.int64)
    cols = []
    vals = []
    for i in range(n):
        nz = np.nonzero(a[i])[0]
        cols.extend(nz.tolist())
        vals.extend(a[i][nz].tolist())
        off[i + 1] = len(cols)
    g = CSRGraph(n, off, np.asarray(cols, dtype=np.int32))
    g.weights = np.asarray(vals, dtype=np.uint32)
    return g


def parse_matrix_text(text: str) -> np.ndarray:
    """Parse the reference's adjacency-matrix fixture text (first token = n)."""
    toks = text.split()
    n = int(toks[0])
    return np.asarray([int(t) for t in toks[1:1 + n * n]], dtype=np.int64).reshape(n, n)


def edge_weight(u: int, v: int, seed: int) -> int:
    return int(lib().or_edge_weight(u, v, seed))


def fill_weights(g: CSRGraph, seed: int) -> np.ndarray:
    w = np.empty(g.m, dtype=np.uint32)
    lib().or_fill_weights(g.n, g.offsets, g.cols, seed, w)
    g.weights = w
    return w


def digest_i64(x: np.ndarray) -> int:
    x = np.ascontiguousarray(x, dtype=np.int64)
    return int(lib().or_digest_i64(x, len(x)))


def digest_i32(x: np.ndarray) -> int:
    x = np.ascontiguousarray(x, dtype=np.int32)
    return int(lib().or_digest_i32(x, len(x)))


# ------------------------------------------------------------ algorithms --
def bfs(g: CSRGraph, src: int) -> tuple[np.ndarray, int]:
    """Depths (unreached = n+1) and #levels; restates examples/bfs/bfs.cpp:17-56."""
    d = np.empty(g.n, dtype=np.int32)
    lv = lib().or_bfs(g.n, g.offsets, g.cols, src, d)
    return d, int(lv)


def sssp(g: CSRGraph, src: int, weights: np.ndarray | None = None) -> np.ndarray:
    """Dijkstra distances (unreached = n+1); restates examples/sssp/sssp.cpp:24-66."""
    w = np.ascontiguousarray(weights if weights is not None else g.weights, dtype=np.uint32)
    d = np.empty(g.n, dtype=np.uint32)
    lib().or_sssp(g.n, g.offsets, g.cols, w, src, d)
    return d


def cc(g: CSRGraph) -> np.ndarray:
    """Max-id component labels; fixpoint of algorithms/cc.hpp:103-169."""
    lab = np.empty(g.n, dtype=np.uint32)
    lib().or_cc(g.n, g.offsets, g.cols, lab)
    return lab


def tc(g: CSRGraph) -> int:
    """Exact triangle count (uint64); the count algorithms/tc.hpp:67-138 intends."""
    return int(lib().or_tc(g.n, g.offsets, g.cols))


def bc(g: CSRGraph, sources, inverse: Optional[CSRGraph] = None) -> np.ndarray:
    """FP64 Brandes summed over sources, source excluded, no /2 (algorithms/bc.hpp:118-200).
    For a directed graph pass its inverse (in-arc CSR): sigma is pulled over in-arcs."""
    s = np.ascontiguousarray(np.asarray(sources, dtype=np.int32))
    out = np.empty(g.n, dtype=np.float64)
    io = ic = None
    if inverse is not None:
        io = np.ascontiguousarray(inverse.offsets, dtype=np.int64)
        ic = np.ascontiguousarray(inverse.cols, dtype=np.int32)
    lib().or_bc(g.n, g.offsets, g.cols, io.ctypes.data if io is not None else None,
                ic.ctypes.data if ic is not None else None, len(s), s, out)
    return out


def pull_hint(g: CSRGraph, inverse: Optional[CSRGraph] = None) -> np.ndarray:
    """Highest out-degree in-neighbour of every vertex (lowest id on ties, -1 if none)."""
    inv = inverse if inverse is not None else g
    out = np.empty(g.n, dtype=np.int32)
    lib().or_pull_hint(g.n, inv.offsets, inv.cols, g.offsets, out)
    return out


def pull_scan_count(g: CSRGraph, depth: np.ndarray, level: int, hint: Optional[np.ndarray] = None):
    """(probes, scanned vertices, hint misses) of one pull level; see or_pull_scan_count."""
    cnt, missed = C.c_int64(0), C.c_int64(0)
    h = np.ascontiguousarray(hint, dtype=np.int32) if hint is not None else None
    tot = lib().or_pull_scan_count(g.n, g.offsets, g.cols, np.ascontiguousarray(depth, dtype=np.int32), level,
                                   h.ctypes.data if h is not None else None, C.byref(cnt), C.byref(missed))
    return int(tot), int(cnt.value), int(missed.value)


BFS_PUSH, BFS_PULL, BFS_HYBRID = 0, 1, 2


def bfs_plain_omp(g: CSRGraph, src: int, mode: int = BFS_PUSH, alpha: float = 15.0, beta: float = 24.0,
                  nthreads: int = 0, with_dirs: bool = False):
    """OpenMP restatement of the reference plain method (CPU baseline).
    Returns (depth, levels) or, with_dirs, (depth, levels, directions 0=push 1=pull)."""
    d = np.empty(g.n, dtype=np.int32)
    cap = 1 << 16
    dirs = np.zeros(cap, dtype=np.int32)
    lv = lib().or_bfs_plain_omp(g.n, g.offsets, g.cols, src, mode, alpha, beta, nthreads, d, dirs, cap)
    if with_dirs:
        return d, int(lv), dirs[:min(int(lv), cap)].copy()
    return d, int(lv)


def sssp_plain_omp(g: CSRGraph, src: int, nthreads: int = 0) -> tuple[np.ndarray, int]:
    d = np.empty(g.n, dtype=np.uint32)
    it = lib().or_sssp_plain_omp(g.n, g.offsets, g.cols, np.ascontiguousarray(g.weights, dtype=np.uint32), src,
                                 nthreads, d)
    return d, int(it)


def omp_max_threads() -> int:
    return int(lib().or_omp_max_threads())


def teps_edges(g: CSRGraph, depth: np.ndarray) -> int:
    """Reference TEPS numerator: sum of degrees of reached vertices (algorithms/bfs.hpp:88-99)."""
    return int(g.degrees()[depth != g.n + 1].sum())
