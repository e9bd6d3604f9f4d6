"""sygraph_amd -- Python front-end over libsygraph_hip (the C ABI declared in
include/sygraph_hip.h).

The classes mirror the reference's host interface for the traversal core:
``Graph`` ~ ``graph::build::fromCSR`` / ``GraphCSR`` (graph/build.hpp:31-41),
``Frontier`` ~ ``frontier::Frontier<T, mlb>`` (frontier/frontier.hpp:45-101),
``BFS``/``SSSP``/``CC``/``TC``/``BC`` ~ ``algorithms::*`` with ``init``/``run``
and getters (algorithms/bfs.hpp:120-351 ...).  Errors raise ``SygraphError``
(a ``RuntimeError``, as the reference throws ``std::runtime_error``).

Every call goes to the HIP library; there is no CPU fallback.  Importing works
without a GPU (so the ABI can be inspected); opening a device without one
raises.
"""
from __future__ import annotations

import ctypes as C
import sys
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsygraph_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "sygraph_hip.h")

SG_OK = 0
SG_ERR_INVALID = -1
SG_ERR_HIP = -2
SG_ERR_NOMEM = -3
SG_ERR_UNSUPPORTED = -4
SG_ERR_NODEVICE = -5
SG_ERR_STATE = -6

SPACE_HOST, SPACE_DEVICE, SPACE_SHARED = 0, 1, 2
GRAPH_DIRECTED, GRAPH_WEIGHTED, GRAPH_INPUT_DEVICE, GRAPH_WEIGHT_FLOAT = 1, 2, 4, 8
POLICY_PUSH, POLICY_PULL, POLICY_HYBRID, POLICY_LEARNED = 0, 1, 2, 3
LB_WORKITEM, LB_SUBGROUP, LB_WORKGROUP, LB_BUCKETING = 0, 1, 2, 3
MAX_LEVELS_RECORDED = 4096


class SygraphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("compute_units", C.c_int32),
                ("wave_size", C.c_int32), ("max_workgroup", C.c_int32), ("max_waves_per_wg", C.c_int32),
                ("lds_bytes", C.c_int64), ("global_mem_bytes", C.c_int64), ("l2_bytes", C.c_int64),
                ("clock_khz", C.c_int32), ("memory_clock_khz", C.c_int32), ("memory_bus_bits", C.c_int32),
                ("device_ordinal", C.c_int32)]


class GraphView(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64), ("offset_bytes", C.c_int32), ("flags", C.c_int32),
                ("offsets", C.c_void_p), ("cols", C.c_void_p), ("weights", C.c_void_p),
                ("inv_offsets", C.c_void_p), ("inv_cols", C.c_void_p), ("inv_weights", C.c_void_p)]


class FrontierView(C.Structure):
    _fields_ = [("n", C.c_int64), ("words0", C.c_int64), ("words1", C.c_int64), ("l0", C.c_void_p),
                ("l1", C.c_void_p), ("list", C.c_void_p), ("list_size", C.c_void_p)]


class BfsOptions(C.Structure):
    _fields_ = [("policy", C.c_int32), ("load_balancer", C.c_int32), ("alpha", C.c_float), ("beta", C.c_float),
                ("record_levels", C.c_int32), ("sync_per_level", C.c_int32), ("learned", C.c_void_p),
                ("t_medium", C.c_int32), ("t_large", C.c_int32), ("chunk", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("burst_arcs", C.c_int32), ("grid_burst_arcs", C.c_int32), ("grid_burst_workgroups", C.c_int32), ("learned_gate", C.c_int32)]


class MsbfsOptions(C.Structure):
    _fields_ = [("policy", C.c_int32), ("alpha", C.c_float), ("beta", C.c_float), ("learned", C.c_void_p)]


class LevelStat(C.Structure):
    _fields_ = [("direction", C.c_int32), ("load_balancer", C.c_int32), ("frontier_vertices", C.c_int64),
                ("frontier_edges", C.c_int64), ("discovered", C.c_int64), ("edges_inspected", C.c_int64),
                ("vertices_scanned", C.c_int64), ("hint_misses", C.c_int64), ("unexplored_edges", C.c_int64),
                ("max_degree", C.c_int64), ("action", C.c_int32),
                ("kernel_ms", C.c_float), ("convert_ms", C.c_float), ("burst", C.c_int32)]


class SsspStats(C.Structure):
    _fields_ = [("rounds", C.c_int32), ("reached_vertices", C.c_int64), ("reached_edges", C.c_int64),
                ("relaxations", C.c_int64), ("kernel_ms", C.c_float),
                ("round_vertices", C.c_int64 * MAX_LEVELS_RECORDED), ("round_edges", C.c_int64 * MAX_LEVELS_RECORDED),
                ("relax_ms", C.c_float), ("relax_launches", C.c_int32)]


class SsspOptions(C.Structure):
    _fields_ = [("load_balancer", C.c_int32), ("delta", C.c_double), ("heavy_pull", C.c_int32),
                ("burst_arcs", C.c_int32), ("record_relax", C.c_int32)]


class BfsStats(C.Structure):
    _fields_ = [("levels", C.c_int32), ("n_push", C.c_int32), ("n_pull", C.c_int32),
                ("reached_vertices", C.c_int64), ("reached_edges", C.c_int64), ("kernel_ms", C.c_float),
                ("level", LevelStat * MAX_LEVELS_RECORDED)]


class BfsManyStats(C.Structure):
    _fields_ = [("reached_vertices", C.c_int64), ("reached_edges", C.c_int64), ("kernel_ms", C.c_float)]


_lib = None

# name -> (restype, argtypes)
_P = C.c_void_p
_I = C.c_int
_SIGS = {
    "sg_last_error": (C.c_char_p, []),
    "sg_abi_version": (_I, []),
    "sg_device_count": (_I, [C.POINTER(C.c_int)]),
    "sg_device_open": (_I, [_I, C.POINTER(_P)]),
    "sg_device_close": (_I, [_P]),
    "sg_device_info_get": (_I, [_P, C.POINTER(DeviceInfo)]),
    "sg_device_stream": (_I, [_P, C.POINTER(_P)]),
    "sg_device_synchronize": (_I, [_P]),
    "sg_device_set_wait_limit": (_I, [_P, C.c_double]),
    "sg_alloc": (_I, [_P, C.c_size_t, _I, C.POINTER(_P)]),
    "sg_free": (_I, [_P, _P]),
    "sg_memcpy": (_I, [_P, _P, _P, C.c_size_t]),
    "sg_fill": (_I, [_P, _P, C.c_uint64, _I, C.c_size_t]),
    "sg_copy_device": (_I, [_P, _P, _P, C.c_size_t]),
    "sg_gather_device": (_I, [_P, _P, _P, C.c_uint64, C.c_size_t, C.c_uint64]),
    "sg_prefetch": (_I, [_P, _P, C.c_size_t]),
    "sg_sssp_options_default": (None, [_P]),
    "sg_sssp_ex": (_I, [_P, C.c_int32, _P, _P, _P]),
    "sg_policy_create": (_I, [_P, _P]),
    "sg_policy_destroy": (_I, [_P]),
    "sg_policy_set_params": (_I, [_P, _P]),
    "sg_policy_get_params": (_I, [_P, _P]),
    "sg_policy_forward": (_I, [_P, _P, C.c_int32, _P]),
    "sg_policy_gradient": (_I, [_P, _P, _P, C.c_int32, _P, _P]),
    "sg_policy_train_step": (_I, [_P, _P, _P, C.c_int32, C.c_float, _P]),
    "sg_event_create": (_I, [_P, C.POINTER(_P)]),
    "sg_event_destroy": (_I, [_P]),
    "sg_event_record": (_I, [_P]),
    "sg_event_synchronize": (_I, [_P]),
    "sg_event_elapsed_ms": (_I, [_P, _P, C.POINTER(C.c_float)]),
    "sg_graph_create": (_I, [_P, C.c_int64, C.c_int64, _P, _I, _P, _P, _I, C.POINTER(_P)]),
    "sg_graph_destroy": (_I, [_P]),
    "sg_graph_get_view": (_I, [_P, C.POINTER(GraphView)]),
    "sg_graph_generate_rmat": (_I, [_P, _I, _I, C.c_double, C.c_double, C.c_double, C.c_uint64, _I, C.c_uint64,
                                    C.POINTER(_P)]),
    "sg_graph_generate_grid": (_I, [_P, _I, _I, C.c_double, C.c_uint64, C.c_uint64, C.POINTER(_P)]),
    "sg_graph_generate_er": (_I, [_P, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64, C.POINTER(_P)]),
    "sg_graph_download": (_I, [_P, _P, _P, _P]),
    "sg_graph_digest": (_I, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "sg_digest": (_I, [_P, _P, C.c_int64, C.c_int, C.POINTER(C.c_uint64)]),
    "sg_graph_reached_edges": (_I, [_P, _P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "sg_frontier_create": (_I, [_P, C.c_int64, C.POINTER(_P)]),
    "sg_frontier_destroy": (_I, [_P]),
    "sg_frontier_get_view": (_I, [_P, C.POINTER(FrontierView)]),
    "sg_frontier_check": (_I, [_P, C.c_int64, C.POINTER(C.c_int)]),
    "sg_frontier_insert": (_I, [_P, C.c_int64]),
    "sg_frontier_remove": (_I, [_P, C.c_int64]),
    "sg_frontier_clear": (_I, [_P]),
    "sg_frontier_clear_async": (_I, [_P]),
    "sg_frontier_empty": (_I, [_P, C.POINTER(C.c_int)]),
    "sg_frontier_size": (_I, [_P, C.POINTER(C.c_int64)]),
    "sg_frontier_merge": (_I, [_P, _P]),
    "sg_frontier_intersect": (_I, [_P, _P]),
    "sg_frontier_and": (_I, [_P, _P, _P]),
    "sg_frontier_compact": (_I, [_P, _I, C.POINTER(C.c_int64)]),
    "sg_frontier_compact_async": (_I, [_P, _I]),
    "sg_frontier_save": (_I, [_P, _P]),
    "sg_frontier_load": (_I, [_P, _P]),
    "sg_frontier_swap": (_I, [_P, _P]),
    "sg_bfs_options_default": (None, [C.POINTER(BfsOptions)]),
    "sg_bfs": (_I, [_P, C.c_int32, C.POINTER(BfsOptions), _P, C.POINTER(BfsStats)]),
    "sg_bfs_many": (_I, [_P, C.c_int32, _P, C.POINTER(BfsOptions), _P, C.POINTER(BfsManyStats)]),
    "sg_sssp": (_I, [_P, C.c_int32, C.c_int32, _P, C.POINTER(SsspStats)]),
    "sg_graph_prepare": (_I, [_P, C.POINTER(C.c_float)]),
    "sg_cc": (_I, [_P, _P, C.POINTER(C.c_int32)]),
    "sg_cc_noalloc": (_I, [_P, _P, C.POINTER(C.c_int32)]),
    "sg_tc": (_I, [_P, C.POINTER(C.c_uint64), _P]),
    "sg_bc": (_I, [_P, C.c_int32, _P, _P]),
    "sg_bc_batched": (_I, [_P, C.c_int32, _P, _P]),
    "sg_bfs_multi": (_I, [_P, C.c_int32, _P, _P]),
    "sg_msbfs_options_default": (None, [C.POINTER(MsbfsOptions)]),
    "sg_bfs_multi_ex": (_I, [_P, C.c_int32, _P, C.POINTER(MsbfsOptions), _P]),
}


def lib():
    """Load libsygraph_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m sygraph_amd.build` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    """Entry points declared in include/sygraph_hip.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(sg_[a-z0-9_]+)\s*\(", txt, re.M)))


def last_error() -> str:
    return lib().sg_last_error().decode()


def check(code: int) -> None:
    if code != SG_OK:
        raise SygraphError(code, last_error())


def device_count() -> int:
    c = C.c_int(0)
    check(lib().sg_device_count(C.byref(c)))
    return c.value



def _finalizing() -> bool:
    """True while the interpreter shuts down: handles are then left to process
    exit, since module globals (and the library handle) may already be gone."""
    return sys is None or sys.is_finalizing()

# ------------------------------------------------------------------ device --
class Device:
    """A HIP device plus its stream (replaces ``sycl::queue``)."""

    def __init__(self, ordinal: int = 0):
        h = C.c_void_p()
        check(lib().sg_device_open(ordinal, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().sg_device_close(self.h)
            self.h = None

    def __del__(self):
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> DeviceInfo:
        i = DeviceInfo()
        check(lib().sg_device_info_get(self.h, C.byref(i)))
        return i

    def synchronize(self):
        check(lib().sg_device_synchronize(self.h))

    # ---- memory
    def alloc(self, nbytes: int, space: int = SPACE_DEVICE) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes, space)

    def to_device(self, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, arr.nbytes, SPACE_DEVICE, dtype=arr.dtype, shape=arr.shape)
        b.upload(arr)
        return b

    def event(self) -> "Event":
        return Event(self)


class DeviceBuffer:
    def __init__(self, dev: Device, nbytes: int, space: int = SPACE_DEVICE, dtype=np.uint8, shape=None):
        self.dev = dev
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = shape if shape is not None else (self.nbytes // self.dtype.itemsize,)
        p = C.c_void_p()
        check(lib().sg_alloc(dev.h, max(self.nbytes, 1), space, C.byref(p)))
        self.ptr = p.value

    def free(self):
        if self.ptr:
            lib().sg_free(self.dev.h, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        if _finalizing():
            return
        try:
            self.free()
        except Exception:
            pass

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().sg_memcpy(self.dev.h, C.c_void_p(self.ptr), arr.ctypes.data_as(C.c_void_p), arr.nbytes))

    def download(self, dtype=None, count: Optional[int] = None) -> np.ndarray:
        dt = np.dtype(dtype) if dtype is not None else self.dtype
        cnt = count if count is not None else self.nbytes // dt.itemsize
        out = np.empty(cnt, dtype=dt)
        check(lib().sg_memcpy(self.dev.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes))
        return out

    def fill(self, value: int, pattern_bytes: int, count: int):
        check(lib().sg_fill(self.dev.h, C.c_void_p(self.ptr), value, pattern_bytes, count))

    def digest(self, count: int, elem_bytes: int = 4, first: int = 0) -> int:
        """Order-sensitive 64-bit digest of `count` elements from element
        `first`, on the device."""
        assert (first + count) * elem_bytes <= self.nbytes
        h = C.c_uint64()
        check(lib().sg_digest(self.dev.h, C.c_void_p(self.ptr + first * elem_bytes), count, elem_bytes, C.byref(h)))
        return h.value


class Event:
    def __init__(self, dev: Device):
        h = C.c_void_p()
        check(lib().sg_event_create(dev.h, C.byref(h)))
        self.h = h

    def record(self):
        check(lib().sg_event_record(self.h))

    def elapsed_ms(self, stop: "Event") -> float:
        ms = C.c_float()
        check(lib().sg_event_elapsed_ms(self.h, stop.h, C.byref(ms)))
        return ms.value

    def __del__(self):
        if _finalizing():
            return
        try:
            lib().sg_event_destroy(self.h)
        except Exception:
            pass


# ------------------------------------------------------------------- graph --
class Graph:
    """Device CSR graph (GraphCSR, graph/impls/graph_csr.hpp:158-439)."""

    def __init__(self, dev: Device, handle: C.c_void_p):
        self.dev = dev
        self.h = handle
        v = self.view()
        self.n = int(v.n)
        self.m = int(v.m)

    @classmethod
    def from_csr(cls, dev: Device, offsets: np.ndarray, cols: np.ndarray, weights: Optional[np.ndarray] = None,
                 directed: bool = False, weight_float: bool = False, offset_bytes: int = 8) -> "Graph":
        off = np.ascontiguousarray(offsets, dtype=np.int64 if offset_bytes == 8 else np.uint32)
        col = np.ascontiguousarray(cols, dtype=np.int32)
        n = len(off) - 1
        m = len(col)
        flags = GRAPH_DIRECTED if directed else 0
        wptr = None
        if weights is not None:
            w = np.ascontiguousarray(weights, dtype=np.float32 if weight_float else np.uint32)
            flags |= GRAPH_WEIGHTED | (GRAPH_WEIGHT_FLOAT if weight_float else 0)
            wptr = w.ctypes.data_as(C.c_void_p)
        h = C.c_void_p()
        check(lib().sg_graph_create(dev.h, n, m, off.ctypes.data_as(C.c_void_p), offset_bytes,
                                    col.ctypes.data_as(C.c_void_p), wptr, flags, C.byref(h)))
        return cls(dev, h)

    @classmethod
    def rmat(cls, dev: Device, scale: int, edge_factor: int, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1, relabel: bool = True, weight_seed: int = 0) -> "Graph":
        h = C.c_void_p()
        check(lib().sg_graph_generate_rmat(dev.h, scale, edge_factor, a, b, c, seed, 1 if relabel else 0,
                                           weight_seed, C.byref(h)))
        return cls(dev, h)

    @classmethod
    def grid(cls, dev: Device, rows: int, cols: int, p_delete: float = 0.0, seed: int = 4,
             weight_seed: int = 0) -> "Graph":
        h = C.c_void_p()
        check(lib().sg_graph_generate_grid(dev.h, rows, cols, p_delete, seed, weight_seed, C.byref(h)))
        return cls(dev, h)

    @classmethod
    def erdos_renyi(cls, dev: Device, n: int, m_pairs: int, seed: int, weight_seed: int = 0) -> "Graph":
        h = C.c_void_p()
        check(lib().sg_graph_generate_er(dev.h, n, m_pairs, seed, weight_seed, C.byref(h)))
        return cls(dev, h)

    def view(self) -> GraphView:
        v = GraphView()
        check(lib().sg_graph_get_view(self.h, C.byref(v)))
        return v

    def download(self, weights: bool = False):
        off = np.empty(self.n + 1, dtype=np.int64)
        col = np.empty(self.m, dtype=np.int32)
        w = np.empty(self.m, dtype=np.uint32) if weights else None
        check(lib().sg_graph_download(self.h, off.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p),
                                      w.ctypes.data_as(C.c_void_p) if w is not None else None))
        return off, col, w

    def prepare(self) -> float:
        """Build the BFS helpers (in-degree-0 bitmap, pull hint); returns their device time in ms."""
        ms = C.c_float()
        check(lib().sg_graph_prepare(self.h, C.byref(ms)))
        return float(ms.value)

    def digest(self) -> tuple[int, int]:
        a, b = C.c_uint64(), C.c_uint64()
        check(lib().sg_graph_digest(self.h, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def reached_edges(self, depth: DeviceBuffer) -> tuple[int, int]:
        e, v = C.c_int64(), C.c_int64()
        check(lib().sg_graph_reached_edges(self.h, C.c_void_p(depth.ptr), C.byref(e), C.byref(v)))
        return int(e.value), int(v.value)

    def destroy(self):
        if getattr(self, "h", None):
            lib().sg_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        if _finalizing():
            return
        try:
            self.destroy()
        except Exception:
            pass


# ---------------------------------------------------------------- frontier --
class Frontier:
    """MLB frontier (frontier/impls/mlb_frontier.hpp:155-449)."""

    def __init__(self, dev: Device, n: int):
        self.dev = dev
        h = C.c_void_p()
        check(lib().sg_frontier_create(dev.h, n, C.byref(h)))
        self.h = h
        self.n = n

    def view(self) -> FrontierView:
        v = FrontierView()
        check(lib().sg_frontier_get_view(self.h, C.byref(v)))
        return v

    def insert(self, i: int):
        check(lib().sg_frontier_insert(self.h, i))

    def remove(self, i: int):
        check(lib().sg_frontier_remove(self.h, i))

    def check(self, i: int) -> bool:
        p = C.c_int()
        check(lib().sg_frontier_check(self.h, i, C.byref(p)))
        return bool(p.value)

    def clear(self):
        check(lib().sg_frontier_clear(self.h))

    def empty(self) -> bool:
        e = C.c_int()
        check(lib().sg_frontier_empty(self.h, C.byref(e)))
        return bool(e.value)

    def size(self) -> int:
        s = C.c_int64()
        check(lib().sg_frontier_size(self.h, C.byref(s)))
        return int(s.value)

    def merge(self, other: "Frontier"):
        check(lib().sg_frontier_merge(self.h, other.h))

    def intersect(self, other: "Frontier"):
        check(lib().sg_frontier_intersect(self.h, other.h))

    def compact(self, mode: int = 0) -> np.ndarray:
        cnt = C.c_int64()
        check(lib().sg_frontier_compact(self.h, mode, C.byref(cnt)))
        v = self.view()
        out = np.empty(cnt.value, dtype=np.int32)
        if cnt.value:
            check(lib().sg_memcpy(self.dev.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(v.list), out.nbytes))
        return out

    def words(self) -> tuple[np.ndarray, np.ndarray]:
        v = self.view()
        buf = np.empty(v.words0 + v.words1, dtype=np.uint64)
        check(lib().sg_frontier_save(self.h, buf.ctypes.data_as(C.c_void_p)))
        return buf[:v.words0].copy(), buf[v.words0:].copy()

    def save(self) -> np.ndarray:
        v = self.view()
        buf = np.empty(v.words0 + v.words1, dtype=np.uint64)
        check(lib().sg_frontier_save(self.h, buf.ctypes.data_as(C.c_void_p)))
        return buf

    def load(self, state: np.ndarray):
        st = np.ascontiguousarray(state, dtype=np.uint64)
        check(lib().sg_frontier_load(self.h, st.ctypes.data_as(C.c_void_p)))

    def active(self) -> list[int]:
        l0, _ = self.words()
        bits = np.unpackbits(l0.view(np.uint8), bitorder="little")[: self.n]
        return np.nonzero(bits)[0].tolist()

    def destroy(self):
        if getattr(self, "h", None):
            lib().sg_frontier_destroy(self.h)
            self.h = None

    def __del__(self):
        if _finalizing():
            return
        try:
            self.destroy()
        except Exception:
            pass


def swap(a: Frontier, b: Frontier):
    check(lib().sg_frontier_swap(a.h, b.h))


# -------------------------------------------------------------- algorithms --
def bfs_options(policy: int = POLICY_PUSH, load_balancer: int = LB_BUCKETING, alpha: float = 15.0,
                beta: float = 24.0, record_levels: bool = False, sync_per_level: bool = False,
                learned: Optional["Policy"] = None, tuning: Optional[dict] = None) -> BfsOptions:
    o = BfsOptions()
    lib().sg_bfs_options_default(C.byref(o))
    o.policy = policy
    o.load_balancer = load_balancer
    o.alpha = alpha
    o.beta = beta
    o.record_levels = 1 if record_levels else 0
    o.sync_per_level = 1 if sync_per_level else 0
    o.learned = learned.h if learned is not None else None
    for k, v in (tuning or {}).items():
        setattr(o, k, int(v))
    return o


class BFS:
    """algorithms::BFS (algorithms/bfs.hpp:120-351): init(source), run(direction, alpha, beta),
    getDistances().  Unreached vertices keep depth n+1."""

    def __init__(self, g: Graph):
        self.g = g
        self.depth: Optional[DeviceBuffer] = None
        self.source: Optional[int] = None
        self.stats = BfsStats()

    def init(self, source: int):
        self.source = int(source)
        if self.depth is None:
            self.depth = DeviceBuffer(self.g.dev, 4 * max(self.g.n, 1), dtype=np.int32, shape=(self.g.n,))

    def reset(self):
        self.source = None

    def run(self, policy: int = POLICY_PUSH, alpha: float = 1.0, beta: float = 1.0,
            load_balancer: int = LB_BUCKETING, sync_per_level: bool = False,
            learned: Optional["Policy"] = None, tuning: Optional[dict] = None,
            record_levels: bool = True) -> BfsStats:
        """tuning: optional knobs t_medium, t_large, chunk, blocks_per_cu, burst_arcs, grid_burst_arcs,
        grid_burst_workgroups, learned_gate.  record_levels: time each step's traversal kernel and its
        helpers on their own (stats.level[i].kernel_ms / convert_ms; one more event per step)."""
        if self.source is None:
            raise SygraphError(SG_ERR_STATE, "BFS instance not initialized")
        o = bfs_options(policy, load_balancer, alpha, beta, record_levels=record_levels,
                        sync_per_level=sync_per_level, learned=learned, tuning=tuning)
        check(lib().sg_bfs(self.g.h, self.source, C.byref(o), C.c_void_p(self.depth.ptr), C.byref(self.stats)))
        return self.stats

    def get_distances(self) -> np.ndarray:
        return self.depth.download(np.int32, self.g.n)

    getDistances = get_distances

    def directions(self) -> list[str]:
        return ["push" if self.stats.level[i].direction == 0 else "pull" for i in range(self.stats.levels)]


class SSSP:
    """algorithms::SSSP (algorithms/sssp.hpp:103-222): init(source), run(), getDistance(v).
    Distances have the weight type (uint32 or float32); unreached = n+1."""

    def __init__(self, g: Graph):
        self.g = g
        v = g.view()
        if not v.weights:
            raise SygraphError(SG_ERR_STATE, "graph has no edge weights")
        self.dtype = np.float32 if (v.flags & GRAPH_WEIGHT_FLOAT) else np.uint32
        self.dist: Optional[DeviceBuffer] = None
        self.source: Optional[int] = None
        self.stats = SsspStats()

    def init(self, source: int):
        self.source = int(source)
        if self.dist is None:
            self.dist = DeviceBuffer(self.g.dev, 4 * max(self.g.n, 1), dtype=self.dtype, shape=(self.g.n,))

    def reset(self):
        self.source = None

    def run(self, load_balancer: int = LB_BUCKETING, delta: float = 0.0, heavy_pull: int = 0,
            burst_arcs: int = 0, record_relax: bool = False) -> SsspStats:
        """delta: bucket width of the near/far split (0 automatic: half the
        mean arc weight on a graph of at least 2^25 arcs and 8 arcs per
        vertex, none otherwise; negative or inf for the
        reference's frontier Bellman-Ford).  heavy_pull: direction of a
        bucket's heavy round (0 automatic, 1 pull, -1 push).  burst_arcs:
        small rounds of the reference's schedule run inside one workgroup (0
        default 4096, negative never).  record_relax: time each relax launch on its own
        (stats.relax_ms, stats.relax_launches)."""
        if self.source is None:
            raise SygraphError(SG_ERR_STATE, "SSSP instance not initialized")
        o = SsspOptions()
        lib().sg_sssp_options_default(C.byref(o))
        o.load_balancer = load_balancer
        o.delta = float(delta)
        o.heavy_pull = int(heavy_pull)
        o.burst_arcs = int(burst_arcs)
        o.record_relax = 1 if record_relax else 0
        check(lib().sg_sssp_ex(self.g.h, self.source, C.byref(o), C.c_void_p(self.dist.ptr), C.byref(self.stats)))
        return self.stats

    def get_distances(self) -> np.ndarray:
        return self.dist.download(self.dtype, self.g.n)

    def getDistance(self, v: int):
        return self.get_distances()[v]


class CC:
    """algorithms::CC (algorithms/cc.hpp:74-192): run() leaves every vertex with the largest
    vertex id of its component (undirected) / the largest id that reaches it (directed)."""

    def __init__(self, g: Graph):
        self.g = g
        self.labels: Optional[DeviceBuffer] = None
        self.rounds = 0

    def init(self, source: int = 0):
        if self.labels is None:
            self.labels = DeviceBuffer(self.g.dev, 4 * max(self.g.n, 1), dtype=np.uint32, shape=(self.g.n,))

    def run(self) -> int:
        if self.labels is None:
            raise SygraphError(SG_ERR_STATE, "CC instance not initialized")
        r = C.c_int32()
        check(lib().sg_cc(self.g.h, C.c_void_p(self.labels.ptr), C.byref(r)))
        self.rounds = r.value
        return self.rounds

    def get_labels(self) -> np.ndarray:
        return self.labels.download(np.uint32, self.g.n)


class TC:
    """algorithms::TC (algorithms/tc.hpp:57-143): exact uint64 triangle count."""

    def __init__(self, g: Graph):
        self.g = g
        self.count: Optional[int] = None
        self.per_vertex: Optional[DeviceBuffer] = None

    def init(self):
        self.count = None

    def run(self, per_vertex: bool = False) -> int:
        c = C.c_uint64()
        pv = None
        if per_vertex:
            self.per_vertex = DeviceBuffer(self.g.dev, 8 * max(self.g.n, 1), dtype=np.uint64, shape=(self.g.n,))
            pv = C.c_void_p(self.per_vertex.ptr)
        check(lib().sg_tc(self.g.h, C.byref(c), pv))
        self.count = int(c.value)
        return self.count

    def getNumTriangles(self) -> int:
        if self.count is None:
            raise SygraphError(SG_ERR_STATE, "TC instance not initialized")
        return self.count


class BC:
    """algorithms::BC (algorithms/bc.hpp:24-224): Brandes dependencies summed
    over the given sources into a float vector (source excluded, no halving)."""

    def __init__(self, g: Graph):
        self.g = g
        self.values: Optional[DeviceBuffer] = None
        self.sources: list[int] = []

    def init(self, sources):
        self.sources = [int(s) for s in (sources if hasattr(sources, "__len__") else [sources])]
        self.values = DeviceBuffer(self.g.dev, 4 * max(self.g.n, 1), dtype=np.float32, shape=(self.g.n,))
        self.values.fill(0, 4, max(self.g.n, 1))

    def run(self, batched: bool = False):
        """batched: 64 sources per bit-lane batch (sg_bc_batched, FP32 path
        counts with per-vertex exponents); otherwise one source after another
        (sg_bc, FP64).  Equal up to FP32 rounding."""
        if self.values is None:
            raise SygraphError(SG_ERR_STATE, "BC instance not initialized")
        s = np.ascontiguousarray(np.asarray(self.sources, dtype=np.int32))
        fn = lib().sg_bc_batched if batched else lib().sg_bc
        check(fn(self.g.h, len(s), s.ctypes.data_as(C.c_void_p), C.c_void_p(self.values.ptr)))

    def get_values(self) -> np.ndarray:
        return self.values.download(np.float32, self.g.n)


MSBFS_AUTO = -1  # pull when a level holds more than n / 20 vertices


def msbfs_options(policy: int = MSBFS_AUTO, alpha: float = 15.0, beta: float = 24.0,
                  learned: Optional["Policy"] = None) -> MsbfsOptions:
    """Direction rule of a multi-source batch: MSBFS_AUTO (default), POLICY_PUSH / POLICY_PULL
    (always), POLICY_HYBRID (alpha / beta on the batch's totals) or POLICY_LEARNED."""
    o = MsbfsOptions()
    lib().sg_msbfs_options_default(C.byref(o))
    o.policy = policy
    o.alpha = alpha
    o.beta = beta
    o.learned = learned.h if learned is not None else None
    return o


def bfs_many(g: Graph, sources, policy: int = POLICY_PUSH, alpha: float = 15.0, beta: float = 24.0,
             out: Optional["DeviceBuffer"] = None, learned: Optional["Policy"] = None,
             tuning: Optional[dict] = None) -> BfsManyStats:
    """Single-source BFS from each source in turn, back to back on the device
    (sg_bfs_many): row i of out (int32[len(sources)][n], allocated when None
    and kept in the returned stats' .depth) holds the depths from sources[i]."""
    s = np.ascontiguousarray(np.asarray(sources, dtype=np.int32))
    buf = out if out is not None else DeviceBuffer(g.dev, 4 * max(len(s) * g.n, 1), dtype=np.int32)
    o = bfs_options(policy, LB_BUCKETING, alpha, beta, learned=learned, tuning=tuning)
    st = BfsManyStats()
    check(lib().sg_bfs_many(g.h, len(s), s.ctypes.data_as(C.c_void_p), C.byref(o), C.c_void_p(buf.ptr),
                            C.byref(st)))
    st.depth = buf
    return st


def bfs_multi(g: Graph, sources, on_device: bool = False, policy: int = MSBFS_AUTO,
              learned: Optional["Policy"] = None):
    """Depths from several sources, row i = sources[i] (unreached = n + 1).
    on_device: return the device buffer (int32[len(sources)][n]) instead.
    policy / learned: the per-level direction rule (msbfs_options)."""
    s = np.ascontiguousarray(np.asarray(sources, dtype=np.int32))
    out = DeviceBuffer(g.dev, 4 * max(len(s) * g.n, 1), dtype=np.int32)
    o = msbfs_options(policy, learned=learned)
    check(lib().sg_bfs_multi_ex(g.h, len(s), s.ctypes.data_as(C.c_void_p), C.byref(o), C.c_void_p(out.ptr)))
    if on_device:
        return out
    return out.download(np.int32, len(s) * g.n).reshape(len(s), g.n)


# ------------------------------------------------------------ learned policy --
POLICY_PARAMS, POLICY_FEATURES, POLICY_ACTIONS = 1542, 8, 6


class Policy:
    """The learned per-level BFS policy (8-32-32-6 MLP) on a device; training
    and inference run in HIP kernels."""

    def __init__(self, dev: Device, params: Optional[np.ndarray] = None):
        self.dev = dev
        h = C.c_void_p()
        check(lib().sg_policy_create(dev.h, C.byref(h)))
        self.h = h
        self._grad = DeviceBuffer(dev, 4 * POLICY_PARAMS, dtype=np.float32)
        if params is not None:
            self.set_params(params)

    def set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == POLICY_PARAMS
        check(lib().sg_policy_set_params(self.h, p.ctypes.data_as(C.c_void_p)))

    def get_params(self) -> np.ndarray:
        p = np.empty(POLICY_PARAMS, dtype=np.float32)
        check(lib().sg_policy_get_params(self.h, p.ctypes.data_as(C.c_void_p)))
        return p

    def forward(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        xb = self.dev.to_device(X)
        out = DeviceBuffer(self.dev, 4 * len(X) * POLICY_ACTIONS, dtype=np.float32)
        check(lib().sg_policy_forward(self.h, C.c_void_p(xb.ptr), len(X), C.c_void_p(out.ptr)))
        return out.download(np.float32).reshape(len(X), POLICY_ACTIONS)

    def gradient(self, X: np.ndarray, y: np.ndarray):
        X = np.ascontiguousarray(X, dtype=np.float32)
        xb, yb = self.dev.to_device(X), self.dev.to_device(np.ascontiguousarray(y, dtype=np.int32))
        loss = C.c_float()
        check(lib().sg_policy_gradient(self.h, C.c_void_p(xb.ptr), C.c_void_p(yb.ptr), len(X),
                                       C.c_void_p(self._grad.ptr), C.byref(loss)))
        return float(loss.value), self._grad.download(np.float32)

    def train_step(self, xb: DeviceBuffer, yb: DeviceBuffer, batch: int, lr: float, first: int = 0) -> float:
        """One Adam step on samples [first, first + batch) of the device arrays."""
        loss = C.c_float()
        check(lib().sg_policy_train_step(self.h, C.c_void_p(xb.ptr + 4 * POLICY_FEATURES * first),
                                         C.c_void_p(yb.ptr + 4 * first), batch, lr, C.byref(loss)))
        return float(loss.value)

    def destroy(self):
        if getattr(self, "h", None):
            lib().sg_policy_destroy(self.h)
            self.h = None

    def __del__(self):
        if _finalizing():
            return
        try:
            self.destroy()
        except Exception:
            pass
