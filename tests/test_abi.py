"""Host-side checks of the C ABI that need no GPU: the library loads, exports
exactly what include/sygraph_hip.h declares, and its argument checks fail with
status codes and messages instead of crashing (the reference throws
std::runtime_error on the same conditions)."""
import ctypes as C
import re
import subprocess

import pytest


def test_library_exports_every_declared_symbol(sg):
    declared = sg.declared_symbols()
    assert len(declared) >= 50
    out = subprocess.run(["nm", "-D", "--defined-only", sg.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(sg_[a-z0-9_]+)$", out, re.M))
    missing = [s for s in declared if s not in exported]
    assert not missing, f"declared but not exported: {missing}"
    for s in declared:  # and every one resolves through ctypes
        getattr(sg.lib(), s)


def test_abi_version(sg):
    assert sg.lib().sg_abi_version() == 1


def test_null_arguments_return_invalid(sg):
    L = sg.lib()
    assert L.sg_device_open(0, None) == sg.SG_ERR_INVALID
    assert "null" in sg.last_error()
    assert L.sg_device_info_get(None, None) == sg.SG_ERR_INVALID
    assert L.sg_graph_create(None, 0, 0, None, 8, None, None, 0, None) == sg.SG_ERR_INVALID
    assert L.sg_bfs(None, 0, None, None, None) == sg.SG_ERR_INVALID
    assert L.sg_frontier_create(None, 10, None) == sg.SG_ERR_INVALID
    assert L.sg_frontier_insert(None, 0) == sg.SG_ERR_INVALID
    # destroying NULL handles is a no-op, as releaseUSM on nullptr (utils/memory.hpp:41-46)
    assert L.sg_graph_destroy(None) == sg.SG_OK
    assert L.sg_frontier_destroy(None) == sg.SG_OK
    assert L.sg_device_close(None) == sg.SG_OK


def test_device_count_never_fails(sg):
    assert sg.device_count() >= 0


def test_open_without_device_reports_nodevice(sg):
    if sg.device_count() > 0:
        pytest.skip("a HIP device is visible")
    h = C.c_void_p()
    rc = sg.lib().sg_device_open(0, C.byref(h))
    assert rc == sg.SG_ERR_NODEVICE
    assert "no HIP device" in sg.last_error()
    with pytest.raises(sg.SygraphError):
        sg.Device(0)


def test_bfs_options_default(sg):
    o = sg.BfsOptions()
    sg.lib().sg_bfs_options_default(C.byref(o))
    assert o.policy == sg.POLICY_PUSH and o.load_balancer == sg.LB_BUCKETING
    assert o.alpha == 15.0 and o.beta == 24.0  # examples/bfs/bfs.cpp:105-106
