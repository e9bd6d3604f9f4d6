"""Memory spaces of the C ABI on the MI355X: device, shared (managed) and
host (pinned) allocations are written by a kernel and read by the host, and
the reverse (the USM spaces of utils/memory.hpp:12-39 that the reference's
tests and examples use: tests/operators/advance.cpp:19-20 shares `visited`
between host loops and device functors)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _host_view(buf, n, dtype=np.uint32):
    ctype = {np.uint32: C.c_uint32, np.uint64: C.c_uint64}[dtype]
    return np.ctypeslib.as_array((ctype * n).from_address(buf.ptr))


@pytest.mark.parametrize("space_name", ["SPACE_SHARED", "SPACE_HOST"])
def test_host_accessible_spaces(sg, device, space_name):
    n = 100_003
    space = getattr(sg, space_name)
    buf = device.alloc(4 * n, space)
    buf.fill(0xDEADBEEF, 4, n)  # device kernel writes (sg_fill synchronizes)
    view = _host_view(buf, n)
    assert (view == 0xDEADBEEF).all()
    view[:] = np.arange(n, dtype=np.uint32)  # host writes, device reads
    dev_copy = device.alloc(4 * n, sg.SPACE_DEVICE)
    sg.lib().sg_memcpy(device.h, C.c_void_p(dev_copy.ptr), C.c_void_p(buf.ptr), 4 * n)
    assert np.array_equal(dev_copy.download(np.uint32), np.arange(n, dtype=np.uint32))


def test_device_space_roundtrip(sg, device):
    x = np.random.default_rng(1).integers(0, 2**63, size=77_777, dtype=np.uint64)
    b = device.to_device(x)
    assert np.array_equal(b.download(np.uint64), x)
    b.fill(7, 8, 10)
    y = b.download(np.uint64)
    assert (y[:10] == 7).all() and np.array_equal(y[10:], x[10:])


def test_copy_device_kernel(sg, device):
    x = np.random.default_rng(3).integers(0, 2**32, size=(1 << 22) + 4, dtype=np.uint32)
    a = device.to_device(x)
    b = device.alloc(x.nbytes)
    sg.check(sg.lib().sg_copy_device(device.h, C.c_void_p(b.ptr), C.c_void_p(a.ptr), x.nbytes))
    assert np.array_equal(b.download(np.uint32), x)
    # sizes and pointers must be 16-byte multiples
    assert sg.lib().sg_copy_device(device.h, C.c_void_p(b.ptr), C.c_void_p(a.ptr), 20) == sg.SG_ERR_INVALID


def _splitmix(i, seed):
    with np.errstate(over="ignore"):
        z = i.astype(np.uint64) + np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def test_gather_device_kernel(sg, device):
    """The random-gather calibration kernel reads src at splitmix64(i, seed)
    mod src_words (the published splitmix64 finaliser, restated in numpy)."""
    src = np.random.default_rng(4).integers(0, 2**32, size=1001, dtype=np.uint32)
    a = device.to_device(src)
    count = 5000
    b = device.alloc(4 * count)
    sg.check(sg.lib().sg_gather_device(device.h, C.c_void_p(b.ptr), C.c_void_p(a.ptr), src.size, count, 7))
    idx = _splitmix(np.arange(count), 7) % np.uint64(src.size)
    assert np.array_equal(b.download(np.uint32, count), src[idx.astype(np.int64)])
    assert sg.lib().sg_gather_device(device.h, C.c_void_p(b.ptr), C.c_void_p(a.ptr), 0, count, 7) == sg.SG_ERR_INVALID


def test_prefetch_shared_memory(sg, device):
    """sg_prefetch migrates managed memory to the device (a hint: memory it
    cannot migrate is accepted too) and leaves the contents intact."""
    n = 1 << 18
    buf = device.alloc(4 * n, sg.SPACE_SHARED)
    view = _host_view(buf, n)
    view[:] = np.arange(n, dtype=np.uint32)
    sg.check(sg.lib().sg_prefetch(device.h, C.c_void_p(buf.ptr), 4 * n))
    dev_copy = device.alloc(4 * n, sg.SPACE_DEVICE)
    sg.check(sg.lib().sg_memcpy(device.h, C.c_void_p(dev_copy.ptr), C.c_void_p(buf.ptr), 4 * n))
    assert np.array_equal(dev_copy.download(np.uint32), np.arange(n, dtype=np.uint32))
    pinned = device.alloc(4 * n, sg.SPACE_HOST)
    sg.check(sg.lib().sg_prefetch(device.h, C.c_void_p(pinned.ptr), 4 * n))
    assert sg.lib().sg_prefetch(device.h, None, 16) != 0  # null pointer rejected


def test_device_digest_matches_oracle(sg, device, oracle_mod):
    """sg_digest (the autotuner's output comparison) equals the oracle's
    order-sensitive digest for 4- and 8-byte arrays."""
    rng = np.random.default_rng(9)
    x32 = rng.integers(0, 2**31, size=123_457, dtype=np.int32)
    x64 = rng.integers(0, 2**62, size=54_321, dtype=np.int64)
    assert device.to_device(x32).digest(len(x32), 4) == oracle_mod.digest_i32(x32)
    assert device.to_device(x64).digest(len(x64), 8) == oracle_mod.digest_i64(x64)
    y = x32.copy()
    y[[0, 1]] = y[[1, 0]]  # order-sensitive
    assert device.to_device(y).digest(len(y), 4) != oracle_mod.digest_i32(x32)
