"""The allocation-free exclusive scan (gb_exclusive_scan_i32_noalloc) through
the C ABI: bit-exact with the allocating gb_exclusive_scan_i32 on the same
input, on consecutive calls where the later call is shorter (the shared
device array keeps the earlier call's partials), in place, at the largest
length it accepts, and a clear GB_EINVAL beyond it."""
import numpy as np
import pytest

import gpu_graphs_amd as gg
from gpu_graphs_amd import _ffi

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def wrap_exclusive_scan(v):
    """Exclusive scan in 32-bit wrap-around arithmetic (numpy)."""
    out = np.zeros(v.size, dtype=np.uint32)
    if v.size > 1:
        out[1:] = np.cumsum(v[:-1].view(np.uint32), dtype=np.uint32)
    return out.view(np.int32)


def scan_both(dev, values):
    lib = gg.lib()
    d_in = gg.DeviceArray.from_numpy(dev, values)
    d_ref = gg.DeviceArray(dev, np.int32, values.size)
    d_new = gg.DeviceArray(dev, np.int32, values.size)
    _ffi.check(lib.gb_exclusive_scan_i32(dev, d_in.ptr, d_ref.ptr, values.size))
    _ffi.check(lib.gb_exclusive_scan_i32_noalloc(dev, d_in.ptr, d_new.ptr, values.size))
    return d_ref.numpy(), d_new.numpy()


def test_noalloc_scan_equals_the_allocating_scan_on_consecutive_calls(gpu):
    rng = np.random.default_rng(20260002)
    # The second call is shorter: its tiles reuse the first call's partials.
    for n in (5_000_003, 70_001, 4096, 1):
        log(f"noalloc scan n={n}")
        v = rng.integers(-(2**31), 2**31, size=n, dtype=np.int64).astype(np.int32)  # wraps
        ref, new = scan_both(gpu, v)
        np.testing.assert_array_equal(new, ref, err_msg=f"n={n}")
        np.testing.assert_array_equal(new, wrap_exclusive_scan(v), err_msg=f"n={n}")


def test_noalloc_scan_in_place_and_empty(gpu):
    lib = gg.lib()
    v = np.arange(1, 100_001, dtype=np.int32)
    d = gg.DeviceArray.from_numpy(gpu, v)
    _ffi.check(lib.gb_exclusive_scan_i32_noalloc(gpu, d.ptr, d.ptr, v.size))
    np.testing.assert_array_equal(d.numpy(), wrap_exclusive_scan(v))
    _ffi.check(lib.gb_exclusive_scan_i32_noalloc(gpu, d.ptr, d.ptr, 0))


def test_noalloc_scan_at_its_maximum_and_beyond(gpu):
    lib = gg.lib()
    nmax = lib.gb_exclusive_scan_noalloc_max()
    assert nmax == 2**30
    log(f"noalloc scan n={nmax} (maximum)")
    d = gg.DeviceArray.from_numpy(gpu, np.ones(nmax, dtype=np.int32))
    out = gg.DeviceArray(gpu, np.int32, nmax)
    _ffi.check(lib.gb_exclusive_scan_i32_noalloc(gpu, d.ptr, out.ptr, nmax))
    got = out.numpy()
    assert got[0] == 0 and got[-1] == nmax - 1
    np.testing.assert_array_equal(got, np.arange(nmax, dtype=np.int32))
    del got
    rc = lib.gb_exclusive_scan_i32_noalloc(gpu, d.ptr, out.ptr, nmax + 1)
    assert rc == _ffi.GB_EINVAL
    assert b"allocation-free maximum" in lib.gb_last_error()
