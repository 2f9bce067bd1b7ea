"""hg_conservation_sums_noalloc, the entry point that allocates nothing (its partial sums in the
library's fixed device scratch array, the result written to a device buffer the caller owns,
no synchronisation), against hg_conservation_sums on the same handle: bit for bit, on two calls
in a row, the second on a smaller mesh, and again after the first mesh's state has changed; a
mesh beyond the array's size is refused."""
import ctypes

import numpy as np
import pytest

from horses2d_amd import cases as C

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Four doubles of device memory from the HIP runtime the library itself uses (the test owns
    this allocation; the entry point under test allocates nothing)."""

    def __init__(self, fill):
        from horses2d_amd import gpu
        gpu.load_library()
        self.hip = ctypes.CDLL("libamdhip64.so.7")   # the runtime already loaded with the library
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), 32) == 0
        self.write(np.full(4, fill))

    def write(self, values):
        host = np.ascontiguousarray(values, dtype=np.float64)
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data, 32, 1) == 0    # host to device

    def read(self):
        host = np.empty(4)
        assert self.hip.hipMemcpy(host.ctypes.data, self.ptr, 32, 2) == 0    # device to host
        return host

    def free(self):
        self.hip.hipFree(self.ptr)


def device_sums(dg, out):
    """Calls the entry point through the C ABI with a device output buffer, then reads it back
    after the handle's stream is idle."""
    from horses2d_amd import gpu
    L = gpu.load_library()
    rc = L.hg_conservation_sums_noalloc(dg.handle, out.ptr)
    assert rc == 0, gpu.last_error()
    dg.synchronize()
    return out.read()


def test_noalloc_conservation_sums_match_the_allocating_entry_point():
    from horses2d_amd import gpu
    out = DeviceBuffer(np.nan)
    big = gpu.DGSEM(C.ns_kh_case(64, 7), "fp64")            # 4096 elements, 1024 partial sums
    small = gpu.DGSEM(C.euler_vortex_case(5, 3), "fp64")     # 25 elements, 2 partial sums
    a = device_sums(big, out)
    np.testing.assert_array_equal(a, big.conservation_sums())
    b = device_sums(small, out)
    np.testing.assert_array_equal(b, small.conservation_sums())
    assert not np.array_equal(a, b)
    dt = big.estimate_dt(1.0, big.case.cfl)
    big.rk_step(0.0, dt)
    c = device_sums(big, out)
    np.testing.assert_array_equal(c, big.conservation_sums())
    out.free()


def test_noalloc_rejects_a_mesh_beyond_its_scratch_array():
    """At N = 10 the sum kernel takes two elements per workgroup: 363 x 363 elements need 65885
    workgroups' partial sums, above the array's 65536.  The call fails with a message and leaves
    the output buffer untouched."""
    from horses2d_amd import gpu
    L = gpu.load_library()
    out = DeviceBuffer(-7.0)
    dg = gpu.DGSEM(C.euler_vortex_case(363, 10), "fp64")
    rc = L.hg_conservation_sums_noalloc(dg.handle, out.ptr)
    assert rc != 0
    assert "partial sums" in gpu.last_error()
    dg.synchronize()
    assert (out.read() == -7.0).all()
    out.free()
    dg.close()
