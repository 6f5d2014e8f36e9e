"""g5_step_noalloc (the step that allocates nothing) against the entry points it is a variant of:
g5_do_time_step_buffers followed by g5_compute_dt_buffer on the result.  Two calls in a row, the
second with a smaller input (the scratch keeps the last call's contents), must give bitwise the
same new state, next dt and failure bits; a size beyond the scratch maximum and RK2_SSP are
refused before anything runs."""
import ctypes

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


class DeviceArray:
    """n doubles of device memory owned by the test (hipMalloc of the HIP runtime)"""

    def __init__(self, n):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.n = n
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(8 * n)) == 0
        self.ptr = p.value

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.size == self.n
        assert self.hip.hipMemcpy(ctypes.c_void_p(self.ptr), ctypes.c_void_p(U.ctypes.data),
                                  ctypes.c_size_t(8 * self.n), 1) == 0

    def download(self):
        U = np.empty(self.n)
        assert self.hip.hipMemcpy(ctypes.c_void_p(U.ctypes.data), ctypes.c_void_p(self.ptr),
                                  ctypes.c_size_t(8 * self.n), 2) == 0
        return U

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.hip.hipFree(ctypes.c_void_p(self.ptr))


def reference(g5, cfg, U0, dt):
    """U2 and next dt through g5_do_time_step_buffers + g5_compute_dt_buffer"""
    with g5.Context(cfg) as ctx, DeviceArray(U0.size) as U, DeviceArray(U0.size) as U2:
        ctx.upload(np.zeros_like(U0))
        U.upload(U0)
        ctx.do_time_step_buffers(U.ptr, U2.ptr, dt)
        ctx.synchronize()
        dn = ctx.compute_dt_buffer(U2.ptr)
        return U2.download(), dn


def noalloc(g5, cfg, U0, dt):
    with DeviceArray(U0.size) as U, DeviceArray(U0.size) as U2:
        U.upload(U0)
        dn, f = g5.step_noalloc(cfg, U.ptr, U2.ptr, dt)
        return U2.download(), dn, f


def test_noalloc_step_equals_the_allocating_entry_points(g5):
    big = cases.air_water_3d(nb=4)                                   # 64 blocks of 16^3, THINC
    small = cases.air_helium_2d(nbx=6, nby=3, slope_type=2.0)        # 18 blocks of 16^2
    fast = cases.air_water_3d(nb=2, precision=g5.PREC_FP64_FAST)     # 8 blocks, reassociated FP64
    for label, (cfg, U0) in (("3-D 64 blocks", big), ("2-D 18 blocks", small), ("3-D 8 blocks fast", fast)):
        with g5.Context(cfg) as ctx:
            ctx.upload(U0)
            dt = ctx.compute_dt()
        ref, dref = reference(g5, cfg, U0, dt)
        got, dn, f = noalloc(g5, cfg, U0, dt)
        print(f"[gpu] no-alloc {label}: dt {dt:.6e} next {dn:.6e}", flush=True)
        assert f == 0 and dn == dref, label
        assert np.array_equal(got, ref), label


def test_noalloc_refusals(g5):
    cfg, U0 = cases.air_helium_2d(nbx=4, nby=2, integrator="RK2_SSP")
    with pytest.raises(g5.G5Error) as e:
        g5.step_noalloc(cfg, 1 << 20, 2 << 20, 1e-3)  # refused on the host, before any device access
    assert e.value.code == g5.G5_ERR_UNSUPPORTED
    cfg = g5.make_config(3, (64, 32, 17), 1.0 / 1024)  # 34816 blocks > 32768
    with pytest.raises(g5.G5Error) as e:
        g5.step_noalloc(cfg, 1 << 20, 2 << 20, 1e-3)
    assert e.value.code == g5.G5_ERR_TOO_LARGE
