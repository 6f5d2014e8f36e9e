"""The caller's block order at the C ABI (g5_set_block_order, g5_do_time_step_buffers,
g5_compute_dt_buffer): a state handed over in the Morton order of the blocks' brick coordinates
(the order of an AMR manager's orchard keys, scheme/GodunovImplemV0.cpp:265-275) steps to the same
bits as the brick-ordered run, through the context's own state and through device arrays the
caller owns (allocated with hipMalloc)."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def morton_order(nb, dim):
    """brick coordinates (n, 3) of the blocks sorted by their Morton (bit-interleaved) key"""
    coords = []
    for z in range(nb[2] if dim == 3 else 1):
        for y in range(nb[1]):
            for x in range(nb[0]):
                key = 0
                for b in range(10):
                    key |= ((x >> b) & 1) << (dim * b)
                    key |= ((y >> b) & 1) << (dim * b + 1)
                    if dim == 3:
                        key |= ((z >> b) & 1) << (dim * b + 2)
                coords.append((key, x, y, z))
    coords.sort()
    return np.array([c[1:] for c in coords], dtype=np.int32)


def lex_index(coords, nb):
    return coords[:, 0] + nb[0] * (coords[:, 1] + nb[1] * coords[:, 2])


def to_order(U, coords, nb, nvar, ncell):
    """brick-ordered state -> the order of coords"""
    return U.reshape(-1, nvar, ncell)[lex_index(coords, nb)].ravel()


def from_order(U, coords, nb, nvar, ncell):
    out = np.empty_like(U.reshape(-1, nvar, ncell))
    out[lex_index(coords, nb)] = U.reshape(-1, nvar, ncell)
    return out.ravel()


class DeviceArray:
    """n doubles of device memory owned by the caller (hipMalloc of the HIP runtime)"""

    def __init__(self, n):
        import ctypes
        self._ct = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.n = n
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(8 * n)) == 0
        self.ptr = p.value

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.size == self.n
        assert self.hip.hipMemcpy(self._ct.c_void_p(self.ptr), self._ct.c_void_p(U.ctypes.data),
                                  self._ct.c_size_t(8 * self.n), 1) == 0  # host to device

    def download(self):
        U = np.empty(self.n)
        assert self.hip.hipMemcpy(self._ct.c_void_p(U.ctypes.data), self._ct.c_void_p(self.ptr),
                                  self._ct.c_size_t(8 * self.n), 2) == 0  # device to host
        return U

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.hip.hipFree(self._ct.c_void_p(self.ptr))


CASES = {
    "2d air-helium 8x4 blocks": lambda: cases.air_helium_2d(nbx=8, nby=4),
    "3d air-water THINC 4^3 blocks": lambda: cases.air_water_3d(nb=4),
    "3d air-water THINC RK2_SSP, walls": lambda: cases.air_water_3d(nb=3, integrator="RK2_SSP",
                                                                     bc=("WALL",) * 6),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_morton_order_equals_brick_order(g5, name):
    cfg, U0 = CASES[name]()
    dim = cfg.dim
    nb = [cfg.nb[0], cfg.nb[1], cfg.nb[2] if dim == 3 else 1]
    nvar, ncell = (8, 256) if dim == 2 else (9, 4096)
    coords = morton_order(nb, dim)
    assert not np.array_equal(lex_index(coords, nb), np.arange(len(coords)))  # a real permutation
    nsteps = 6
    say(f"{name}: brick order vs Morton order, {nsteps} steps")
    with g5.Context(cfg) as ctx:
        ctx.upload(U0)
        t_ref, _ = ctx.run(nsteps)
        ref = ctx.download()
    # the context's own state in the caller's order
    with g5.Context(cfg) as ctx:
        ctx.set_block_order(coords)
        ctx.upload(to_order(U0, coords, nb, nvar, ncell))
        t, _ = ctx.run(nsteps)
        got = ctx.download()
        assert t == t_ref and np.array_equal(from_order(got, coords, nb, nvar, ncell), ref)
    # the order set after the upload: the state moves with its blocks
    with g5.Context(cfg) as ctx:
        ctx.upload(U0)
        ctx.set_block_order(coords)
        assert np.array_equal(ctx.download(), to_order(U0, coords, nb, nvar, ncell))
    # caller-owned device arrays: dt from U, U2 <- step(U), swap (the caller's U <- U2)
    with g5.Context(cfg) as ctx, DeviceArray(U0.size) as A, DeviceArray(U0.size) as B:
        ctx.set_block_order(coords)
        ctx.upload(np.zeros_like(U0))  # the context's own state is not used by the array path
        A.upload(to_order(U0, coords, nb, nvar, ncell))
        U, U2 = A, B
        tt = 0.0
        for _ in range(nsteps):
            dt = ctx.compute_dt_buffer(U.ptr)
            ctx.do_time_step_buffers(U.ptr, U2.ptr, dt)
            ctx.synchronize()
            tt += dt
            U, U2 = U2, U
        got = U.download()
        assert tt == t_ref
        assert np.array_equal(from_order(got, coords, nb, nvar, ncell), ref)


def test_block_order_is_validated(g5):
    cfg, U0 = cases.air_helium_2d(nbx=4, nby=2)
    with g5.Context(cfg) as ctx:
        bad = morton_order([4, 2, 1], 2)
        bad[3] = bad[2]
        with pytest.raises(g5.G5Error) as e:
            ctx.set_block_order(bad)
        assert e.value.code == g5.G5_ERR_INVALID
        bad = morton_order([4, 2, 1], 2)
        bad[0, 0] = 4
        with pytest.raises(g5.G5Error):
            ctx.set_block_order(bad)
        with pytest.raises(g5.G5Error):
            ctx.set_block_order(morton_order([4, 2, 1], 2)[:7])
        with pytest.raises(g5.G5Error):
            ctx.do_time_step_buffers(0, 0, 1e-3)
