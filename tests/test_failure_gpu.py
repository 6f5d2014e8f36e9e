"""Failure detection of diverging solves (every failure is planted from the host as input data;
the kernels index memory by block and cell coordinates only, so a NaN state runs like any other).

- A non-finite state (a NaN in one cell) stops its solve with G5_FAIL_NONFINITE, at the first dt
  when it is there from the start, after the next step when it appears mid-run; the solves
  batched with it in the same pool keep stepping and end bitwise equal to their alone runs.
- With abort_negative_pressure ([hydro] abort_when_negative_pressure) a negative pressure stops the
  solve with G5_FAIL_NEGATIVE_PRESSURE; without it a stiffened gas under tension (p < 0 < p + pinf)
  steps normally.
- A single-solve context reports the failure from g5_run and g5_compute_dt (G5_ERR_DIVERGED).
- A solve larger than a pool's whole budget is refused with G5_ERR_TOO_LARGE, not G5_ERR_NOMEM.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def three_solves(g5):
    a = cases.air_helium_2d(nbx=6, nby=4)
    b = cases.air_helium_2d(nbx=5, nby=3, slope_type=2.0)
    c = cases.air_helium_2d(nbx=4, nby=4, thinc=False, periodic=(1, 0, 0),
                            bc=("PERIODIC", "PERIODIC", "WALL", "WALL", "NONE", "NONE"))
    return [a, b, c]


def alone(g5, cfg, U0, nsteps):
    with g5.Pool(2, budget_bytes=1 << 26) as pool:
        sid = pool.admit(cfg, max_steps=nsteps)
        pool.upload(sid, U0)
        pool.start([sid])
        pool.step(nsteps)
        return pool.download(sid)


def test_nan_at_start_stops_only_its_solve(g5):
    solves = three_solves(g5)
    ref = [alone(g5, c, u, 10) for c, u in solves]
    say("NaN in one cell of solve 1 of 3 at the start")
    with g5.Pool(2, budget_bytes=1 << 26) as pool:
        ids = []
        for k, (cfg, U0) in enumerate(solves):
            sid = pool.admit(cfg, max_steps=10)
            U = U0.copy()
            if k == 1:
                U[8 * 256 * 2 + 37] = np.nan  # block 2, density of cell 37
            pool.upload(sid, U)
            ids.append(sid)
        pool.start(ids)
        pool.step(10)
        assert pool.failure(ids[1]) == g5.FAIL_NONFINITE
        with pytest.raises(g5.G5Error) as e:
            pool.status(ids[1])
        assert e.value.code == g5.G5_ERR_DIVERGED
        assert ids[1] in pool.finished()
        for k in (0, 2):
            assert pool.failure(ids[k]) == 0
            t, n, active = pool.status(ids[k])
            assert n == 10 and not active
            assert np.array_equal(pool.download(ids[k]), ref[k]), k


def test_nan_mid_run_stops_after_the_next_step(g5):
    solves = three_solves(g5)
    ref = [alone(g5, c, u, 12) for c, u in solves]
    say("NaN planted in solve 0 after 5 steps")
    with g5.Pool(2, budget_bytes=1 << 26) as pool:
        ids = []
        for cfg, U0 in solves:
            sid = pool.admit(cfg, max_steps=12)
            pool.upload(sid, U0)
            ids.append(sid)
        pool.start(ids)
        pool.step(5)
        U = pool.download(ids[0])
        U[3 * 8 * 256 + 1 * 256 + 100] = np.inf  # energy of a cell of block 3
        pool.upload(ids[0], U)
        pool.step(7)
        assert pool.failure(ids[0]) == g5.FAIL_NONFINITE
        with pytest.raises(g5.G5Error):
            pool.status(ids[0])
        # the failing step is counted, then the solve stops
        import ctypes
        t, n, a = ctypes.c_double(), ctypes.c_int32(), ctypes.c_int32()
        rc = g5.lib().g5_pool_status(pool._h, ids[0], ctypes.byref(t), ctypes.byref(n), ctypes.byref(a))
        assert rc == g5.G5_ERR_DIVERGED and n.value == 6 and a.value == 0
        for k in (1, 2):
            assert pool.failure(ids[k]) == 0
            assert np.array_equal(pool.download(ids[k]), ref[k]), k


def water_under_tension(g5, abort):
    # uniform water at rest (stiffened gas, pinf = 6e8) at 1e5 Pa with one cell at -1e5 Pa
    cfg, U0, pressure = cases.interface_advection_2d(alpha_min=0.0, alpha_max=0.0, pinf=6e8,
                                                     velocity=(0.0, 0.0), nb=(2, 2))
    cfg.abort_negative_pressure = 1 if abort else 0
    V = U0.reshape(-1, 8, 256).copy()
    xi, Pi = 1 / 3.4, 4.4 * 6e8 / 3.4
    V[1, 1, 40] = -1e5 * xi + Pi  # E of the cell at p = -1e5 (at rest, alpha0 = 0)
    assert pressure(V.ravel())[1, 40] == pytest.approx(-1e5, rel=1e-6)
    return cfg, V.ravel()


@pytest.mark.parametrize("abort", [True, False])
def test_negative_pressure(g5, abort):
    cfg, U0 = water_under_tension(g5, abort)
    say(f"negative pressure, abort_negative_pressure={abort}")
    with g5.Pool(2, budget_bytes=1 << 24) as pool:
        sid = pool.admit(cfg, max_steps=8)
        pool.upload(sid, U0)
        pool.start([sid])
        pool.step(8)
        if abort:
            assert pool.failure(sid) == g5.FAIL_NEGATIVE_PRESSURE
            with pytest.raises(g5.G5Error):
                pool.status(sid)
        else:
            assert pool.failure(sid) == 0
            t, n, active = pool.status(sid)
            assert n == 8 and np.all(np.isfinite(pool.download(sid)))


def test_context_reports_divergence(g5):
    cfg, U0 = cases.air_water_3d(nb=2)
    U = U0.copy()
    U[5 * 9 * 4096 + 0 * 4096 + 2000] = np.nan
    say("context: g5_run and g5_compute_dt on a NaN state")
    with g5.Context(cfg) as ctx:
        ctx.upload(U)
        with pytest.raises(g5.G5Error) as e:
            ctx.compute_dt()
        assert e.value.code == g5.G5_ERR_DIVERGED
        with pytest.raises(g5.G5Error) as e:
            ctx.run(5)
        assert e.value.code == g5.G5_ERR_DIVERGED and "non-finite" in str(e.value)
        # a healthy state uploaded afterwards runs again
        ctx.upload(U0)
        t, n = ctx.run(3)
        assert n == 3


def test_solve_larger_than_the_pool_is_too_large(g5):
    cfg, U0 = cases.air_helium_2d(nbx=8, nby=8)  # 64 blocks, 1 MiB per state copy
    with g5.Pool(2, budget_bytes=1 << 20) as pool:  # room for 32 blocks of two copies
        with pytest.raises(g5.G5Error) as e:
            pool.admit(cfg)
        assert e.value.code == g5.G5_ERR_TOO_LARGE
        small, _ = cases.air_helium_2d(nbx=4, nby=4)
        other, _ = cases.air_helium_2d(nbx=5, nby=4)
        sid = pool.admit(small)
        assert pool.try_admit(other) is None  # 16 + 20 blocks do not fit now: NOMEM
        pool.release(sid)
        assert pool.try_admit(other) is not None
