"""Independent pins of the oracle's time-step control, its RK2_SSP stage combination and its
choice of tanh (the checker of the HIP path; CPU only).

  - dt (scheme/ComputeDtFiveEqFunctor.cpp:85-134, SolverGodunovFiveEq.cpp:368-390): on uniform
    states whose sound speed is known in closed form (c = 1 for gamma = 1.4, p = 1, rho = 1.4),
    dt equals cfl / sum_d (c + |u_d|)/dx; on a two-state mesh it is set by the cell with the
    largest sum over directions, not the largest single direction, and ghost cells do not count.
  - RK2_SSP (SolverGodunovFiveEq.cpp:496-524): on a smooth periodic problem with a fixed mesh,
    halving dt shrinks the difference between successive solutions at the same final time about
    4x (second order in time); RK1 shrinks it about 2x.  A wrong stage buffer or combination
    coefficient makes RK2_SSP first order.
  - tanh: the oracle and the HIP path use a deterministic expm1-based tanh where the reference
    calls the C library's; the oracle built with std::tanh (oracle/Makefile, libg5oracle_libm_tanh)
    differs after whole THINC runs by at most 1e-12 relative L1 per conserved variable (measured
    3e-14 on the 3-D air-water case, 8e-15 on the 2-D one, 2e-14 on the 2-D air-helium one).
"""
import numpy as np
import pytest

import cases


def uniform_state(cfg, rho, p, vel, gamma=1.4):
    """Pure material 0 (alpha0 = 1), ideal gas: U[block][var][cell] of a uniform state."""
    dim = cfg.dim
    nblk = int(np.prod(list(cfg.nb)[:dim]))
    ncell = 16 ** dim
    e = p / (gamma - 1.0) + 0.5 * rho * sum(v * v for v in vel[:dim])
    vals = [rho, e] + [rho * v for v in vel[:dim]] + [1.0, rho, 0.0, 0.0]
    U = np.empty((nblk, len(vals), ncell))
    for k, v in enumerate(vals):
        U[:, k, :] = v
    return U


def ideal_cfg(g5, dim, nb, dx, **kw):
    opts = dict(periodic=(1, 1, 1), bc=("PERIODIC",) * 6, gamma=(1.4, 1.4), pinf=(0.0, 0.0), cfl=0.8)
    opts.update(kw)
    return g5.make_config(dim, nb, dx, **opts)


# ---------------------------------------------------------------- dt
@pytest.mark.parametrize("dim", [2, 3])
def test_dt_closed_form_uniform(g5, orc, dim):
    dx = 1.0 / 64
    cfg = ideal_cfg(g5, dim, (2, 2, 2), dx, cfl=0.8)
    vel = (0.5, -0.25, 2.0)  # c = sqrt(1.4 * 1 / 1.4) = 1; |v| enters with its sign dropped
    U = uniform_state(cfg, 1.4, 1.0, vel)
    o = orc.Oracle(cfg)
    o.set_state(U.ravel())
    rate = sum(1.0 + abs(v) for v in vel[:dim]) / dx  # 2-D: 2.75 * 64, 3-D: 5.75 * 64
    dt = o.compute_dt()
    assert dt == pytest.approx(0.8 / rate, rel=1e-14)
    assert o.compute_inv_dt() == pytest.approx(rate, rel=1e-14)


def test_dt_max_over_cells_of_sum_over_directions(g5, orc):
    # state A in the left blocks has the largest single-direction speed (c + |u| = 4, sum 5),
    # state B in the right blocks the largest sum over directions (3 + 3 = 6): dt follows B
    dx = 1.0 / 32
    cfg = ideal_cfg(g5, 2, (2, 1), dx, periodic=(0, 0, 0), bc=("WALL",) * 4 + ("NONE", "NONE"))
    UA = uniform_state(cfg, 1.4, 1.0, (3.0, 0.0, 0.0))
    UB = uniform_state(cfg, 1.4, 1.0, (2.0, -2.0, 0.0))
    U = UA.copy()
    U[1] = UB[1]
    o = orc.Oracle(cfg)
    o.set_state(U.ravel())
    assert o.compute_dt() == pytest.approx(0.8 / (6.0 / dx), rel=1e-14)
    # only owned cells count: the walls' mirror images (ghosts of the outside ring) of a faster
    # state do not; the same states swapped give the same dt
    U2 = UB.copy()
    U2[1] = UA[1]
    o.set_state(U2.ravel())
    assert o.compute_dt() == pytest.approx(0.8 / (6.0 / dx), rel=1e-14)
    # one fast cell inside a block sets the maximum
    U3 = UA.copy()
    U3[1] = UA[1]
    U3[1, :, 77] = UB[1, :, 77]
    o.set_state(U3.ravel())
    assert o.compute_dt() == pytest.approx(0.8 / (6.0 / dx), rel=1e-14)


# ---------------------------------------------------------------- RK2_SSP temporal order
def smooth_wave(g5, integrator, slope_type):
    nb = (4, 4)
    dx = 1.0 / 64
    cfg = ideal_cfg(g5, 2, nb, dx, integrator=integrator, slope_type=slope_type)
    i = np.arange(16)
    X = (np.arange(nb[0])[:, None] * 16 + i[None, :] + 0.5) * dx
    Y = (np.arange(nb[1])[:, None] * 16 + i[None, :] + 0.5) * dx
    x = X[None, :, None, :]
    y = Y[:, None, :, None]
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)
    ux, uy, p = 1.0, 0.5, 1.0
    E = p / 0.4 + 0.5 * rho * (ux * ux + uy * uy)
    U = np.zeros((nb[1], nb[0], 8, 16, 16))
    for v, f in enumerate([rho, E, rho * ux, rho * uy, 1.0, rho, 0.0, 0.0]):
        U[:, :, v] = f
    return cfg, U.reshape(nb[0] * nb[1], 8, 256).ravel()


def solution_at(orc, cfg, U0, nsteps, T):
    o = orc.Oracle(cfg)
    o.set_state(U0)
    for _ in range(nsteps):
        o.godunov_unsplit(T / nsteps)
    return o.get_state()


def temporal_ratios(g5, orc, integrator, slope_type):
    cfg, U0 = smooth_wave(g5, integrator, slope_type)
    T = 0.02  # 16 steps: sum-over-directions CFL about 0.6
    sols = [solution_at(orc, cfg, U0, n, T) for n in (16, 32, 64, 128)]
    d = [np.abs(sols[k] - sols[k + 1]).sum() for k in range(3)]
    return d[0] / d[1], d[1] / d[2]


@pytest.mark.parametrize("slope_type", [0.0, 1.0])
def test_rk2_ssp_is_second_order_in_time(g5, orc, slope_type):
    r1, r2 = temporal_ratios(g5, orc, "RK2_SSP", slope_type)  # measured 4.00-4.04
    assert 3.8 <= r1 <= 4.2 and 3.8 <= r2 <= 4.2, (r1, r2)


def test_rk1_is_first_order_in_time(g5, orc):
    r1, r2 = temporal_ratios(g5, orc, "RK1", 0.0)  # measured 2.00
    assert 1.9 <= r1 <= 2.1 and 1.9 <= r2 <= 2.1, (r1, r2)


def test_rk2_ssp_step_is_the_heun_composition_of_rk1_steps(g5, orc):
    # one RK2_SSP step of the oracle equals 1/2 U + 1/2 RK1(RK1(U)) assembled from two of its own
    # RK1 steps (each with its border fill), bit for bit
    cfg2, U0 = smooth_wave(g5, "RK2_SSP", 1.0)
    cfg1, _ = smooth_wave(g5, "RK1", 1.0)
    dt = 0.02 / 16
    a = solution_at(orc, cfg2, U0, 1, dt)
    u1 = solution_at(orc, cfg1, U0, 1, dt)
    u2 = solution_at(orc, cfg1, u1, 1, dt)
    assert np.array_equal(a, 0.5 * U0 + 0.5 * u2)


# ---------------------------------------------------------------- tanh
def _rel_l1(a, b, dim):
    nv = 8 if dim == 2 else 9
    A = a.reshape(-1, nv, 16 ** dim)
    B = b.reshape(-1, nv, 16 ** dim)
    return np.array([np.abs(A[:, v] - B[:, v]).sum() / np.abs(B[:, v]).sum() for v in range(nv)])


def _det_vs_libm(orc, cfg, U0, nsteps):
    out = []
    for variant in ("", "libm_tanh"):
        o = orc.Oracle(cfg, variant)
        o.set_state(U0)
        o.run(nsteps)
        out.append(o.get_state())
    return _rel_l1(out[0], out[1], cfg.dim)


def test_det_tanh_vs_libm_tanh_3d_air_water(orc):
    cfg, U0 = cases.air_water_3d(nb=4)
    err = _det_vs_libm(orc, cfg, U0, 50)
    assert err.max() <= 1e-12, err


def test_det_tanh_vs_libm_tanh_2d(g5, orc):
    rho2, u2, p2, _ = g5.rankine_hugoniot(4.4, 6e8, 1000.0, 1e5, p2=1e9)
    cfg = g5.make_config(2, (16, 16), 1.0 / 256, bc=("WALL",) * 4 + ("NONE", "NONE"), cfl=0.5,
                         gamma=(1.4, 4.4), pinf=(0.0, 6e8), thinc=True, thinc_beta=1.6,
                         thinc_epsilon=1e-3)
    regs = [g5.region(0.0, (0.0, rho2), p2, u2), g5.region(0.0, (0.0, 1000.0), 1e5),
            g5.region(1.0, (1.0, 0.0), 1e5)]
    U0 = g5.init_shock_bubble(cfg, regs, 0.25, (0.5, 0.5), 0.2)
    err = _det_vs_libm(orc, cfg, U0, 200)
    assert err.max() <= 1e-12, err
    cfg, U0 = cases.air_helium_2d(thinc=True, thinc_beta=1.6)
    err = _det_vs_libm(orc, cfg, U0, 200)
    assert err.max() <= 1e-12, err
