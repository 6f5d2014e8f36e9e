"""Two-level static mesh in 3-D on the GPU (g5_create_two_level with a 3-D configuration) against
the oracle's 3-D two-level mesh (pinned in tests/test_oracle_amr.py: uniform state, closed forms on
linear fields, conservation across the levels, children averages).

The coupling is that of the 2-D tests (tests/test_amr_gpu.py) with eight-child restriction, the
3-D prolongations, four fine faces per coarse face summed in the reference's order and divided by
8, and z faces between the levels at the bottom and top of the z-march.  Bar: the FP64 path
bit-identical to the oracle (asserted as relative L1 <= 1e-12 per conserved variable and exactly),
the reassociated FP64 path within relative L1 1e-12 (momentum components against the momentum
vector's norm).  The refined region is a 3-D corner of four blocks, so coarse cells meet fine
ones on faces of all three orientations, and the shock crosses into it during the run.
"""
import ctypes
import sys

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

RTOL_L1 = 1e-12
CORNER = [[1, 1, 1], [2, 1, 1], [1, 2, 1], [1, 1, 2]]


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def rel_l1(a, b, momentum_norm=False):
    A = a.reshape(-1, 9, 4096)
    B = b.reshape(-1, 9, 4096)
    mom_den = sum(np.abs(B[:, m]).sum() for m in (2, 3, 4))
    out = []
    for v in range(9):
        den = mom_den if (momentum_norm and v in (2, 3, 4)) else np.abs(B[:, v]).sum()
        out.append(np.abs(A[:, v] - B[:, v]).sum() / den if den > 0 else np.abs(A[:, v] - B[:, v]).sum())
    return np.array(out)


def two_level_state3(nb, layout, U_coarse, U_fine):
    Uc = U_coarse.reshape(-1, 9, 4096)
    Uf = U_fine.reshape(-1, 9, 4096)
    out = []
    for x, y, z, lv in layout:
        out.append(Uc[x + nb * (y + nb * z)] if lv == 0 else Uf[x + 2 * nb * (y + 2 * nb * z)])
    return np.stack(out).ravel()


def case3(g5, maker, refined, prolongation, **kw):
    cfg, Uc = maker(nb=4, **kw)
    _, Uf = maker(nb=8, **kw)
    ctx = g5.Context(cfg, refined=refined, prolongation=prolongation)
    return cfg, ctx, two_level_state3(4, ctx.block_layout(), Uc, Uf)


def oracle3(orc, cfg, refined, prolongation):
    return orc.Oracle(cfg, refined=np.array(refined, dtype=np.int32), prolongation=prolongation)


def test_3d_layout_matches_the_oracle(g5, orc):
    cfg, ctx, _ = case3(g5, cases.air_helium_3d, CORNER, 0)
    with ctx:
        lay = ctx.block_layout()
        o = oracle3(orc, cfg, CORNER, 0)
        assert ctx.nblocks == o.nowned == 60 + 32
        assert np.array_equal(lay[:, 3], o.block_levels()[: o.nowned])
        assert np.array_equal(lay[:, :3], o.block_coords()[: o.nowned])


@pytest.mark.parametrize("prolongation", [0, 1, 2, 3])
def test_3d_shock_bubble_two_level(g5, orc, prolongation):
    # 45 steps: the Mach 1.22 shock enters the refined corner.  With the order-4 interpolation,
    # which is not monotone, the helium density of the 3-D stencil's ghosts reaches 0 and the state
    # turns non-finite after 5 steps in the oracle too: the GPU stops the solve with
    # G5_ERR_DIVERGED after that same step.
    label = f"3-D two-level air-helium prolongation={prolongation}"
    _announce(label)
    steps = 45
    cfg, ctx, U0 = case3(g5, cases.air_helium_3d, CORNER, prolongation)
    o = oracle3(orc, cfg, CORNER, prolongation)
    o.set_state(U0)
    with ctx:
        ctx.upload(U0)
        assert ctx.compute_dt() == o.compute_dt()
        tt, done = ctypes.c_double(0.0), ctypes.c_int32()
        rc = g5.lib().g5_run(ctx._h, steps, ctypes.byref(tt), 1e300, ctypes.byref(done))
        t_g, n_g = tt.value, done.value
        U_g = ctx.download()
    if rc == g5.G5_ERR_DIVERGED:
        for k in range(1, n_g + 1):
            o.godunov_unsplit(o.compute_dt())
            assert np.all(np.isfinite(o.get_state())) == (k < n_g), (k, n_g)
        _announce(f"{label}: stopped after {n_g} steps, as the oracle's state turns non-finite")
        assert prolongation == 3 and n_g == 5
        return
    assert rc == g5.G5_OK, rc
    t_o, n_o = o.run(steps)
    U_o = o.get_state()
    assert n_g == n_o == steps and t_g == t_o, (t_g, t_o)
    fine = np.repeat(o.block_levels()[: o.nowned], 9 * 4096) == 1
    assert np.abs(U_g - U0)[fine].max() > 0.1  # the shock is in the fine blocks
    err = rel_l1(U_g, U_o)
    exact = np.array_equal(U_g, U_o)
    _announce(f"{label}: max rel L1 = {err.max():.3e} bit-exact={exact}")
    assert err.max() <= RTOL_L1 and exact, err


def test_3d_thinc_walls_host_dt(g5, orc):
    label = "3-D two-level air-water THINC walls, host dt"
    _announce(label)
    cfg, ctx, U0 = case3(g5, cases.air_water_3d, CORNER, 1, bc=("WALL",) * 6)
    o = oracle3(orc, cfg, CORNER, 1)
    o.set_state(U0)
    with ctx:
        ctx.upload(U0)
        for _ in range(8):
            dt = o.compute_dt()
            o.godunov_unsplit(dt)
            ctx.do_time_step(dt)
        U_g = ctx.download()
    err = rel_l1(U_g, o.get_state())
    exact = np.array_equal(U_g, o.get_state())
    _announce(f"{label}: max rel L1 = {err.max():.3e} bit-exact={exact}")
    assert err.max() <= RTOL_L1 and exact, err


def test_3d_rk2_periodic(g5, orc):
    label = "3-D two-level RK2_SSP periodic, MILLER_PUCKETT"
    _announce(label)
    cfg, ctx, U0 = case3(g5, cases.air_helium_3d, [[1, 1, 1], [2, 2, 2]], 2, integrator="RK2_SSP",
                         periodic=(1, 1, 1), bc=("PERIODIC",) * 6)
    cfg.source_term = g5.SOURCE_MILLER_PUCKETT
    ctx.close()
    ctx = g5.Context(cfg, refined=[[1, 1, 1], [2, 2, 2]], prolongation=2)
    o = oracle3(orc, cfg, [[1, 1, 1], [2, 2, 2]], 2)
    o.set_state(U0)
    with ctx:
        ctx.upload(U0)
        t_g, n_g = ctx.run(12)
        U_g = ctx.download()
    t_o, n_o = o.run(12)
    err = rel_l1(U_g, o.get_state())
    exact = np.array_equal(U_g, o.get_state())
    _announce(f"{label}: max rel L1 = {err.max():.3e} bit-exact={exact}")
    assert n_g == n_o == 12 and t_g == t_o and err.max() <= RTOL_L1 and exact, err


@pytest.mark.parametrize("prolongation", [0, 1])
def test_3d_fast_path_two_level(g5, orc, prolongation):
    label = f"3-D two-level FP64_FAST prolongation={prolongation}"
    _announce(label)
    cfg, ctx, U0 = case3(g5, cases.air_helium_3d, CORNER, prolongation, precision=g5.PREC_FP64_FAST)
    o = oracle3(orc, cfg, CORNER, prolongation)
    o.set_state(U0)
    with ctx:
        ctx.upload(U0)
        t_g, _ = ctx.run(45)
        U_g = ctx.download()
    t_o, _ = o.run(45)
    err = rel_l1(U_g, o.get_state(), momentum_norm=True)
    _announce(f"{label}: max rel L1 = {err.max():.3e}")
    assert err.max() <= RTOL_L1 and abs(t_g - t_o) <= RTOL_L1 * t_o, err
