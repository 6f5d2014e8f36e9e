"""Two-level static mesh on the GPU (g5_create_two_level) against the oracle's two-level mesh.

The coupling under test is the reference's AMR path of the V0 step on a 2-D mesh whose refined
blocks are fixed: ghosts of a coarse block inside a refined neighbour come from the restriction
of the four fine cells, ghosts of a fine block outside the fine blocks from the prolongation of
the coarse parent (SIMPLE_COPY, EXTRAPOLATE_LINEAR_MINMOD, or the conservative interpolations of
order 2 and 4) (scheme/ConvertToPrimitivesVariablesFunctor.cpp:214-250,339-490,553-827,884-918); a coarse cell on a face
shared with fine blocks is updated with the sum of the fine faces' fluxes over 4
(scheme/ReadFluxesAndConservativeUpdateFunctor.cpp:137-199, .h:256-307); fluxes and the CFL rate
use each level's cell length.  The oracle's restatement of it is pinned in
tests/test_oracle_amr.py (closed forms on linear fields, uniform-state preservation, conservation
across the levels).

Bars: the FP64 path bit-identical to the oracle (asserted as relative L1 <= 1e-12 per conserved
variable, the FP64 parity tolerance, and reported exact or not), the same final time; the
reassociated FP64 path within relative L1 1e-12 (momentum components against the momentum
vector's L1 norm) and conserving mass, energy and momentum to round-off in a periodic box.
Refined regions are L-shaped, so coarse cells meet fine ones on faces of both orientations and
at a concave corner, and the moving features (a shock, an advected material blob) cross the level
interface during the run.
"""
import ctypes
import sys

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

RTOL_L1 = 1e-12
L_SHAPE_8x4 = [[3, 1], [4, 1], [3, 2]]
L_SHAPE_6x6 = [[2, 2], [3, 2], [2, 3]]


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def rel_l1(a, b, momentum_norm=False):
    A = a.reshape(-1, 8, 256)
    B = b.reshape(-1, 8, 256)
    mom_den = np.abs(B[:, 2]).sum() + np.abs(B[:, 3]).sum()
    out = []
    for v in range(8):
        den = mom_den if (momentum_norm and v in (2, 3)) else np.abs(B[:, v]).sum()
        out.append(np.abs(A[:, v] - B[:, v]).sum() / den if den > 0 else np.abs(A[:, v] - B[:, v]).sum())
    return np.array(out)


def two_level_state(nbx, layout, U_coarse, U_fine):
    """The state of a two-level mesh in the context's block order from the one-level states on the
    coarse grid (nbx x nby blocks) and on the fine grid (2nbx x 2nby blocks of half the cell length)"""
    Uc = U_coarse.reshape(-1, 8, 256)
    Uf = U_fine.reshape(-1, 8, 256)
    return np.stack([Uc[x + nbx * y] if lv == 0 else Uf[x + 2 * nbx * y] for x, y, _, lv in layout]).ravel()


def shock_bubble_two_level(g5, refined, prolongation, **kw):
    cfg, Uc = cases.air_helium_2d(nbx=8, nby=4, **kw)
    _, Uf = cases.air_helium_2d(nbx=16, nby=8, **kw)
    ctx = g5.Context(cfg, refined=refined, prolongation=prolongation)
    return cfg, ctx, two_level_state(8, ctx.block_layout(), Uc, Uf)


def oracle_of(orc, cfg, refined, prolongation):
    r = np.array([[x, y, 0] for x, y in refined], dtype=np.int32)
    return orc.Oracle(cfg, refined=r, prolongation=prolongation)


def test_layout_matches_the_oracle(g5, orc):
    cfg, ctx, _ = shock_bubble_two_level(g5, L_SHAPE_8x4, 0)
    with ctx:
        lay = ctx.block_layout()
        o = oracle_of(orc, cfg, L_SHAPE_8x4, 0)
        assert ctx.nblocks == o.nowned == 29 + 12
        assert np.array_equal(lay[:, 3], o.block_levels()[: o.nowned])
        assert np.array_equal(lay[:, :2], o.block_coords()[: o.nowned, :2])


@pytest.mark.parametrize("prolongation", [0, 1, 2, 3])
def test_shock_crosses_the_level_interface(g5, orc, prolongation):
    # air-helium shock-bubble on an 8x4 brick (256x64 coarse cells) with an L-shaped refined region
    # around the bubble: the Mach 1.22 shock enters the fine blocks after about 100 steps
    label = f"two-level shock-bubble prolongation={prolongation}"
    _announce(label)
    cfg, ctx, U0 = shock_bubble_two_level(g5, L_SHAPE_8x4, prolongation)
    o = oracle_of(orc, cfg, L_SHAPE_8x4, prolongation)
    o.set_state(U0)
    with ctx:
        ctx.upload(U0)
        dt_g = ctx.compute_dt()
        assert dt_g == o.compute_dt(), (dt_g, o.compute_dt())
        t_g, n_g = ctx.run(150)
        U_g = ctx.download()
    t_o, n_o = o.run(150)
    U_o = o.get_state()
    assert n_g == n_o == 150 and t_g == t_o, (t_g, t_o)
    assert np.all(np.isfinite(U_g))
    # the shock is inside the fine blocks: their state has changed by O(1)
    fine = np.repeat(o.block_levels()[: o.nowned], 8 * 256) == 1
    assert np.abs(U_g - U0)[fine].max() > 0.1
    err = rel_l1(U_g, U_o)
    exact = np.array_equal(U_g, U_o)
    _announce(f"{label}: max rel L1 = {err.max():.3e} bit-exact={exact}")
    assert err.max() <= RTOL_L1, err
    assert exact


def test_host_dt_steps_thinc_walls(g5, orc):
    # THINC on the two-level mesh with walls on every side, stepped with host dts (do_time_step)
    label = "two-level THINC host-dt"
    _announce(label)
    kw = dict(thinc=True, thinc_beta=1.6, bc=("WALL",) * 4 + ("NONE",) * 2)
    cfg, ctx, U0 = shock_bubble_two_level(g5, L_SHAPE_8x4, 1, **kw)
    o = oracle_of(orc, cfg, L_SHAPE_8x4, 1)
    o.set_state(U0)
    with ctx:
        ctx.upload(U0)
        for _ in range(25):
            dt = o.compute_dt()
            o.godunov_unsplit(dt)
            ctx.do_time_step(dt)
        U_g = ctx.download()
    U_o = o.get_state()
    err = rel_l1(U_g, U_o)
    exact = np.array_equal(U_g, U_o)
    _announce(f"{label}: max rel L1 = {err.max():.3e} bit-exact={exact}")
    assert err.max() <= RTOL_L1 and exact, err


def advection_two_level(g5, refined, prolongation, **kw):
    cfg, Uc, pressure = cases.interface_advection_2d(nb=(6, 6), **kw)
    _, Uf, _ = cases.interface_advection_2d(nb=(12, 12), **kw)
    ctx = g5.Context(cfg, refined=refined, prolongation=prolongation)
    return cfg, ctx, two_level_state(6, ctx.block_layout(), Uc, Uf)


def test_rk2_thinc_periodic_advection(g5, orc):
    # a material blob advected diagonally through an L-shaped refined region, THINC, RK2_SSP,
    # periodic in x and y, minmod prolongation
    label = "two-level periodic advection RK2_SSP THINC"
    _announce(label)
    cfg, ctx, U0 = advection_two_level(g5, L_SHAPE_6x6, 1, thinc=True, integrator="RK2_SSP")
    o = oracle_of(orc, cfg, L_SHAPE_6x6, 1)
    o.set_state(U0)
    with ctx:
        ctx.upload(U0)
        t_g, n_g = ctx.run(30)
        U_g = ctx.download()
    t_o, n_o = o.run(30)
    U_o = o.get_state()
    assert n_g == n_o == 30 and t_g == t_o
    err = rel_l1(U_g, U_o)
    exact = np.array_equal(U_g, U_o)
    _announce(f"{label}: max rel L1 = {err.max():.3e} bit-exact={exact}")
    assert err.max() <= RTOL_L1 and exact, err


def totals(ctx, cfg, U):
    lv = ctx.block_layout()[:, 3]
    V = U.reshape(-1, 8, 256)
    vol = (cfg.dx / 2.0 ** lv) ** 2
    return np.array([(V[:, v].sum(axis=1) * vol).sum() for v in range(8)])


@pytest.mark.parametrize("prolongation", [0, 1, 2])
def test_fast_path_two_level(g5, orc, prolongation):
    # the reassociated FP64 path on both two-level cases: the 1e-12 bar against the oracle, and
    # conservation of mass, energy and momentum across the levels in the periodic box.  (With the
    # order-4 conservative interpolation, which is not monotone, the shock case is sensitive to
    # round-off: the bit-exact path and the reassociated one agree to 1e-17 after 10 steps and
    # then separate by about 1.6x per step, 1e-5 after 60 -- as two round-off seeds of the
    # bit-exact path would.  The bit-exact path still equals the oracle there (above); the
    # reassociated path's order-4 bar is checked on the smooth case below.)
    label = f"two-level FP64_FAST prolongation={prolongation}"
    _announce(label)
    cfg, ctx, U0 = shock_bubble_two_level(g5, L_SHAPE_8x4, prolongation, precision=g5.PREC_FP64_FAST)
    o = oracle_of(orc, cfg, L_SHAPE_8x4, prolongation)
    o.set_state(U0)
    with ctx:
        ctx.upload(U0)
        t_g, _ = ctx.run(150)
        U_g = ctx.download()
    t_o, _ = o.run(150)
    err = rel_l1(U_g, o.get_state(), momentum_norm=True)
    _announce(f"{label} shock-bubble: max rel L1 = {err.max():.3e}")
    assert err.max() <= RTOL_L1 and abs(t_g - t_o) <= RTOL_L1 * t_o, err
    cfg, ctx, U0 = advection_two_level(g5, L_SHAPE_6x6, prolongation, precision=g5.PREC_FP64_FAST)
    with ctx:
        ctx.upload(U0)
        ctx.run(40)
        U_g = ctx.download()
        T0, T1 = totals(ctx, cfg, U0), totals(ctx, cfg, U_g)
    for v in (0, 1, 2, 3):
        assert abs(T1[v] - T0[v]) <= 1e-14 * abs(T0[v]), (v, T0[v], T1[v])


@pytest.mark.parametrize("prolongation", [2, 3])
def test_fast_path_conservative_interpolation_smooth(g5, orc, prolongation):
    # the advected blob through the L-shaped refined region with the conservative interpolations:
    # the reassociated path within the 1e-12 bar of the oracle
    label = f"two-level FP64_FAST advection prolongation={prolongation}"
    _announce(label)
    cfg, ctx, U0 = advection_two_level(g5, L_SHAPE_6x6, prolongation, precision=g5.PREC_FP64_FAST)
    o = oracle_of(orc, cfg, L_SHAPE_6x6, prolongation)
    o.set_state(U0)
    with ctx:
        ctx.upload(U0)
        t_g, _ = ctx.run(40)
        U_g = ctx.download()
    t_o, _ = o.run(40)
    err = rel_l1(U_g, o.get_state(), momentum_norm=True)
    _announce(f"{label}: max rel L1 = {err.max():.3e}")
    assert err.max() <= RTOL_L1 and abs(t_g - t_o) <= RTOL_L1 * t_o, err


def random_two_level(g5, seed):
    """A seeded two-level configuration: a 2-D brick of 4-7 x 3-6 blocks with 1-5 refined blocks
    (some adjacent) at least one block inside its border, per-face ZERO_GRADIENT / WALL or periodic
    directions, THINC or MUSCL, minmod or MC, HANCOCK / RK1 / RK2_SSP, one of the four
    prolongations, gases or a stiffened gas, sometimes the MILLER_PUCKETT source term; a
    shock-bubble state sampled on both levels.  Returns (cfg, refined, prolongation, U0 on the
    coarse and fine grids, steps, label)."""
    rng = np.random.default_rng(1000 + seed)
    nbx, nby = int(rng.integers(4, 8)), int(rng.integers(3, 7))
    inner = [(x, y) for x in range(1, nbx - 1) for y in range(1, nby - 1)]
    k = int(rng.integers(1, min(5, len(inner)) + 1))
    refined = [list(inner[i]) for i in sorted(rng.choice(len(inner), size=k, replace=False))]
    periodic = [int(rng.random() < 0.3), int(rng.random() < 0.3), 0]
    bc = []
    for d in range(2):
        for _ in range(2):
            bc.append("PERIODIC" if periodic[d] else ("WALL" if rng.random() < 0.5 else "ZERO_GRADIENT"))
    bc += ["NONE", "NONE"]
    thinc = bool(rng.random() < 0.5)
    slope = 2.0 if rng.random() < 0.4 else 1.0
    integrator = str(rng.choice(["HANCOCK", "HANCOCK", "RK1", "RK2_SSP"]))
    prolong = int(rng.integers(0, 4))
    water = bool(rng.random() < 0.3)
    mp = bool(rng.random() < 0.25) or seed % 5 == 4
    kw = dict(periodic=periodic, bc=tuple(bc), thinc=thinc, slope_type=slope, integrator=integrator,
              cfl=float(rng.uniform(0.3, 0.8)), thinc_beta=float(rng.uniform(1.2, 2.0)))
    if water:
        kw.update(gamma=(1.4, 4.4), pinf=(0.0, 6e8))
        rho2, u2, p2, _ = g5.rankine_hugoniot(4.4, 6e8, 1000.0, 1e5, p2=float(rng.uniform(2e8, 1e9)))
        regs = [g5.region(0.0, (0.0, rho2), p2, u2, float(rng.uniform(-20, 20))), g5.region(0.0, (0.0, 1000.0), 1e5),
                g5.region(1.0, (1.0, 0.0), 1e5)]
    else:
        kw.update(gamma=(1.4, float(rng.uniform(1.1, 1.9))), pinf=(0.0, 0.0))
        rho2, u2, p2, _ = g5.rankine_hugoniot(1.4, 0.0, 1.0, 1.0, mach=float(rng.uniform(1.1, 3.0)))
        regs = [g5.region(1.0, (rho2, 0.0), p2, u2, float(rng.uniform(-0.3, 0.3))), g5.region(1.0, (1.0, 0.0), 1.0),
                g5.region(0.0, (0.0, float(rng.uniform(0.1, 3.0))), 1.0)]
    dx = 1.0 / (16 * nby)
    ext = (16 * nbx * dx, 16 * nby * dx)
    centre = (float(rng.uniform(0.3, 0.7)) * ext[0], float(rng.uniform(0.3, 0.7)) * ext[1])
    radius = float(rng.uniform(0.15, 0.3)) * ext[1]
    front = float(rng.uniform(0.05, 0.35)) * ext[0]
    cfg = g5.make_config(2, (nbx, nby), dx, **kw)
    cfg_f = g5.make_config(2, (2 * nbx, 2 * nby), dx / 2, **kw)
    if mp:
        cfg.source_term = cfg_f.source_term = g5.SOURCE_MILLER_PUCKETT
    Uc = g5.init_shock_bubble(cfg, regs, front, centre, radius)
    Uf = g5.init_shock_bubble(cfg_f, regs, front, centre, radius)
    steps = int(rng.integers(6, 16))
    label = (f"seed {seed}: nb=({nbx},{nby}) refined={refined} periodic={periodic[:2]} bc={bc[:4]} thinc={thinc} "
             f"slope={slope} {integrator} prolong={prolong} water={water} mp={mp} steps={steps}")
    return cfg, refined, prolong, Uc, Uf, steps, label


@pytest.mark.parametrize("seed", range(16))
def test_random_two_level_configurations(g5, orc, seed):
    # bit-exact against the oracle; a configuration whose state turns non-finite (the order-4
    # interpolation, which is not monotone, across a shock into water: seeds 1 and 9) stops with
    # G5_ERR_DIVERGED after the step whose state the oracle also finds non-finite
    cfg, refined, prolong, Uc, Uf, steps, label = random_two_level(g5, seed)
    _announce(label)
    with g5.Context(cfg, refined=refined, prolongation=prolong) as ctx:
        U0 = two_level_state(cfg.nb[0], ctx.block_layout(), Uc, Uf)
        ctx.upload(U0)
        tt, done = ctypes.c_double(0.0), ctypes.c_int32()
        rc = g5.lib().g5_run(ctx._h, steps, ctypes.byref(tt), 1e300, ctypes.byref(done))
        t_g, n_g = tt.value, done.value
        U_g = ctx.download()
    o = oracle_of(orc, cfg, refined, prolong)
    o.set_state(U0)
    if rc == g5.G5_ERR_DIVERGED:
        for k in range(1, n_g + 1):
            o.godunov_unsplit(o.compute_dt())
            assert np.all(np.isfinite(o.get_state())) == (k < n_g), (label, k, n_g)
        assert prolong == 3 and n_g < steps
        return
    assert rc == g5.G5_OK, rc
    t_o, n_o = o.run(steps)
    U_o = o.get_state()
    assert n_g == n_o == steps and t_g == t_o, (t_g, t_o)
    err = rel_l1(U_g, U_o)
    assert err.max() <= RTOL_L1 and np.array_equal(U_g, U_o), (label, err)


def test_two_level_solves_batched_with_one_level_ones(g5):
    # one pool, one launch list: a one-level shock-bubble, the two-level shock-bubble and the
    # two-level periodic advection (admitted after a released solve, so its slots are not one
    # run); each ends bitwise equal to its run alone in a context of its own
    cfg_a, Ua = cases.air_helium_2d(nbx=8, nby=4)
    cfg_b, ctx_b, Ub = shock_bubble_two_level(g5, L_SHAPE_8x4, 1)
    cfg_c, ctx_c, Uc = advection_two_level(g5, L_SHAPE_6x6, 2)
    steps = 40
    alone = {}
    for name, ctx, U0 in (("b", ctx_b, Ub), ("c", ctx_c, Uc)):
        with ctx:
            ctx.upload(U0)
            ctx.run(steps)
            alone[name] = ctx.download()
    with g5.Context(cfg_a) as ctx:
        ctx.upload(Ua)
        ctx.run(steps)
        alone["a"] = ctx.download()
    _announce("pool: one-level + two two-level solves, fragmented slots")
    with g5.Pool(2, precision=g5.PREC_FP64, budget_bytes=64 << 20) as pool:
        filler = pool.admit(g5.make_config(2, (2, 2), 1.0 / 32))
        sa = pool.admit(cfg_a, max_steps=steps)
        sb = pool.admit_two_level(cfg_b, L_SHAPE_8x4, 1, max_steps=steps)
        pool.release(filler)
        sc = pool.admit_two_level(cfg_c, L_SHAPE_6x6, 2, max_steps=steps)
        assert len(pool.block_layout(sc)) == 33 + 12 and pool.block_layout(sb)[-1, 3] == 1
        for sid, U0 in ((sa, Ua), (sb, Ub), (sc, Uc)):
            pool.upload(sid, U0)
            assert np.array_equal(pool.download(sid), U0)
        pool.start([sa, sb, sc])
        pool.step(steps + 2)
        for sid, name in ((sa, "a"), (sb, "b"), (sc, "c")):
            t, n, active = pool.status(sid)
            assert n == steps and not active
            assert np.array_equal(pool.download(sid), alone[name]), name
