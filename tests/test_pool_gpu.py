"""Batched solves in one pool (the batching and serving forms): each solve is bit-identical to
the same solve run alone and to the oracle, whatever else shares the launch, including solves
admitted while others are in flight; memory is reused after release; the budget is enforced;
the device initial condition equals the host one."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def mixed_2d_solves(g5):
    a = cases.air_helium_2d(nbx=8, nby=4)
    b = cases.air_helium_2d(nbx=5, nby=3, slope_type=2.0, periodic=(1, 0, 0),
                            bc=("PERIODIC", "PERIODIC", "WALL", "WALL", "NONE", "NONE"))
    rho2, u2, p2, _ = g5.rankine_hugoniot(4.4, 6e8, 1000.0, 1e5, p2=1e9)
    cfg = g5.make_config(2, (6, 6), 1.0 / 96, bc=("ZERO_GRADIENT",) * 4 + ("NONE", "NONE"), cfl=0.5,
                         gamma=(1.4, 4.4), pinf=(0.0, 6e8))
    regs = [g5.region(0.0, (0.0, rho2), p2, u2), g5.region(0.0, (0.0, 1000.0), 1e5),
            g5.region(1.0, (1.0, 0.0), 1e5)]
    c = (cfg, g5.init_shock_bubble(cfg, regs, 0.25, (0.5, 0.5), 0.2))
    return [a, b, c]


def oracle_run(orc, cfg, U0, n):
    o = orc.Oracle(cfg)
    o.set_state(U0)
    o.run(n)
    return o.get_state()


def test_pool_mix_matches_alone_and_oracle(g5, orc):
    say("pool: three 2-D solves in one launch")
    solves = mixed_2d_solves(g5)
    nsteps = 12
    with g5.Pool(2, budget_bytes=1 << 28) as pool:
        ids = []
        for cfg, U0 in solves:
            sid = pool.admit(cfg, max_steps=nsteps)
            pool.upload(sid, U0)
            ids.append(sid)
        pool.start(ids)
        pool.step(nsteps + 3)  # extra steps: finished solves must not move
        for sid, (cfg, U0) in zip(ids, solves):
            t, n, active = pool.status(sid)
            assert n == nsteps and not active
            U = pool.download(sid)
            assert np.array_equal(U, oracle_run(orc, cfg, U0, nsteps)), sid
    for cfg, U0 in solves[:1]:
        with g5.Context(cfg) as ctx:
            ctx.upload(U0)
            ctx.run(nsteps)
            assert np.array_equal(ctx.download(), oracle_run(orc, cfg, U0, nsteps))


def test_pool_admission_in_flight_and_reuse(g5, orc):
    say("pool: admission while another solve runs, release and reuse")
    (ca, Ua), (cb, Ub), (cc, Uc) = mixed_2d_solves(g5)
    na, nb_, nc = (c.nb[0] * c.nb[1] for c in (ca, cb, cc))
    with g5.Pool(2, budget_bytes=max(na + nb_, nb_ + nc) * 2 * 2048 * 8) as pool:
        a = pool.admit(ca, max_steps=10)
        pool.upload(a, Ua)
        pool.start([a])
        pool.step(4)
        b = pool.admit(cb, max_steps=6)     # joins while a is in flight
        pool.upload(b, Ub)
        pool.start([b])
        assert pool.try_admit(cc) is None   # budget full
        pool.step(6)
        assert sorted(pool.finished()) == sorted([a, b])
        assert np.array_equal(pool.download(a), oracle_run(orc, ca, Ua, 10))
        assert np.array_equal(pool.download(b), oracle_run(orc, cb, Ub, 6))
        pool.release(a)
        c = pool.admit(cc, max_steps=5)     # reuses a's memory (scattered slots)
        pool.upload(c, Uc)
        pool.start([c])
        pool.step(5)
        assert np.array_equal(pool.download(c), oracle_run(orc, cc, Uc, 5))
        assert np.array_equal(pool.download(b), oracle_run(orc, cb, Ub, 6))  # untouched while finished


def test_pool_3d_mix(g5, orc):
    say("pool: 3-D THINC solves")
    a = cases.air_water_3d(nb=2)
    b = cases.air_water_3d(nb=3)
    with g5.Pool(3, thinc=True, budget_bytes=1 << 30) as pool:
        ids = []
        for cfg, U0 in (a, b):
            sid = pool.admit(cfg, max_steps=4)
            pool.upload(sid, U0)
            ids.append(sid)
        pool.start(ids)
        pool.step(4)
        for sid, (cfg, U0) in zip(ids, (a, b)):
            assert np.array_equal(pool.download(sid), oracle_run(orc, cfg, U0, 4))


def test_pool_3d_inlet_and_plain_solves(g5, orc):
    # a 3-D jet (inflow border) and a plain solve in one launch list: the step kernel with the
    # inflow code runs both; after the jet is released the list has no inlet and the plain kernel
    # continues the other solve -- every result bit-identical to the oracle
    say("pool: 3-D inlet + plain solves, then plain alone")
    a = cases.jet(3, inlet="jet")
    b = cases.air_water_3d(nb=2)
    with g5.Pool(3, thinc=True, budget_bytes=1 << 30) as pool:
        ida = pool.admit(a[0], max_steps=6)
        pool.upload(ida, a[1])
        idb = pool.admit(b[0], max_steps=10)
        pool.upload(idb, b[1])
        pool.start([ida, idb])
        pool.step(6)
        assert np.array_equal(pool.download(ida), oracle_run(orc, a[0], a[1], 6))
        pool.release(ida)
        pool.step(4)
        assert pool.status(idb)[1] == 10
        assert np.array_equal(pool.download(idb), oracle_run(orc, b[0], b[1], 10))


@pytest.mark.parametrize("dim", [2, 3])
def test_pool_miller_puckett_beside_allaire_kokh(g5, orc, dim):
    # one launch list with a MILLER_PUCKETT solve and an ALLAIRE_KOKH solve: the kernel variant
    # with the factors runs both, each block taking its own solve's model; both bit-identical
    say(f"pool: MILLER_PUCKETT beside ALLAIRE_KOKH, dim={dim}")
    if dim == 2:
        a = cases.air_helium_2d(nbx=6, nby=3, source_term="MILLER_PUCKETT")
        b = cases.air_helium_2d(nbx=4, nby=4)
    else:
        a = cases.air_water_3d(nb=2, source_term="MILLER_PUCKETT")
        b = cases.air_water_3d(nb=2)
    with g5.Pool(dim, thinc=bool(a[0].thinc_enabled), budget_bytes=1 << 29) as pool:
        ids = []
        for cfg, U0 in (a, b):
            sid = pool.admit(cfg, max_steps=6)
            pool.upload(sid, U0)
            ids.append(sid)
        pool.start(ids)
        pool.step(6)
        for sid, (cfg, U0) in zip(ids, (a, b)):
            assert np.array_equal(pool.download(sid), oracle_run(orc, cfg, U0, 6))


def test_pool_rk2_ssp_solves(g5, orc):
    # RK2_SSP in a pool with the stage buffer: two solves of different sizes and step budgets
    # (the first finishes while the second goes on), each bit-identical to the oracle
    say("pool: RK2_SSP solves")
    a = cases.air_helium_2d(nbx=6, nby=3, integrator="RK2_SSP")
    b = cases.air_helium_2d(nbx=4, nby=5, integrator="RK2_SSP", slope_type=2.0, thinc=False)
    with g5.Pool(2, budget_bytes=1 << 27, rk2=True) as pool:
        ids = []
        for (cfg, U0), n in ((a, 5), (b, 9)):
            sid = pool.admit(cfg, max_steps=n)
            pool.upload(sid, U0)
            ids.append(sid)
        pool.start(ids)
        pool.step(12)
        for sid, (cfg, U0), n in zip(ids, (a, b), (5, 9)):
            assert pool.status(sid)[1] == n
            assert np.array_equal(pool.download(sid), oracle_run(orc, cfg, U0, n))
    with g5.Pool(2, budget_bytes=1 << 27) as pool:  # without the stage buffer: refused
        with pytest.raises(g5.G5Error) as e:
            pool.admit(a[0])
        assert e.value.code == g5.G5_ERR_UNSUPPORTED


@pytest.mark.parametrize("dim", [2, 3])
def test_device_initial_condition_matches_host(g5, dim):
    say(f"device IC dim={dim}")
    cfg, U0 = cases.air_helium_2d(nbx=4, nby=3) if dim == 2 else cases.air_water_3d(nb=2)
    if dim == 2:
        rho2, u2, p2, _ = g5.rankine_hugoniot(1.4, 0.0, 1.0, 1.0, mach=1.22)
        regs = [g5.region(1.0, (rho2, 0.0), p2, u2), g5.region(1.0, (1.0, 0.0), 1.0),
                g5.region(0.0, (0.0, 0.138), 1.0)]
        args = (0.6, (1.0, 0.5 * 3 * 16 * cfg.dx), 0.25)
    else:
        rho2, u2, p2, _ = g5.rankine_hugoniot(4.4, 6e8, 1000.0, 1e5, p2=1e9)
        regs = [g5.region(0.0, (0.0, rho2), p2, u2), g5.region(0.0, (0.0, 1000.0), 1e5),
                g5.region(1.0, (1.0, 0.0), 1e5)]
        args = (0.25, (0.5, 0.5, 0.5), 0.2)
    host = g5.init_shock_bubble(cfg, regs, *args)
    with g5.Pool(dim, thinc=bool(cfg.thinc_enabled), budget_bytes=1 << 28) as pool:
        sid = pool.admit(cfg)
        pool.init_shock_bubble(sid, regs, args[0], args[1], args[2], shape=0)
        dev = pool.download(sid)
    assert np.array_equal(dev, host)


def test_device_triple_point_and_box(g5):
    say("device IC triple point / box")
    cfg = g5.make_config(2, (7, 3), 1.0 / 48, gamma=(1.5, 1.4))
    regs = [g5.region(1.0, (1.0, 0.0), 1.0), g5.region(1.0, (1.0, 0.0), 0.1), g5.region(0.0, (0.0, 0.125), 0.1)]
    with g5.Pool(2, budget_bytes=1 << 26) as pool:
        sid = pool.admit(cfg)
        pool.init_triple_point(sid, regs, 1.0, 1.5 / 2)
        U = pool.download(sid).reshape(7 * 3, 8, 256)
        x = np.zeros((7 * 3, 256)); y = np.zeros((7 * 3, 256))
        for b in range(21):
            bx, by = b % 7, b // 7
            for c in range(256):
                x[b, c] = (bx * 16 + c % 16 + 0.5) / 48
                y[b, c] = (by * 16 + c // 16 + 0.5) / 48
        a0 = U[:, 4]
        assert np.all(a0[x <= 1.0] == 1.0) and np.all(a0[(x > 1.0) & (y <= 0.75)] == 1.0)
        assert np.all(a0[(x > 1.0) & (y > 0.75)] == 0.0)
        # the host generator of the single-solve path gives the same bits
        assert np.array_equal(pool.download(sid), g5.init_triple_point(cfg, regs, 1.0, 1.5 / 2))
        sid2 = pool.admit(cfg)
        pool.init_shock_bubble(sid2, regs, 0.3, (1.5, 0.5), (0.25, 0.2), shape=1)
        V = pool.download(sid2).reshape(21, 8, 256)
        area = (1 - V[:, 4]).sum() / 48 ** 2
        assert area == pytest.approx(0.5 * 0.4, rel=1e-12)


def test_device_initial_conditions_match_oracle(g5, orc):
    # the device generators against the oracle's restatement of InitShockBubble / InitTriplePoint:
    # bit for bit away from the bubble surface and for the triple point, cut-cell fractions within
    # the sub-cell quadrature's distance from the tangent-plane value
    import math

    import test_oracle_init as T

    say("device IC vs oracle: shock-bubble 2-D/3-D, triple point 2-D/3-D")
    cfg = g5.make_config(2, (16, 8), 1.0 / 128, gamma=(1.4, 1.648))
    regs = T.regions_air_helium(g5)
    cfg3, _ = cases.air_water_3d(nb=2)
    regs3 = T.regions_air_water(g5)
    for c, rg, args, a_pre, a_bub, tol, exact in (
            (cfg, regs, (0.6, (1.0, 0.5), 0.25), 1.0, 0.0, 0.02, math.pi * (0.25 * 128) ** 2),
            (cfg3, regs3, (0.25, (0.5, 0.5, 0.5), 0.2), 0.0, 1.0, 0.04, 4.0 / 3.0 * math.pi * (0.2 * 32) ** 3)):
        with g5.Pool(c.dim, thinc=bool(c.thinc_enabled), budget_bytes=1 << 28) as pool:
            sid = pool.admit(c)
            pool.init_shock_bubble(sid, T.g5_regions(g5, rg), args[0], args[1], args[2], shape=0)
            dev = pool.download(sid)
        T.compare_bubble(dev, orc.init_shock_bubble(c, rg, *args), c, a_pre, a_bub, tol, exact)
    for dim, nb in ((2, (7, 3)), (3, (4, 2, 2))):
        c = g5.make_config(dim, nb, 1.0 / 16 if dim == 3 else 1.0 / 48, gamma=(1.5, 4.4), pinf=(0.0, 6e8))
        rg = [(1.0, (1.0, 0.0), 1.0, 0.5), (0.0, (0.0, 1000.0), 1e5), (1.0, (0.125, 0.0), 0.1)]
        yd = 1.5 if dim == 3 else 0.75
        with g5.Pool(dim, budget_bytes=1 << 27) as pool:
            sid = pool.admit(c)
            pool.init_triple_point(sid, T.g5_regions(g5, rg), 1.0, yd)
            dev = pool.download(sid)
        assert np.array_equal(dev, orc.init_triple_point(c, rg, 1.0, yd)), dim


def test_device_triple_point_3d_matches_host(g5):
    say("device IC triple point 3d vs host")
    cfg = g5.make_config(3, (4, 2, 2), 1.0 / 16, gamma=(1.5, 1.4))
    regs = [g5.region(1.0, (1.0, 0.0), 1.0), g5.region(0.0, (0.0, 1.0), 0.1), g5.region(1.0, (0.125, 0.0), 0.1)]
    with g5.Pool(3, budget_bytes=1 << 27) as pool:
        sid = pool.admit(cfg)
        pool.init_triple_point(sid, regs, 1.0, 1.5)
        dev = pool.download(sid)
    assert np.array_equal(dev, g5.init_triple_point(cfg, regs, 1.0, 1.5))


def test_launch_geometries_bit_identical(g5):
    # every launch geometry the autotuner may pick computes the same bits; the 2-D persistent
    # geometries loop over several blocks per workgroup here (1920 blocks, 256..1024 workgroups);
    # the 3-D z-columns of 2, 4 and 8 blocks march through 3 blocks in z (padded columns)
    say("launch geometries: 2-D FP64/mixed, 3-D FP64/mixed/reassociated")
    for dim, prec in ((2, g5.PREC_FP64), (2, g5.PREC_MIXED), (3, g5.PREC_FP64), (3, g5.PREC_MIXED),
                      (3, g5.PREC_FP64_FAST)):
        if dim == 2:
            cfg, U0 = cases.air_helium_2d(nbx=48, nby=40, thinc=True, thinc_beta=1.6, precision=prec)
            geoms = [(0, 1, 1, 0, 1), (0, 0, 1, 0, 1), (1, 1, 1, 0, 1), (2, 1, 1, 0, 1), (4, 0, 1, 0, 1),
                     (0, 1, 1, 1, 1), (0, 0, 1, 1, 1)]
        else:
            cfg, U0 = cases.air_water_3d(nb=3, precision=prec)
            geoms = [(0, 1, 1, 0, 1), (0, 0, 1, 0, 1), (0, 1, 0, 0, 1), (0, 0, 0, 0, 1), (0, 1, 1, 0, 2),
                     (0, 1, 1, 0, 4), (0, 0, 0, 0, 4), (0, 1, 0, 0, 2), (0, 1, 1, 0, 8)]
        ref = None
        for wg, remap, zs, rw, cl in geoms:
            say(f"dim={dim} prec={prec} launch wg_per_cu={wg} remap={remap} zsplit={zs} ring_wave={rw} col_len={cl}")
            with g5.Pool(dim, thinc=True, precision=prec, budget_bytes=1 << 28) as pool:
                pool.set_launch(wg, remap, zs, rw, cl)
                sid = pool.admit(cfg, max_steps=6)
                pool.upload(sid, U0)
                pool.start([sid])
                pool.step(6)
                U = pool.download(sid)
                assert pool.launch()["wg_per_cu"] == wg and pool.launch()["col_len"] == cl
            if ref is None:
                ref = U
            else:
                assert np.array_equal(U, ref), (dim, prec, wg, remap, zs, rw, cl)
        # the autotuned pool gives the same bits as well
        with g5.Pool(dim, thinc=True, precision=prec, budget_bytes=1 << 28) as pool:
            sid = pool.admit(cfg, max_steps=6)
            pool.upload(sid, U0)
            pool.start([sid])
            pool.step(6)
            assert np.array_equal(pool.download(sid), ref)
            say(f"dim={dim} prec={prec} autotuned launch {pool.launch()}")


def test_served_3d_request_bits_do_not_depend_on_the_launch(g5):
    # a 3-D request of the seeded serving set (MUSCL, downsized to 4^3 blocks) in mixed precision
    # and in the reassociated FP64 form: the split and the in-place z solve and every z-column
    # length give the same bits.  The mixed kernels once rounded the two z-solve placements
    # differently for this request, when the compiler contracted a*b+c across statements (each
    # inlined copy of the z solve then fused according to the code around it); contraction within
    # an expression rounds every copy alike.
    spec = g5.serve_specs(20251203, 32, 3000.0, 0.25)[11]
    assert spec.dim == 3
    spec.nb[0] = spec.nb[1] = spec.nb[2] = 4
    spec.steps = 9
    for prec in (g5.PREC_MIXED, g5.PREC_FP64_FAST):
        ref = None
        for zs in (0, 1):
            for cl in (1, 2, 4, 8):
                say(f"served 3-D request, prec={prec} zsplit={zs} col_len={cl}")
                with g5.Pool(3, precision=prec, budget_bytes=1 << 28) as pool:
                    pool.set_launch(0, 1, zs, 0, cl)
                    sid = pool.admit_spec(spec)
                    pool.start([sid])
                    pool.step(spec.steps + 3)
                    t, n, _ = pool.status(sid)
                    U = pool.download(sid)
                assert n == spec.steps and np.all(np.isfinite(U))
                if ref is None:
                    ref = U
                else:
                    assert np.array_equal(U, ref), (prec, zs, cl)


def test_z_columns_in_a_pool_of_uneven_solves(g5):
    # three 3-D solves of 3, 5 and 2 blocks in z in one pool, marched in z-columns of 2, 4, 8 blocks
    # (padded columns, columns of different solves side by side in the list): each ends with the
    # bits of its run alone with one block per workgroup
    say("z-columns: pool of 3-D solves with 3, 5 and 2 blocks in z")
    shapes = [(3, 2, 3), (2, 2, 5), (2, 3, 2)]
    alone = []
    runs = []
    for nb in shapes:
        cfg, U0 = cases.air_water_3d(nb=max(nb))
        cfg = g5.make_config(3, nb, cfg.dx, bc=("ZERO_GRADIENT",) * 6, cfl=0.5, gamma=(1.4, 4.4), pinf=(0.0, 6e8),
                             thinc=True, thinc_beta=1.6, thinc_epsilon=1e-3, precision=g5.PREC_FP64_FAST)
        rho2, u2, p2, _ = g5.rankine_hugoniot(4.4, 6e8, 1000.0, 1e5, p2=1e9)
        regs = [g5.region(0.0, (0.0, rho2), p2, u2), g5.region(0.0, (0.0, 1000.0), 1e5), g5.region(1.0, (1.0, 0.0), 1e5)]
        centre = tuple(0.5 * 16 * n * cfg.dx for n in nb)
        U0 = g5.init_shock_bubble(cfg, regs, 0.25 * 16 * nb[0] * cfg.dx, centre, 0.3 * 16 * min(nb) * cfg.dx)
        runs.append((cfg, U0))
        with g5.Pool(3, thinc=True, precision=g5.PREC_FP64_FAST, budget_bytes=1 << 28) as pool:
            pool.set_launch(0, 1, 1, 0, 1)
            sid = pool.admit(cfg, max_steps=5)
            pool.upload(sid, U0)
            pool.start([sid])
            pool.step(5)
            alone.append(pool.download(sid))
    for cl in (2, 4, 8):
        with g5.Pool(3, thinc=True, precision=g5.PREC_FP64_FAST, budget_bytes=1 << 29) as pool:
            pool.set_launch(0, 1, 1, 0, cl)
            sids = []
            for cfg, U0 in runs:
                sids.append(pool.admit(cfg, max_steps=5))
                pool.upload(sids[-1], U0)
            pool.start(sids)
            pool.step(5)
            for k, sid in enumerate(sids):
                assert np.array_equal(pool.download(sid), alone[k]), (cl, k)


def test_z_columns_on_every_step_path(g5, orc):
    # z-columns of 8 blocks against one block per workgroup on the paths the step kernels branch on:
    # RK2_SSP (stage buffer, two launches per step), a jet inflow border (the INLET instances, with
    # the per-plane border resolution of inflow blocks), the MILLER_PUCKETT source term, walls on
    # every face and a periodic z direction (column ends wrap); bit-exact FP64 against the oracle
    say("z-columns: RK2_SSP, jet inlet, MILLER_PUCKETT, periodic z")
    runs = [("rk2", cases.air_water_3d(nb=2, integrator="RK2_SSP"), True),
            ("jet", cases.jet(3, nb=(3, 2, 5)), False),
            ("mp", cases.air_water_3d(nb=3, source_term="MILLER_PUCKETT"), False),
            ("walls-periodic-z", cases.air_helium_3d(nb=3, periodic=(0, 0, 1),
                                                     bc=("WALL",) * 4 + ("PERIODIC", "PERIODIC")), False)]
    for label, (cfg, U0), rk2 in runs:
        ref = oracle_run(orc, cfg, U0, 5)
        for cl in (1, 2, 8):
            with g5.Pool(3, thinc=bool(cfg.thinc_enabled), budget_bytes=1 << 28, rk2=rk2) as pool:
                pool.set_launch(0, 1, 1, 0, cl)
                sid = pool.admit(cfg, max_steps=5)
                pool.upload(sid, U0)
                pool.start([sid])
                pool.step(5)
                assert np.array_equal(pool.download(sid), ref), (label, cl)


def test_3d_persistent_launch_is_rejected(g5):
    with g5.Pool(3, thinc=True, budget_bytes=1 << 26) as pool:
        with pytest.raises(g5.G5Error):
            pool.set_launch(1, 1, 1)
        with pytest.raises(g5.G5Error):  # the ring wavefront is a 2-D form
            pool.set_launch(0, 1, 1, 1)
        with pytest.raises(g5.G5Error):  # z-columns of 1, 2, 4 or 8 blocks
            pool.set_launch(0, 1, 1, 0, 3)
    with g5.Pool(2, thinc=True, budget_bytes=1 << 26) as pool:
        with pytest.raises(g5.G5Error):  # ... with one workgroup per block
            pool.set_launch(2, 1, 1, 1)
        with pytest.raises(g5.G5Error):  # z-columns are a 3-D form
            pool.set_launch(0, 1, 1, 0, 2)


def test_checkpoint_restart_is_bit_identical(g5, tmp_path):
    # a solve saved after 7 steps, released, and restored into another pool (different slots)
    # finishes its 19-step budget with the bits of the uninterrupted run, FP64 and mixed
    from godunov_five_eq_amd import checkpoint

    for prec in (g5.PREC_FP64, g5.PREC_MIXED):
        say(f"checkpoint/restart precision={prec}")
        cfg, U0 = cases.air_helium_2d(nbx=7, nby=5, thinc=True, thinc_beta=1.6, precision=prec)
        with g5.Pool(2, thinc=True, precision=prec, budget_bytes=1 << 26) as pool:
            sid = pool.admit(cfg, max_steps=19)
            pool.upload(sid, U0)
            pool.start([sid])
            pool.step(25)
            ref = pool.download(sid)
            t_ref, n_ref, _ = pool.status(sid)
        prefix = str(tmp_path / f"solve{prec}")
        with g5.Pool(2, thinc=True, precision=prec, budget_bytes=1 << 26) as pool:
            other = pool.admit(cfg)  # occupies the first slots, so the solve lands elsewhere
            pool.upload(other, U0)
            sid = pool.admit(cfg, max_steps=19)
            pool.upload(sid, U0)
            pool.start([sid])
            pool.step(7)
            checkpoint.save(prefix, pool, sid)
        with g5.Pool(2, thinc=True, precision=prec, budget_bytes=1 << 26) as pool:
            sid = checkpoint.load(prefix, pool)
            pool.step(25)
            t, n, active = pool.status(sid)
            assert n == n_ref == 19 and t == t_ref and not active
            assert np.array_equal(pool.download(sid), ref)


def test_hdf5_checkpoint_restart_is_bit_identical(g5, tmp_path):
    # the HDF5 form: a 3-D THINC solve saved after 5 steps (variables by name, time and step count
    # in the file) and restored into a fresh pool finishes its 12-step budget with the bits of the
    # uninterrupted run
    from godunov_five_eq_amd import checkpoint

    say("HDF5 checkpoint/restart 3d")
    cfg, U0 = cases.air_water_3d(nb=3)
    with g5.Pool(3, thinc=True, budget_bytes=1 << 28) as pool:
        sid = pool.admit(cfg, max_steps=12)
        pool.upload(sid, U0)
        pool.start([sid])
        pool.step(15)
        ref = pool.download(sid)
        t_ref, n_ref, _ = pool.status(sid)
    with g5.Pool(3, thinc=True, budget_bytes=1 << 28) as pool:
        sid = pool.admit(cfg, max_steps=12)
        pool.upload(sid, U0)
        pool.start([sid])
        pool.step(5)
        path = checkpoint.save_hdf5(str(tmp_path), "solve_0000005", pool, sid)
    with g5.Pool(3, thinc=True, budget_bytes=1 << 28) as pool:
        sid = checkpoint.load_hdf5(path, pool, cfg, 1e300, 12)
        assert pool.status(sid)[1] == 5
        pool.step(15)
        t, n, active = pool.status(sid)
        assert n == n_ref == 12 and t == t_ref and not active
        assert np.array_equal(pool.download(sid), ref)


def test_pool_error_paths(g5):
    say("pool error paths")
    cfg2, U0 = cases.air_helium_2d(nbx=4, nby=2)
    cfg3, _ = cases.air_helium_3d(nb=2)
    with g5.Pool(2, budget_bytes=1 << 22) as pool:  # 128 slots of 2-D blocks
        with pytest.raises(g5.G5Error) as e:
            pool.admit(cfg3)
        assert e.value.code == g5.G5_ERR_INVALID
        rk2, _ = cases.air_helium_2d(nbx=4, nby=2, integrator="RK2_SSP")
        with pytest.raises(g5.G5Error) as e:
            pool.admit(rk2)
        assert e.value.code == g5.G5_ERR_UNSUPPORTED
        # one integrator per pool: the step graph applies it to the whole launch list
        first = pool.admit(cfg2)
        rk1, _ = cases.air_helium_2d(nbx=4, nby=2, integrator="RK1")
        with pytest.raises(g5.G5Error) as e:
            pool.admit(rk1)
        assert e.value.code == g5.G5_ERR_INVALID and "integrator" in str(e.value)
        pool.release(first)
        pool.release(pool.admit(rk1))  # accepted once the pool is empty
        big, _ = cases.air_helium_2d(nbx=16, nby=16)  # 256 blocks > 128 slots: can never fit
        with pytest.raises(g5.G5Error) as e:
            pool.try_admit(big)
        assert e.value.code == g5.G5_ERR_TOO_LARGE
        sid = pool.admit(cfg2, max_steps=3)  # the pool still works after the refusals
        pool.upload(sid, U0)
        pool.start([sid])
        with pytest.raises(g5.G5Error) as e:
            pool.set_clock(sid, 0.5, 2)
        assert e.value.code == g5.G5_ERR_STATE
        pool.step(5)
        assert pool.status(sid)[1] == 3
        pool.release(sid)
        with pytest.raises(g5.G5Error):
            pool.release(sid)
        with pytest.raises(g5.G5Error):
            pool.status(sid)
        assert pool.usage()[0] == 0
