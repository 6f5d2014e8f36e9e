"""Workload definitions shared by the tests, smoke() and bench.py.

shock_bubble_2d_air_helium: 2-D air-helium shock-bubble of app/test_shock_bubble_air_helium_2d.ini
  at 512x128 cells (domain 4x1, 16x16 blocks), air gamma=1.4, helium gamma=1.648, bubble radius
  0.25 at (1.0, 0.5), density ratio 0.138, Mach 1.22 shock, minmod, HLLC, CFL 0.8, HANCOCK,
  ZERO_GRADIENT in x, WALL in y.
shock_bubble_3d_air_water: 3-D air bubble (radius 0.2) in water (gamma 4.4, pinf 6e8 Pa,
  rho 1000) hit by a 1e9 Pa planar shock, THINC beta 1.6, HLLC, CFL 0.5, HANCOCK.  Unit cube,
  bubble centre (0.5, 0.5, 0.5), shock front at x = 0.25, pre-shock pressure 1e5 Pa, air rho 1.
"""
import godunov_five_eq_amd as g5


def air_helium_2d(nbx=32, nby=8, **kw):
    dx = 1.0 / (16 * nby)
    rho2, u2, p2, _ = g5.rankine_hugoniot(1.4, 0.0, 1.0, 1.0, mach=1.22)
    opts = dict(bc=("ZERO_GRADIENT", "ZERO_GRADIENT", "WALL", "WALL", "NONE", "NONE"), cfl=0.8,
                slope_type=1.0, gamma=(1.4, 1.648), pinf=(0.0, 0.0))
    opts.update(kw)
    cfg = g5.make_config(2, (nbx, nby), dx, **opts)
    regs = [g5.region(1.0, (rho2, 0.0), p2, u2), g5.region(1.0, (1.0, 0.0), 1.0),
            g5.region(0.0, (0.0, 0.138), 1.0)]
    # bubble of radius 0.25 at x=1, mid height; shock front at x=0.6
    yc = 0.5 * nby * 16 * dx
    U0 = g5.init_shock_bubble(cfg, regs, 0.6, (1.0, yc), 0.25)
    return cfg, U0


def air_water_3d(nb=16, thinc=True, **kw):
    dx = 1.0 / (16 * nb)
    pinf_w = 6e8
    rho2, u2, p2, _ = g5.rankine_hugoniot(4.4, pinf_w, 1000.0, 1e5, p2=1e9)
    opts = dict(bc=("ZERO_GRADIENT",) * 6, cfl=0.5, slope_type=1.0, gamma=(1.4, 4.4),
                pinf=(0.0, pinf_w), thinc=thinc, thinc_beta=1.6, thinc_epsilon=1e-3)
    opts.update(kw)
    cfg = g5.make_config(3, (nb, nb, nb), dx, **opts)
    # material 0 = air, material 1 = water
    regs = [g5.region(0.0, (0.0, rho2), p2, u2), g5.region(0.0, (0.0, 1000.0), 1e5),
            g5.region(1.0, (1.0, 0.0), 1e5)]
    U0 = g5.init_shock_bubble(cfg, regs, 0.25, (0.5, 0.5, 0.5), 0.2)
    return cfg, U0


def air_helium_3d(nb=4, **kw):
    """Small 3-D air-helium shock-bubble (MUSCL) for parity tests."""
    dx = 1.0 / (16 * nb)
    rho2, u2, p2, _ = g5.rankine_hugoniot(1.4, 0.0, 1.0, 1.0, mach=1.22)
    opts = dict(bc=("ZERO_GRADIENT", "ZERO_GRADIENT", "WALL", "WALL", "WALL", "WALL"), cfl=0.8,
                slope_type=1.0, gamma=(1.4, 1.648), pinf=(0.0, 0.0))
    opts.update(kw)
    cfg = g5.make_config(3, (nb, nb, nb), dx, **opts)
    regs = [g5.region(1.0, (rho2, 0.0), p2, u2), g5.region(1.0, (1.0, 0.0), 1.0),
            g5.region(0.0, (0.0, 0.138), 1.0)]
    U0 = g5.init_shock_bubble(cfg, regs, 0.2, (0.5, 0.5, 0.5), 0.25)
    return cfg, U0


def interface_advection_2d(alpha_min=0.1, alpha_max=0.9, pinf=0.0, rho0=1000.0, rho1=1000.0,
                           velocity=(100.0, 50.0), p0=1e5, nb=(4, 4), **kw):
    """Periodic 2-D box with uniform pressure p0 and velocity, and a smooth material blob
    (alpha0 from alpha_min outside to alpha_max inside a circle of radius 0.25, tanh width 0.03):
    air (gamma 1.4) and a stiffened gas (gamma 4.4, pinf) with phase densities rho0 and rho1.
    Returns (cfg, U0, pressure(U) -> per-cell pressure of the mixture EOS)."""
    import numpy as np

    dx = 1.0 / (16 * nb[1])
    opts = dict(periodic=(1, 1, 0), bc=("PERIODIC",) * 6, cfl=0.5, gamma=(1.4, 4.4), pinf=(0.0, pinf),
                thinc_beta=1.6, thinc_epsilon=1e-3)
    opts.update(kw)
    cfg = g5.make_config(2, nb, dx, **opts)
    g1 = np.array([1 / 0.4, 1 / 3.4])
    gp = np.array([0.0, 4.4 * pinf / 3.4])
    ux, uy = velocity
    nblk = nb[0] * nb[1]
    i = np.arange(16)
    X = (np.arange(nb[0])[:, None] * 16 + i[None, :] + 0.5) * dx  # [bx, i]
    Y = (np.arange(nb[1])[:, None] * 16 + i[None, :] + 0.5) * dx  # [by, j]
    U = np.zeros((nb[1], nb[0], 8, 16, 16))  # [by, bx, var, j, i]
    x = X[None, :, None, :]
    y = Y[:, None, :, None]
    r = np.hypot(x - 0.5 * nb[0] * 16 * dx, y - 0.5)
    a0 = alpha_min + (alpha_max - alpha_min) * 0.5 * (1 + np.tanh((0.25 - r) / 0.03))
    ar0, ar1 = a0 * rho0, (1 - a0) * rho1
    rho = ar0 + ar1
    xi = a0 * g1[0] + (1 - a0) * g1[1]
    Pi = a0 * gp[0] + (1 - a0) * gp[1]
    E = p0 * xi + Pi + 0.5 * rho * (ux * ux + uy * uy)
    for v, f in enumerate([rho, E, rho * ux, rho * uy, a0, ar0, 1 - a0, ar1]):
        U[:, :, v] = f
    U0 = U.reshape(nblk, 8, 256).ravel()

    def pressure(S):
        V = S.reshape(-1, 8, 256)
        d, a = V[:, 0], V[:, 4]
        vx, vy = V[:, 2] / d, V[:, 3] / d
        xi_ = a * g1[0] + (1 - a) * g1[1]
        Pi_ = a * gp[0] + (1 - a) * gp[1]
        return (V[:, 1] - 0.5 * d * (vx * vx + vy * vy) - Pi_) / xi_

    return cfg, U0, pressure


def jet(dim=2, nb=None, inlet="jet", jet_geom=(0.5, 0.5, 0.25), **kw):
    """Fuel jet into nitrogen after app/test_jet_2d_fuel_N2.ini (gammas, p_inf, region states),
    on a fixed single-level mesh with WALL borders: region 0 (nitrogen, u = 2) everywhere, and
    the outside cells with x < xmin take region 1 (fuel, u = 18) inside the jet (inlet "jet")
    or all of them take region 0 (inlet "shock_bubble").  Returns (cfg, U0)."""
    import numpy as np

    nb = nb or ((6, 2) if dim == 2 else (3, 2, 2))
    dx = 1.0 / (16 * nb[1])
    r0 = g5.region(1.0, (32.2, 0.0), 3.83982e5, 2.0)
    r1 = g5.region(0.0, (0.0, 623.0), 3.83982e5, 18.0)
    opts = dict(bc=("WALL",) * 6, cfl=0.8, slope_type=2.0, gamma=(1.4, 4.4), pinf=(0.0, 8e7),
                thinc=True, thinc_beta=1.5, thinc_epsilon=1e-3)
    opts.update(kw)
    base = g5.make_config(dim, nb, dx, **opts)
    cfg = g5.make_config(dim, nb, dx, inlet=(inlet, r1 if inlet == "jet" else r0), jet=jet_geom, **opts)
    u0 = g5.region_conservative(base, r0)
    ncell = 16 ** dim
    nblocks = int(np.prod(nb))
    U0 = np.repeat(u0[None, :, None], nblocks, axis=0).repeat(ncell, axis=2).ravel()
    return cfg, U0


def random_case(seed):
    """A seeded small configuration for randomized parity: 2-D or 3-D, block counts from 1,
    per-face ZERO_GRADIENT / WALL or periodic directions, THINC or MUSCL, minmod or MC, HANCOCK /
    RK1 / RK2_SSP, ideal or stiffened gases, a shock-bubble initial state with transverse
    velocity, sometimes an inflow border and sometimes the MILLER_PUCKETT source term.
    Returns (cfg, U0, steps, label)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    dim = 3 if rng.random() < 0.4 else 2
    nb = [int(rng.integers(1, 6 if dim == 2 else 4)) for _ in range(dim)]
    periodic = [int(rng.random() < 0.3) for _ in range(dim)] + [0] * (3 - dim)
    bc = []
    for d in range(3):
        for _ in range(2):
            if d >= dim:
                bc.append("NONE")
            elif periodic[d]:
                bc.append("PERIODIC")
            else:
                bc.append("WALL" if rng.random() < 0.5 else "ZERO_GRADIENT")
    thinc = bool(rng.random() < 0.5)
    slope = 2.0 if rng.random() < 0.4 else 1.0
    integrator = str(rng.choice(["HANCOCK", "HANCOCK", "RK1", "RK2_SSP"]))
    water = bool(rng.random() < 0.4)
    dx = 1.0 / (16 * max(nb))
    ext = [16 * n * dx for n in nb]
    if water:
        gamma, pinf = (float(rng.uniform(1.2, 1.7)), 4.4), (0.0, 6e8)
        rho2, u2, p2, _ = g5.rankine_hugoniot(4.4, 6e8, 1000.0, 1e5, p2=float(rng.uniform(2e8, 1e9)))
        regs = [g5.region(0.0, (0.0, rho2), p2, u2, float(rng.uniform(-20, 20)), float(rng.uniform(-20, 20))),
                g5.region(0.0, (0.0, 1000.0), 1e5), g5.region(1.0, (1.0, 0.0), 1e5)]
    else:
        gamma, pinf = (1.4, float(rng.uniform(1.1, 1.9))), (0.0, 0.0)
        rho2, u2, p2, _ = g5.rankine_hugoniot(1.4, 0.0, 1.0, 1.0, mach=float(rng.uniform(1.1, 3.0)))
        regs = [g5.region(1.0, (rho2, 0.0), p2, u2, float(rng.uniform(-0.3, 0.3)), float(rng.uniform(-0.3, 0.3))),
                g5.region(1.0, (1.0, 0.0), 1.0), g5.region(0.0, (0.0, float(rng.uniform(0.1, 3.0))), 1.0)]
    inlet = None
    if not periodic[0] and rng.random() < 0.25:
        inlet = (str(rng.choice(["jet", "shock_bubble"])), regs[0])
    cfg = g5.make_config(dim, nb, dx, periodic=periodic, bc=tuple(bc), integrator=integrator, thinc=thinc,
                         cfl=float(rng.uniform(0.3, 0.8)), slope_type=slope, gamma=gamma, pinf=pinf,
                         thinc_beta=float(rng.uniform(1.2, 2.0)), inlet=inlet,
                         jet=(0.5 * ext[1], 0.5 * ext[2] if dim == 3 else 0.5, 0.4 * ext[1]))
    centre = [float(rng.uniform(0.3, 0.7)) * e for e in ext]
    radius = float(rng.uniform(0.15, 0.35)) * min(ext)
    x_front = float(rng.uniform(0.05, 0.35)) * ext[0]
    U0 = g5.init_shock_bubble(cfg, regs, x_front, centre, radius)
    steps = int(rng.integers(5, 13)) if dim == 2 else int(rng.integers(3, 7))
    mp = bool(rng.random() < 0.25)  # drawn last, so the draws above do not depend on it
    if mp:
        cfg.source_term = g5.SOURCE_MILLER_PUCKETT
    label = (f"seed {seed}: {dim}d nb={nb} periodic={periodic[:dim]} bc={bc[:2 * dim]} thinc={thinc} "
             f"slope={slope} {integrator} water={water} inlet={inlet[0] if inlet else None} "
             f"source={'MILLER_PUCKETT' if mp else 'ALLAIRE_KOKH'} steps={steps}")
    return cfg, U0, steps, label


def haas_sturtevant(level=5, **kw):
    """The air-helium shock-bubble of app/test_shock_bubble_air_helium_2d/
    test_shock_bubble_air_helium_2d_measure.ini on a fixed mesh: domain 445 x 89 mm (brick 5x1,
    scaling 89), 16x16 blocks at `level` (dx = 89 / (16 * 2^level) mm), Mach 1.22 shock at
    x = 170 mm, helium-air bubble of radius 25 mm at (220, 44.5), MC slopes, HLLC, THINC
    beta 1.5, CFL 0.5, zero-gradient x borders and walls in y.  Units mm, ms, g and kPa-consistent
    (velocities in mm/ms = m/s).  Returns (cfg, U0)."""
    n = 2 ** level
    dx = 89.0 / (16 * n)
    opts = dict(bc=("ZERO_GRADIENT", "ZERO_GRADIENT", "WALL", "WALL", "NONE", "NONE"), cfl=0.5, slope_type=2.0,
                gamma=(1.4, 1.648), pinf=(0.0, 0.0), thinc=True, thinc_beta=1.5, thinc_epsilon=1e-3)
    opts.update(kw)
    cfg = g5.make_config(2, (5 * n, n), dx, **opts)
    regs = [g5.region(1.0, (1.6571422230442021, 0.0), 159059.985, 114.51006445833735),
            g5.region(1.0, (1.1839, 0.0), 101325.0),
            g5.region(0.0, (0.0, 0.2193078324225865), 101325.0)]
    U0 = g5.init_shock_bubble(cfg, regs, 170.0, (220.0, 44.5), 25.0)
    return cfg, U0


def wave_fronts(cfg, U):
    """Positions (mm) of the five fronts of the reference's plot.py on a 2-D state of the
    haas_sturtevant case: along the centreline y = 44.5 the upstream ("jet") and downstream
    interfaces (helium fraction 1/2), the refracted shock inside the bubble and the transmitted
    shock ahead of it (rightmost pressure rise of 5% above ambient on either side of the
    downstream interface); along the bottom wall row the incident shock (same criterion)."""
    import numpy as np

    nbx, nby = cfg.nb[0], cfg.nb[1]
    A = U.reshape(nby, nbx, 8, 16, 16).transpose(2, 0, 3, 1, 4).reshape(8, nby * 16, nbx * 16)  # var, y, x
    ny, nx = A.shape[1], A.shape[2]
    x = (np.arange(nx) + 0.5) * cfg.dx
    rho, E, mu, mv, a0 = A[0], A[1], A[2], A[3], A[4]
    g1 = a0 / 0.4 + (1 - a0) / 0.648
    p = (E - 0.5 * (mu * mu + mv * mv) / rho) / g1
    jc = ny // 2  # the centreline lies between rows jc-1 and jc
    he = 1.0 - 0.5 * (a0[jc - 1] + a0[jc])
    pc = 0.5 * (p[jc - 1] + p[jc])

    def crossing(f, level, i0, i1, rightmost):
        idx = np.nonzero(f[i0:i1] > level)[0]
        if idx.size == 0:
            return float("nan")
        k = i0 + (idx[-1] if rightmost else idx[0])
        # linear interpolation to the level between cells k and its outer neighbour
        kn = k + 1 if rightmost else k - 1
        if kn < 0 or kn >= nx or f[k] == f[kn]:
            return float(x[k])
        return float(x[k] + (x[kn] - x[k]) * (f[k] - level) / (f[k] - f[kn]))

    inside = np.nonzero(he > 0.5)[0]
    jet = crossing(he, 0.5, 0, nx, False)
    down = crossing(he, 0.5, 0, nx, True)
    idown = int(inside[-1]) if inside.size else nx - 1
    ijet = int(inside[0]) if inside.size else 0
    pref = 101325.0 * 1.05
    refracted = crossing(pc, pref, ijet, idown + 1, True)
    transmitted = crossing(pc, pref, idown + 1, nx, True)
    shock = crossing(p[0], pref, 0, nx, True)
    return {"jet": jet, "downstream_interface": down, "refracted": refracted, "transmitted": transmitted,
            "shock": shock}
