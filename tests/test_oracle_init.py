"""Initial conditions: the oracle's restatement (region states, shock-bubble with the tangent-plane
volume fraction, triple point) pinned by closed forms and an independent exact clipping, and the
product's host generators checked against it (CPU only; the device generators are checked against
the oracle in tests/test_pool_gpu.py).

  - region states (common.h:68-144 get_region_init_state): E = p xi + Pi + rho |u|^2 / 2 with the
    textbook pure-phase limits p/(gamma-1) and (p + gamma pinf)/(gamma-1);
  - the cell volume fraction below a plane (the restated vof::compute_volume_fraction_of_rect_below_plane,
    absent from the reference tree): axis-aligned planes, planes through the centre (1/2), the
    complement rule, and an independent exact value (2-D: the square clipped by the half-plane,
    shoelace area; 3-D: those areas integrated over z slices);
  - the tangent-plane fraction of a sphere (InitShockBubble.cpp:142-150): the half-space below
    the tangent plane contains the sphere, so the summed volume bounds pi R^2 (4/3 pi R^3) from
    above, within 1e-3 at R = 32 dx;
  - shock-bubble and triple-point states (InitShockBubble.cpp:95-209, InitTriplePoint.cpp:66-128):
    pure cells carry the region states; the product's generators (sub-cell quadrature for the
    bubble fraction) equal the oracle's bit for bit away from the bubble surface and the volume
    fraction of its cut cells is within the quadrature error of the tangent-plane value.
"""
import math

import numpy as np
import pytest

import cases


def regions_air_helium(g5):
    rho2, u2, p2, _ = g5.rankine_hugoniot(1.4, 0.0, 1.0, 1.0, mach=1.22)
    return [(1.0, (rho2, 0.0), p2, u2), (1.0, (1.0, 0.0), 1.0), (0.0, (0.0, 0.138), 1.0)]


def regions_air_water(g5):
    rho2, u2, p2, _ = g5.rankine_hugoniot(4.4, 6e8, 1000.0, 1e5, p2=1e9)
    return [(0.0, (0.0, rho2), p2, u2), (0.0, (0.0, 1000.0), 1e5), (1.0, (1.0, 0.0), 1e5)]


def g5_regions(g5, regs):
    return [g5.region(r[0], r[1], r[2], *r[3:]) for r in regs]


def test_region_state_closed_forms(orc, g5):
    cfg = g5.make_config(3, (1, 1, 1), 0.1, gamma=(1.4, 4.4), pinf=(0.0, 6e8))
    # pure ideal gas (material 0)
    u = orc.region_state(cfg, (1.0, (1.2, 0.0), 2.5e5, 3.0, -4.0, 12.0))
    assert u[0] == 1.2 and u[5] == 1.0 and u[6] == 1.2 and u[7] == 0.0 and u[8] == 0.0
    assert u[2] == 3.0 * 1.2 and u[3] == -4.0 * 1.2 and u[4] == 12.0 * 1.2
    assert u[1] == pytest.approx(2.5e5 / 0.4 + 0.5 * 1.2 * (9 + 16 + 144), rel=1e-15)
    # pure stiffened gas (material 1)
    u = orc.region_state(cfg, (0.0, (0.0, 1000.0), 1e9, 50.0))
    assert u[1] == pytest.approx((1e9 + 4.4 * 6e8) / 3.4 + 0.5 * 1000.0 * 2500.0, rel=1e-14)
    assert u[5] == 0.0 and u[7] == 1.0
    # mixture: p xi + Pi, xi = sum alpha_k/(gamma_k - 1), Pi = sum alpha_k gamma_k pinf_k/(gamma_k - 1)
    a = 0.3
    u = orc.region_state(cfg, (a, (0.3, 700.0), 1e5, 0.0, 2.0))
    xi = a / 0.4 + (1 - a) / 3.4
    pi = (1 - a) * 4.4 * 6e8 / 3.4
    assert u[1] == pytest.approx(1e5 * xi + pi + 0.5 * 700.3 * 4.0, rel=1e-14)
    # the product's region state is the same restatement
    assert np.array_equal(g5.region_conservative(cfg, g5.region(a, (0.3, 700.0), 1e5, 0.0, 2.0)), u)


def clipped_square_area(n, alpha, lo, dx):
    """Exact area of the square [lo, lo + dx]^2 where n.x <= alpha: Sutherland-Hodgman clipping
    of the square by the half-plane, then the shoelace formula."""
    poly = [(lo[0], lo[1]), (lo[0] + dx, lo[1]), (lo[0] + dx, lo[1] + dx), (lo[0], lo[1] + dx)]
    f = [n[0] * x + n[1] * y - alpha for x, y in poly]
    out = []
    for k in range(4):
        (x0, y0), (x1, y1), f0, f1 = poly[k], poly[(k + 1) % 4], f[k], f[(k + 1) % 4]
        if f0 <= 0:
            out.append((x0, y0))
        if (f0 < 0) != (f1 < 0) and f0 != f1:
            t = f0 / (f0 - f1)
            out.append((x0 + t * (x1 - x0), y0 + t * (y1 - y0)))
    if len(out) < 3:
        return 0.0
    return 0.5 * abs(sum(out[k][0] * out[(k + 1) % len(out)][1] - out[(k + 1) % len(out)][0] * out[k][1]
                         for k in range(len(out))))


def clipped_cube_volume(n, alpha, lo, dx, nslice=4000):
    """Volume of the cube [lo, lo + dx]^3 where n.x <= alpha: the exact clipped area of each z slice
    (midpoint rule over nslice slices; the area is piecewise quadratic in z)."""
    h = dx / nslice
    v = 0.0
    for s in range(nslice):
        z = lo[2] + (s + 0.5) * h
        v += clipped_square_area(n[:2], alpha - n[2] * z, lo[:2], dx) * h
    return v


def test_box_fraction_below_plane(orc):
    rng = np.random.default_rng(7)
    lo2, lo3, dx = np.array([0.3, -1.1]), np.array([0.3, -1.1, 2.0]), 0.25
    # axis-aligned planes: the fraction is the plane's offset across the cell, clamped
    for t in (-0.5, 0.0, 0.125, 0.5, 0.875, 1.0, 1.5):
        assert orc.box_fraction_below_plane([1.0, 0.0], lo2[0] + t * dx, lo2, dx) == pytest.approx(min(max(t, 0), 1), abs=1e-15)
        assert orc.box_fraction_below_plane([0.0, 0.0, -1.0], -(lo3[2] + t * dx), lo3, dx) == \
            pytest.approx(1.0 - min(max(t, 0), 1), abs=1e-15)
    for _ in range(200):
        for dim, lo in ((2, lo2), (3, lo3)):
            nrm = rng.normal(size=dim)
            nrm /= np.linalg.norm(nrm)
            centre = lo + 0.5 * dx
            # a plane through the centre halves the cell; the complement rule
            assert orc.box_fraction_below_plane(nrm, float(nrm @ centre), lo, dx) == pytest.approx(0.5, abs=1e-12)
            alpha = float(nrm @ centre + rng.uniform(-0.6, 0.6) * dx)
            f = orc.box_fraction_below_plane(nrm, alpha, lo, dx)
            assert f + orc.box_fraction_below_plane(-nrm, -alpha, lo, dx) == pytest.approx(1.0, abs=1e-12)
            if dim == 2:
                assert f == pytest.approx(clipped_square_area(nrm, alpha, lo, dx) / dx ** 2, abs=1e-12)
    for _ in range(6):  # 3-D against the sliced exact areas
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        alpha = float(nrm @ (lo3 + 0.5 * dx) + rng.uniform(-0.6, 0.6) * dx)
        f = orc.box_fraction_below_plane(nrm, alpha, lo3, dx)
        assert f == pytest.approx(clipped_cube_volume(nrm, alpha, lo3, dx) / dx ** 3, abs=1e-6)


@pytest.mark.parametrize("dim", [2, 3])
def test_tangent_sphere_fraction_bounds_the_sphere(orc, dim):
    # the half-space below a tangent plane contains the sphere: the summed fractions bound its
    # volume from above, and at R = 32 dx the excess is the small curvature term
    R, dx = 0.25, 0.25 / 32
    c = np.array([0.5, 0.5, 0.5])[:dim] + 0.3 * dx  # off the grid's symmetry lines
    n = int(round(1.0 / dx))
    lo_idx = int((c[0] - R) / dx) - 2
    hi_idx = int((c[0] + R) / dx) + 2
    rng = range(lo_idx, hi_idx)
    total = 0.0
    if dim == 2:
        for i in rng:
            for j in rng:
                total += orc.tangent_sphere_fraction([i * dx, j * dx], dx, c, R)
        exact = math.pi * R ** 2 / dx ** 2
    else:
        for i in rng:
            for j in rng:
                for k in rng:
                    x = np.array([i + 0.5, j + 0.5, k + 0.5]) * dx - c
                    d = np.linalg.norm(x)
                    if d + dx <= R:
                        total += 1.0
                    elif d - dx < R:
                        total += orc.tangent_sphere_fraction([i * dx, j * dx, k * dx], dx, c, R)
        exact = 4.0 / 3.0 * math.pi * R ** 3 / dx ** 3
    assert n > 0
    assert exact < total < exact * (1 + 1e-3), (total / exact - 1)


def test_oracle_shock_bubble_pure_cells(orc, g5):
    cfg = g5.make_config(2, (16, 8), 1.0 / 128, gamma=(1.4, 1.648))
    regs = regions_air_helium(g5)
    U = orc.init_shock_bubble(cfg, regs, 0.6, (1.0, 0.5), 0.25).reshape(128, 8, 256)
    S = [orc.region_state(cfg, r) for r in regs]
    blk = np.arange(128)[:, None]
    cell = np.arange(256)[None, :]
    xc = ((blk % 16) * 16 + cell % 16 + 0.5) / 128
    yc = ((blk // 16) * 16 + cell // 16 + 0.5) / 128
    d = np.hypot(xc - 1.0, yc - 0.5)
    far = d > 0.25 + 1.0 / 128
    for v in range(8):
        assert np.all(U[:, v][far & (xc <= 0.6)] == S[0][v])
        assert np.all(U[:, v][far & (xc > 0.6)] == S[1][v])
    inside = d < 0.25 - 1.0 / 128
    for v in range(8):
        assert np.allclose(U[:, v][inside], S[2][v], rtol=1e-15, atol=0)
    # cut cells: partial densities and momenta linear in the fraction, alpha in (0, 1)
    cut = ~far & ~inside
    a = U[:, 4][cut]
    phi0 = a  # alpha0 is 1 in region 1 and 0 in the bubble
    assert np.all((a >= 0) & (a <= 1))
    assert np.allclose(U[:, 5][cut], phi0 * S[1][5] + (1 - phi0) * S[2][5], rtol=1e-15)
    assert np.allclose(U[:, 7][cut], phi0 * S[1][7] + (1 - phi0) * S[2][7], rtol=1e-15, atol=1e-17)
    # the bubble's helium mass is its density times the summed fractions, close to pi R^2
    area = (1 - U[:, 4]).sum() / 128 ** 2
    assert math.pi * 0.25 ** 2 < area < math.pi * 0.25 ** 2 * (1 + 2e-3)


def test_oracle_triple_point(orc, g5):
    for dim, nb in ((2, (7, 3)), (3, (4, 2, 2))):
        cfg = g5.make_config(dim, nb, 1.0 / 16 if dim == 3 else 1.0 / 48, gamma=(1.5, 1.4))
        regs = [(1.0, (1.0, 0.0), 1.0), (0.0, (0.0, 1.0), 0.1), (1.0, (0.125, 0.0), 0.1)]
        xd, yd = 1.0, (1.5 if dim == 3 else 0.75)
        U = orc.init_triple_point(cfg, regs, xd, yd)
        nv, nc = (8, 256) if dim == 2 else (9, 4096)
        U = U.reshape(-1, nv, nc)
        S = [orc.region_state(cfg, r) for r in regs]
        for b in range(U.shape[0]):
            bx, by, bz = b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1])
            for c in range(nc):
                ci, cj, ck = c % 16, (c // 16) % 16, c // 256
                x = (bx * 16 + ci + 0.5) * cfg.dx
                y = (by * 16 + cj + 0.5) * cfg.dx
                z = (bz * 16 + ck + 0.5) * cfg.dx
                r = 0 if x <= xd else (1 if (y if dim == 2 else math.sqrt(y * y + z * z)) <= yd else 2)
                assert np.array_equal(U[b, :, c], S[r]), (b, c)


def compare_bubble(U_prod, U_orc, cfg, a_pre, a_bub, tol_cut, exact_cells):
    """Product vs oracle shock-bubble states: bitwise where both volume fractions are 0 or both 1
    (away from the surface), the cut cells' fractions within tol_cut; the oracle's summed fractions
    bound the exact bubble volume from above (tangent planes), the product's (sub-cell quadrature)
    are within 0.5% of it."""
    nv, nc = (8, 256) if cfg.dim == 2 else (9, 4096)
    ia0 = 4 if cfg.dim == 2 else 5
    P = U_prod.reshape(-1, nv, nc)
    O = U_orc.reshape(-1, nv, nc)
    fp = (P[:, ia0] - a_pre) / (a_bub - a_pre)
    fo = (O[:, ia0] - a_pre) / (a_bub - a_pre)
    same = ((fp == 0) & (fo == 0)) | ((fp == 1) & (fo == 1))
    assert same.mean() > 0.9
    for v in range(nv):
        assert np.array_equal(P[:, v][same], O[:, v][same]), v
    worst = np.abs(fp - fo).max()
    assert worst <= tol_cut, worst
    assert exact_cells < fo.sum() < exact_cells * 1.01
    assert abs(fp.sum() / exact_cells - 1) < 5e-3
    return worst


def test_host_shock_bubble_matches_oracle(orc, g5):
    cfg = g5.make_config(2, (16, 8), 1.0 / 128, gamma=(1.4, 1.648))
    regs = regions_air_helium(g5)
    prod = g5.init_shock_bubble(cfg, g5_regions(g5, regs), 0.6, (1.0, 0.5), 0.25)
    compare_bubble(prod, orc.init_shock_bubble(cfg, regs, 0.6, (1.0, 0.5), 0.25), cfg, 1.0, 0.0, 0.02,
                   math.pi * (0.25 * 128) ** 2)
    cfg3, _ = cases.air_water_3d(nb=2)
    regs3 = regions_air_water(g5)
    prod3 = g5.init_shock_bubble(cfg3, g5_regions(g5, regs3), 0.25, (0.5, 0.5, 0.5), 0.2)
    compare_bubble(prod3, orc.init_shock_bubble(cfg3, regs3, 0.25, (0.5, 0.5, 0.5), 0.2), cfg3, 0.0, 1.0, 0.04,
                   4.0 / 3.0 * math.pi * (0.2 * 32) ** 3)


def test_host_triple_point_matches_oracle(orc, g5):
    for dim, nb in ((2, (7, 3)), (3, (4, 2, 2))):
        cfg = g5.make_config(dim, nb, 1.0 / 16 if dim == 3 else 1.0 / 48, gamma=(1.5, 4.4), pinf=(0.0, 6e8))
        regs = [(1.0, (1.0, 0.0), 1.0, 0.5), (0.0, (0.0, 1000.0), 1e5), (1.0, (0.125, 0.0), 0.1)]
        yd = 1.5 if dim == 3 else 0.75
        prod = g5.init_triple_point(cfg, g5_regions(g5, regs), 1.0, yd)
        assert np.array_equal(prod, orc.init_triple_point(cfg, regs, 1.0, yd))
