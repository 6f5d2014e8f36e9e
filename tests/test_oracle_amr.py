"""The oracle's two-level static mesh (CPU): the coarse/fine coupling of the reference's V0 step
(ghost fill by restriction and prolongation, ConvertToPrimitivesVariablesFunctor.cpp:541-918;
fine-face flux sums at coarse faces, ReadFluxesAndConservativeUpdateFunctor.h:246-307 and the
NEIGHBOR_IS_FINER branches of .cpp:137-199), pinned by
  - closed forms: on a field linear in every conservative variable the restriction of a coarse
    cell is its centre value (the mean of the 2^dim fine cells), minmod prolongation reproduces
    the fine cells' values and SIMPLE_COPY the coarse value; the conservative interpolations of
    order 2 and 4 reproduce the fine cell averages of any polynomial of degree <= 2 (<= 4) in each
    variable, and the four children of a coarse cell average back to it (any field);
  - invariants: a uniform state stays uniform bit for bit across the levels, and the totals of
    mass, energy and momentum (cell values times the level's cell volume) are conserved to
    round-off in a periodic box whose refined region is L-shaped (so coarse cells meet fine ones
    on faces of both orientations and at a concave corner);
  - determinism of the layout: coarse leaves first, then fine blocks, each in lexicographic order.
"""
import numpy as np
import pytest

REFINED = [[1, 1, 0], [2, 1, 0], [1, 2, 0]]


def periodic_cfg(g5, dim=2, nb=4, **kw):
    opts = dict(periodic=(1, 1, 1), bc=("PERIODIC",) * 6, gamma=(1.4, 1.4), cfl=0.6)
    opts.update(kw)
    return g5.make_config(dim, (nb,) * dim, 1.0 / (16 * nb), **opts)


def sample(o, cfg, f):
    """U[block][var][cell] of pure material 0 with (rho, u, v, p) = f(x, y) at the cell centres"""
    lv = o.block_levels()
    bc = o.block_coords()[: o.nowned]
    U = np.zeros((o.nowned, 8, 256))
    i = np.arange(16)
    for k in range(o.nowned):
        dx = cfg.dx / (2 ** lv[k])
        X = (bc[k, 0] * 16 + i[None, :] + 0.5) * dx + 0 * i[:, None]
        Y = (bc[k, 1] * 16 + i[:, None] + 0.5) * dx + 0 * i[None, :]
        rho, u, v, p = f(X, Y)
        E = p / 0.4 + 0.5 * rho * (u * u + v * v)
        for var, val in enumerate([rho, E, rho * u, rho * v, 1.0, rho, 0.0, 0.0]):
            U[k, var] = np.broadcast_to(val, (16, 16)).ravel()
    return U.ravel()


def totals(o, cfg, U):
    lv = o.block_levels()
    V = U.reshape(-1, 8, 256)
    vol = (cfg.dx / (2.0 ** lv)) ** 2
    return np.array([(V[:, v].sum(axis=1) * vol).sum() for v in range(8)])


def test_layout(g5, orc):
    o = orc.Oracle(periodic_cfg(g5), refined=REFINED)
    lv, bc = o.block_levels(), o.block_coords()[: o.nowned]
    assert o.nowned == 13 + 12 and list(lv[:13]) == [0] * 13 and list(lv[13:]) == [1] * 12
    coarse = [tuple(c[:2]) for c in bc[:13]]
    assert coarse == sorted(coarse, key=lambda c: (c[1], c[0])) and (1, 1) not in coarse
    fine = [tuple(c[:2]) for c in bc[13:]]
    assert fine == sorted(fine, key=lambda c: (c[1], c[0])) and (2, 2) in fine and (4, 2) in fine


@pytest.mark.parametrize("prolongation", [0, 1])
def test_restriction_and_prolongation_on_a_linear_field(g5, orc, prolongation):
    cfg = periodic_cfg(g5)
    o = orc.Oracle(cfg, refined=REFINED, prolongation=prolongation)
    # every conservative variable linear in x and y (the restriction and the interior prolongation
    # stencils stay away from the periodic wrap)
    a, bx, by = np.array([2.0, 5.0, 0.3, -0.2, 1.0, 2.0, 0.0, 0.0]), np.array([0.5, 1.0, 0.25, 0.5, 0, 0.5, 0, 0]), \
        np.array([-0.25, 0.75, -0.5, 0.125, 0, -0.25, 0, 0])
    lv, bc = o.block_levels(), o.block_coords()[: o.nowned]
    U = np.zeros((o.nowned, 8, 256))
    i = np.arange(16)
    for k in range(o.nowned):
        dx = cfg.dx / 2 ** lv[k]
        X = ((bc[k, 0] * 16 + i[None, :] + 0.5) * dx + 0 * i[:, None]).ravel()
        Y = ((bc[k, 1] * 16 + i[:, None] + 0.5) * dx + 0 * i[None, :]).ravel()
        for v in range(8):
            U[k, v] = a[v] + bx[v] * X + by[v] * Y
    o.set_state(U.ravel())
    dx = cfg.dx
    # restriction: coarse cell (20, 20) lies in refined block (1, 1): the linear field at its centre
    u = o.cons_level(0, (20, 20))
    x, y = 20.5 * dx, 20.5 * dx
    ref = a + bx * x + by * y
    for v in (1, 2, 3, 4, 5):
        assert u[v] == pytest.approx(ref[v], rel=1e-14, abs=1e-14), v
    assert u[6] == 1.0 - u[4] and u[0] == u[5] + u[7]
    # prolongation: fine cell (30, 40) lies in coarse leaf (0, 1) (coarse cell (15, 20))
    u = o.cons_level(1, (30, 40))
    xf, yf = 30.5 * dx / 2, 40.5 * dx / 2
    xc, yc = 15.5 * dx, 20.5 * dx
    ref = a + bx * (xf if prolongation else xc) + by * (yf if prolongation else yc)
    for v in (1, 2, 3, 5):
        assert u[v] == pytest.approx(ref[v], rel=1e-14, abs=1e-14), v
    assert u[0] == u[5] + u[7]


@pytest.mark.parametrize("prolongation", [0, 1, 2, 3])
def test_uniform_state_is_preserved(g5, orc, prolongation):
    cfg = periodic_cfg(g5)
    o = orc.Oracle(cfg, refined=REFINED, prolongation=prolongation)
    U0 = sample(o, cfg, lambda x, y: (1.0 + 0 * x, 0.5 + 0 * x, -0.25 + 0 * x, 1.0 + 0 * x))
    o.set_state(U0)
    o.run(5)
    assert np.array_equal(o.get_state(), U0)


@pytest.mark.parametrize("prolongation", [0, 1, 2, 3])
def test_conservation_across_levels(g5, orc, prolongation):
    cfg = periodic_cfg(g5)
    o = orc.Oracle(cfg, refined=REFINED, prolongation=prolongation)
    U0 = sample(o, cfg, lambda x, y: (1.0 + 0.5 * np.exp(-((x - 0.3) ** 2 + (y - 0.4) ** 2) / 0.01), 1.0 + 0 * x,
                                      0.5 + 0 * x, 1.0 + 0.2 * np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.02)))
    o.set_state(U0)
    T0 = totals(o, cfg, U0)
    _, n = o.run(40)
    U = o.get_state()
    T1 = totals(o, cfg, U)
    assert n == 40 and np.all(np.isfinite(U))
    # mass (ID), energy and momentum: telescoping flux sums, round-off only
    for v in (0, 1, 2, 3):
        assert abs(T1[v] - T0[v]) <= 2e-15 * abs(T0[v]), (v, T0[v], T1[v])
    # the pulse has crossed into and out of the fine region
    assert np.abs(U - U0).max() > 1e-3


def _mono_avg(a, lo, hi):
    return (hi ** (a + 1) - lo ** (a + 1)) / ((a + 1) * (hi - lo))


def _poly_avg(poly, x0, x1, y0, y1):
    return sum(c * _mono_avg(a, x0, x1) * _mono_avg(b, y0, y1) for (a, b), c in poly.items())


POLY2 = {(0, 0): 1.0, (1, 0): 0.5, (0, 1): 0.25, (2, 0): 0.125, (1, 1): -0.0625, (0, 2): 0.03125, (2, 1): 0.01}
POLY4 = {**POLY2, (3, 0): 0.002, (4, 0): -1e-4, (0, 4): 2e-4, (2, 2): -3e-4, (3, 1): 1e-4, (4, 3): 1e-7}


@pytest.mark.parametrize("prolongation,poly", [(2, POLY2), (3, POLY4)])
def test_conservative_interpolation_is_exact_on_polynomials(g5, orc, prolongation, poly):
    # every cell holds the exact average of a polynomial over the cell (coordinates in coarse
    # cells); the interpolated fine cell then holds the fine cell's exact average
    cfg = periodic_cfg(g5)
    o = orc.Oracle(cfg, refined=REFINED, prolongation=prolongation)
    lv, bc = o.block_levels(), o.block_coords()[: o.nowned]
    scale = {1: 3.0, 2: 0.5, 3: -0.25, 4: 0.001, 5: 1.0, 7: 2.0}  # E, rho u, rho v, alpha0, IAD0, IAD1
    U = np.ones((o.nowned, 8, 256))
    for k in range(o.nowned):
        h = 0.5 ** lv[k]
        for c in range(256):
            gx, gy = bc[k, 0] * 16 + c % 16, bc[k, 1] * 16 + c // 16
            a = _poly_avg(poly, gx * h, (gx + 1) * h, gy * h, (gy + 1) * h)
            for v, f in scale.items():
                U[k, v, c] = f * a
    o.set_state(U.ravel())
    # fine cells outside the fine blocks, in their ghost rings: left of x = 32, below y = 32, and
    # at the concave corner (coarse block (2, 2) is not refined: fine x, y >= 64 there)
    # (a stored fine cell would keep ID = 1: the interpolated state has ID = IAD0 + IAD1)
    for fx, fy in [(31, 40), (30, 41), (40, 31), (41, 30), (31, 31), (64, 70), (70, 64), (64, 64), (65, 65)]:
        u = o.cons_level(1, (fx, fy))
        ref = _poly_avg(poly, fx / 2, (fx + 1) / 2, fy / 2, (fy + 1) / 2)
        for v, f in scale.items():
            assert u[v] == pytest.approx(f * ref, rel=1e-12, abs=1e-12), (fx, fy, v)
        assert u[0] == u[5] + u[7] and u[6] == 1.0 - u[4]


@pytest.mark.parametrize("prolongation", [2, 3])
def test_conservative_interpolation_children_average_to_the_parent(g5, orc, prolongation):
    cfg = periodic_cfg(g5)
    o = orc.Oracle(cfg, refined=REFINED, prolongation=prolongation)
    rng = np.random.default_rng(7)
    o.set_state(rng.uniform(0.5, 2.0, o.nowned * 8 * 256))
    # coarse leaf cells next to the refined region, whose four children are fine ghost cells
    for cx, cy in [(15, 20), (20, 15), (15, 15), (23, 40), (40, 23)]:
        kids = [o.cons_level(1, (2 * cx + i, 2 * cy + j)) for j in (0, 1) for i in (0, 1)]
        parent = o.cons_level(0, (cx, cy))
        for v in (1, 2, 3, 4, 5, 7):
            mean = (kids[0][v] + kids[1][v] + kids[2][v] + kids[3][v]) / 4.0
            assert mean == pytest.approx(parent[v], rel=4e-16, abs=0.0), (cx, cy, v)


# ---- 3-D two-level mesh ----------------------------------------------------------------------
REFINED_3D = [[1, 1, 1], [2, 1, 1], [1, 2, 1], [1, 1, 2]]  # a 3-D corner of four blocks


def periodic_cfg3(g5, nb=4, **kw):
    opts = dict(periodic=(1, 1, 1), bc=("PERIODIC",) * 6, gamma=(1.4, 1.4), cfl=0.6)
    opts.update(kw)
    return g5.make_config(3, (nb,) * 3, 1.0 / (16 * nb), **opts)


def fill3(o, cfg, f):
    """U[block][var][cell] of a 3-D two-level mesh: var v = f(v, X, Y, Z) at the cell centres in
    units of the coarse cell"""
    lv, bc = o.block_levels(), o.block_coords()[: o.nowned]
    U = np.zeros((o.nowned, 9, 4096))
    i = np.arange(16)
    for k in range(o.nowned):
        h = 0.5 ** lv[k]
        X = (bc[k, 0] * 16 + i[None, None, :] + 0.5) * h + 0 * i[:, None, None]
        Y = (bc[k, 1] * 16 + i[None, :, None] + 0.5) * h + 0 * i[:, None, None]
        Z = (bc[k, 2] * 16 + i[:, None, None] + 0.5) * h + 0 * i[None, None, :]
        for v in range(9):
            U[k, v] = np.broadcast_to(f(v, X, Y, Z), (16, 16, 16)).ravel()
    return U.ravel()


def test_3d_layout_and_uniform_state(g5, orc):
    cfg = periodic_cfg3(g5)
    for prolongation in (0, 1, 2, 3):
        o = orc.Oracle(cfg, refined=REFINED_3D, prolongation=prolongation)
        lv = o.block_levels()
        assert o.nowned == 64 - 4 + 32 and list(lv[:60]) == [0] * 60 and list(lv[60:]) == [1] * 32
        # rho, E, mu, mv, mw, a0, ar0, a1, ar1 of a uniform moving gas
        vals = [1.0, 3.0, 0.5, -0.25, 0.125, 1.0, 1.0, 0.0, 0.0]
        U0 = fill3(o, cfg, lambda v, X, Y, Z: vals[v] + 0 * X)
        o.set_state(U0)
        o.run(3)
        assert np.array_equal(o.get_state(), U0), prolongation


def test_3d_restriction_and_prolongation_closed_forms(g5, orc):
    cfg = periodic_cfg3(g5)
    a = np.array([2.0, 5.0, 0.3, -0.2, 0.1, 1.0, 2.0, 0.0, 0.0])
    b = np.array([[0.5, 1.0, 0.25, 0.5, 0.125, 0, 0.5, 0, 0], [-0.25, 0.75, -0.5, 0.125, 0.25, 0, -0.25, 0, 0],
                  [0.125, -0.5, 0.25, -0.125, 0.5, 0, 0.125, 0, 0]]) / 16.0
    for prolongation in (0, 1, 2, 3):
        o = orc.Oracle(cfg, refined=REFINED_3D, prolongation=prolongation)
        o.set_state(fill3(o, cfg, lambda v, X, Y, Z: a[v] + b[0, v] * X + b[1, v] * Y + b[2, v] * Z))
        # restriction: coarse cell (20, 20, 20) lies in refined block (1, 1, 1)
        u = o.cons_level(0, (20, 20, 20))
        ref = a + b[0] * 20.5 + b[1] * 20.5 + b[2] * 20.5
        for v in (1, 2, 3, 4, 6):
            assert u[v] == pytest.approx(ref[v], rel=1e-14, abs=1e-14), (prolongation, v)
        assert u[7] == 1.0 - u[5] and u[0] == u[6] + u[8]
        # prolongation: fine cell (30, 40, 40) lies in coarse leaf (0, 1, 1), coarse cell (15, 20, 20)
        u = o.cons_level(1, (30, 40, 40))
        xf = np.array([30.5, 40.5, 40.5]) / 2
        xc = np.array([15.5, 20.5, 20.5])
        pos = xc if prolongation == 0 else xf  # a copy of the parent, or exact on linear fields
        ref = a + b[0] * pos[0] + b[1] * pos[1] + b[2] * pos[2]
        for v in (1, 2, 3, 4, 6):
            assert u[v] == pytest.approx(ref[v], rel=1e-13, abs=1e-13), (prolongation, v)


@pytest.mark.parametrize("prolongation", [0, 1, 2, 3])
def test_3d_conservation_across_levels(g5, orc, prolongation):
    cfg = periodic_cfg3(g5)
    o = orc.Oracle(cfg, refined=REFINED_3D, prolongation=prolongation)

    def f(v, X, Y, Z):
        r2 = (X - 24) ** 2 + (Y - 26) ** 2 + (Z - 28) ** 2
        rho = 1.0 + 0.5 * np.exp(-r2 / 60.0)
        p = 1.0 + 0.2 * np.exp(-r2 / 90.0)
        u, w, s = 1.0, 0.5, -0.25
        return [rho, p / 0.4 + 0.5 * rho * (u * u + w * w + s * s), rho * u, rho * w, rho * s, 1.0 + 0 * X, rho,
                0 * X, 0 * X][v]

    U0 = fill3(o, cfg, f)
    o.set_state(U0)
    lv = o.block_levels()[: o.nowned]
    vol = (0.5 ** lv) ** 3

    def totals(U):
        V = U.reshape(-1, 9, 4096)
        return np.array([(V[:, v].sum(axis=1) * vol).sum() for v in range(9)])

    T0 = totals(U0)
    _, n = o.run(12)
    T1 = totals(o.get_state())
    assert n == 12
    for v in (0, 1, 2, 3, 4):
        assert abs(T1[v] - T0[v]) <= 2e-15 * abs(T0[v]), (v, T0[v], T1[v])


@pytest.mark.parametrize("prolongation", [2, 3])
def test_3d_conservative_interpolation_children_average_to_the_parent(g5, orc, prolongation):
    cfg = periodic_cfg3(g5)
    o = orc.Oracle(cfg, refined=REFINED_3D, prolongation=prolongation)
    rng = np.random.default_rng(11)
    o.set_state(rng.uniform(0.5, 2.0, o.nowned * 9 * 4096))
    for c in [(15, 20, 20), (20, 15, 20), (20, 20, 15), (15, 15, 15), (40, 20, 20)]:
        kids = [o.cons_level(1, (2 * c[0] + i, 2 * c[1] + j, 2 * c[2] + k)) for k in (0, 1) for j in (0, 1)
                for i in (0, 1)]
        parent = o.cons_level(0, c)
        for v in (1, 2, 3, 4, 5, 6, 8):
            mean = sum(kd[v] for kd in kids) / 8.0
            assert mean == pytest.approx(parent[v], rel=8e-16, abs=0.0), (c, v)
