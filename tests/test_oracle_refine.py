"""The oracle's refine flags (CPU): ComputeRefineFlags of scheme/ComputeRefineFlags.cpp -- the
Lohner indicator (split and unsplit) and the simple normalised gradient on the partial densities
IAD0 and IAD1, the per-cell update_flag and the maximum over the block -- pinned by answers the
definitions fix by hand:
  - a step of IAD0 from a to b along x: the Lohner split value |b - a| / (|b - a| + eps (3|a| + |b|))
    in the cell left of the step, (|b - a|) / (|b - a| + eps (|a| + 3|b|)) right of it, 0 along y;
    the unsplit value sqrt(dx2^2 / (tx^2 + ty^2)) with ty = 4 eps |C|; the simple gradient
    |b - a| / max(|a|, |b|), 0 below 0.001;
  - a field linear in x and y with exactly representable values: every second difference is 0 in
    the interior; at a non-periodic border the outside neighbour is the cell itself;
  - a partial density that vanishes with its neighbours gives 0/0: the NaN takes part in no
    comparison, so the cell votes "nothing" and the block cannot coarsen;
  - the flag order (0 initial < 1 coarsen < 2 nothing < 3 refine) under the block maximum, the
    level limits, max-accumulation into the caller's flags, the periodic wrap, and the transposed
    field giving the transposed flags;
  - on a two-level mesh: a coarse cell next to a refined block reads the average of the four fine
    cells (0 second difference on a linear field) and a fine cell next to a coarse leaf reads its
    coarse parent (a known non-zero value on a linear field); in 3-D the same with the eight fine
    cells and a field along each of the three axes in turn.
The flag values and their order are the ones the block maximum needs (the absent core defines
AMRContextBase's constants); they are this library's output values.
"""
import numpy as np
import pytest

LOHNER_SPLIT, LOHNER_UNSPLIT, SIMPLE_GRADIENT = 0, 1, 2
EPS = 0.02


def cfg_of(g5, nb=(4, 2), periodic=(0, 0, 0)):
    bc = tuple("PERIODIC" if periodic[d // 2] else "ZERO_GRADIENT" for d in range(4)) + ("NONE", "NONE")
    return g5.make_config(2, nb, 1.0 / 64, periodic=periodic, bc=bc)


def state(nb, f0, f1):
    """U[block][var][cell] (lexicographic blocks) with IAD0 = f0(gx, gy), IAD1 = f1(gx, gy) at the
    global cell indices, the other variables 1"""
    nbx, nby = nb
    U = np.ones((nby, nbx, 8, 16, 16))
    gx = np.arange(nbx)[None, :, None, None] * 16 + np.arange(16)[None, None, None, :]
    gy = np.arange(nby)[:, None, None, None] * 16 + np.arange(16)[None, None, :, None]
    U[:, :, 5] = np.broadcast_to(f0(gx, gy), (nby, nbx, 16, 16))
    U[:, :, 7] = np.broadcast_to(f1(gx, gy), (nby, nbx, 16, 16))
    return U.reshape(nbx * nby, 8, 256).ravel()


def run(orc, cfg, U, indicator, level_min=0, level_max=5, base_level=2, refine=0.1, coarsen=0.0125, flags=None,
        refined=None):
    o = orc.Oracle(cfg) if refined is None else orc.Oracle(cfg, refined=refined)
    o.set_state(U)
    return o.refine_flags(indicator, level_min, level_max, base_level, refine, coarsen, EPS, flags)


def test_step_closed_forms(g5, orc):
    a, b = 1.0, 3.0
    cfg = cfg_of(g5)
    # step between global x cells 23 and 24 (block 1, local 7|8): blocks 1 and 5 hold it
    U = state((4, 2), lambda x, y: np.where(x < 24, a, b) + 0 * y, lambda x, y: 0.5 + 0 * x + 0 * y)
    left = (b - a) / ((b - a) + EPS * (3 * a + b))
    right = (b - a) / ((b - a) + EPS * (a + 3 * b))
    assert left > right
    f, imax = run(orc, cfg, U, LOHNER_SPLIT)
    assert imax[1] == imax[5] == left
    assert np.all(imax[[0, 2, 3, 4, 6, 7]] == 0.0)
    # unsplit: sqrt(dx2^2 / (tx^2 + ty^2)), ty = eps (|C| + 2|C| + |C|)
    f_u, imax_u = run(orc, cfg, U, LOHNER_UNSPLIT)
    tx = (b - a) + EPS * (3 * a + b)
    assert imax_u[1] == pytest.approx(np.sqrt((b - a) ** 2 / (tx ** 2 + (4 * EPS * a) ** 2)), rel=1e-15)
    f_g, imax_g = run(orc, cfg, U, SIMPLE_GRADIENT)
    assert imax_g[1] == (b - a) / b and np.all(imax_g[[0, 2, 3]] == 0.0)
    # the step blocks refine, the others (level 2 > level_min 0, indicator 0) coarsen
    for fl in (f, f_u, f_g):
        assert list(fl) == [1, 3, 1, 1, 1, 3, 1, 1]


def test_simple_gradient_ignores_tiny_values(g5, orc):
    cfg = cfg_of(g5)
    U = state((4, 2), lambda x, y: np.where(x < 24, 1e-4, 9e-4) + 0 * y, lambda x, y: 0.5 + 0 * x + 0 * y)
    _, imax = run(orc, cfg, U, SIMPLE_GRADIENT)
    assert np.all(imax == 0.0)


def test_linear_field_and_borders(g5, orc):
    cfg = cfg_of(g5)
    U = state((4, 2), lambda x, y: 1.0 + x + 2.0 * y, lambda x, y: 2.0 + 0 * x + 0 * y)
    _, imax = run(orc, cfg, U, LOHNER_SPLIT)
    # interior second differences vanish; at the x border cell gx = 0 the outside neighbour is the
    # cell itself (L = C, R = C + 1): 1 / (1 + eps (4|C| + 1)); at the y border gy = 0 (L = C,
    # R = C + 2): 2 / (2 + eps (4|C| + 2)); both largest at the corner cell, C = 1
    C = 1.0
    x_border = 1.0 / (1.0 + EPS * (4 * C + 1))
    y_border = 2.0 / (2.0 + EPS * (4 * C + 2))
    assert imax[0] == max(x_border, y_border)
    # the block at the far x border (gx = 48..63): there L = C - 1, R = C (outside is self), so
    # 1 / (1 + eps (4|C| - 1)), largest at C = 64; its y border cells give 2 / (2 + eps (4|C| + 2)),
    # largest at gx = 48, C = 49
    assert imax[3] == max(1.0 / (1.0 + EPS * (4 * 64.0 - 1)), 2.0 / (2.0 + EPS * (4 * 49.0 + 2)))
    # a block touching no border sees only zero second differences: none here (2 rows of blocks),
    # so compare with a periodic box, where every second difference but the wrap's vanishes
    cfgp = cfg_of(g5, periodic=(1, 1, 0))
    Up = state((4, 2), lambda x, y: 1.0 + 0.0 * x + 0 * y, lambda x, y: 2.0 + 0 * x + 0 * y)
    _, imax_p = run(orc, cfgp, Up, LOHNER_SPLIT)
    assert np.all(imax_p == 0.0)


def test_vanishing_density_gives_nan_and_blocks_coarsening(g5, orc):
    cfg = cfg_of(g5)
    U = state((4, 2), lambda x, y: 1.0 + 0 * x + 0 * y, lambda x, y: 0.0 * x + 0 * y)
    f, imax = run(orc, cfg, U, LOHNER_SPLIT)
    assert np.all(imax == 0.0) and np.all(f == 2)
    # with IAD1 non-zero the same field coarsens
    U = state((4, 2), lambda x, y: 1.0 + 0 * x + 0 * y, lambda x, y: 1.0 + 0 * x + 0 * y)
    f, _ = run(orc, cfg, U, LOHNER_SPLIT)
    assert np.all(f == 1)


def test_level_limits_and_accumulation(g5, orc):
    cfg = cfg_of(g5)
    U = state((4, 2), lambda x, y: np.where(x < 24, 1.0, 3.0) + 0 * y, lambda x, y: 0.5 + 0 * x + 0 * y)
    # at level_max the step blocks cannot refine: nothing; at level_min the others cannot coarsen
    f, _ = run(orc, cfg, U, LOHNER_SPLIT, level_min=2, level_max=2)
    assert np.all(f == 2)
    # max-accumulation into the caller's flags
    f0 = np.array([3, 0, 2, 0, 0, 0, 0, 0], dtype=np.int32)
    f, _ = run(orc, cfg, U, LOHNER_SPLIT, flags=f0)
    assert list(f) == [3, 3, 2, 1, 1, 3, 1, 1]


def test_periodic_wrap_and_transpose(g5, orc):
    cfgp = cfg_of(g5, nb=(4, 4), periodic=(1, 1, 0))
    # a step at the periodic x seam (gx = 0 | 63): the end blocks of every row see it
    U = state((4, 4), lambda x, y: np.where(x < 32, 1.0, 2.0) + 0 * y, lambda x, y: 1.0 + 0 * x + 0 * y)
    f, imax = run(orc, cfgp, U, LOHNER_SPLIT)
    assert list(f.reshape(4, 4)[0]) == [3, 3, 3, 3]
    # the transposed field gives the transposed flags and indicators
    Ut = state((4, 4), lambda x, y: np.where(y < 32, 1.0, 2.0) + 0 * x, lambda x, y: 1.0 + 0 * x + 0 * y)
    for ind in (LOHNER_SPLIT, LOHNER_UNSPLIT, SIMPLE_GRADIENT):
        f, imax = run(orc, cfgp, U, ind, refine=0.5)
        ft, imaxt = run(orc, cfgp, Ut, ind, refine=0.5)
        assert np.array_equal(f.reshape(4, 4).T, ft.reshape(4, 4))
        assert np.array_equal(imax.reshape(4, 4).T, imaxt.reshape(4, 4))


def test_two_level_neighbours(g5, orc):
    # 4x4 brick, block (1, 1) refined; IAD0 linear in x at the cell centres, in units of the
    # coarse cell: coarse cell gx has x = gx + 0.5, fine cell fx has x = fx / 2 + 0.25
    cfg = g5.make_config(2, (4, 4), 1.0 / 64)
    o = orc.Oracle(cfg, refined=np.array([[1, 1, 0]], dtype=np.int32))
    lv, bc = o.block_levels(), o.block_coords()[: o.nowned]
    U = np.ones((o.nowned, 8, 256))
    i = np.arange(16)
    for k in range(o.nowned):
        g = bc[k, 0] * 16 + i[None, :] + 0 * i[:, None]
        x = g + 0.5 if lv[k] == 0 else g / 2.0 + 0.25
        U[k, 5] = x.ravel()
        U[k, 7] = 1.0
    o.set_state(U.ravel())
    f, imax = o.refine_flags(LOHNER_SPLIT, 0, 5, 2, 0.1, 0.0125, EPS)
    # coarse leaves: restriction of the linear field is the coarse centre value, so interior
    # coarse cells see 0; the brick border cells see the outside-is-self value
    # fine block (2, 2), cell (fx, fy) = (32, 32): along x it reads the coarse parent of fx = 31
    # (coarse cell 15, x = 15.5): L = 15.5, C = 16.25, R = 16.75; along y the parent of fy = 31 is
    # coarse cell (16, 15), x = 16.5: L = 16.5, C = R = 16.25.  The y value is the block's largest.
    def lohner(L, C, R):
        return abs(L - 2 * C + R) / (abs(R - C) + abs(C - L) + EPS * (abs(L) + 2 * abs(C) + abs(R)))
    along_x, along_y = lohner(15.5, 16.25, 16.75), lohner(16.5, 16.25, 16.25)
    assert along_y > along_x > 0.0
    kf = [k for k in range(o.nowned) if lv[k] == 1 and tuple(bc[k, :2]) == (2, 2)][0]
    assert imax[kf] == along_y
    # the coarse leaves (1, 0) and (0, 1) next to the refined block, away from the brick's x and y
    # borders along the field's direction: block (0, 1) has the x border, block (2, 1) does not
    k21 = [k for k in range(o.nowned) if lv[k] == 0 and tuple(bc[k, :2]) == (2, 1)][0]
    assert imax[k21] == 0.0


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_two_level_neighbours_3d(g5, orc, axis):
    # 4x4x4 brick, block (1, 1, 1) refined; IAD0 linear along one axis at the cell centres, in
    # units of the coarse cell (coarse cell g: g + 0.5, fine cell f: f / 2 + 0.25); IAD1 = 1
    cfg = g5.make_config(3, (4, 4, 4), 1.0 / 64)
    o = orc.Oracle(cfg, refined=np.array([[1, 1, 1]], dtype=np.int32))
    lv, bc = o.block_levels(), o.block_coords()[: o.nowned]
    U = np.ones((o.nowned, 9, 16, 16, 16))
    i = np.arange(16)
    shape = [1, 1, 1]
    shape[2 - axis] = 16  # cell arrays are [z][y][x]
    for k in range(o.nowned):
        g = (bc[k, axis] * 16 + i).reshape(shape)
        U[k, 6] = np.broadcast_to(g + 0.5 if lv[k] == 0 else g / 2.0 + 0.25, (16, 16, 16))
    o.set_state(U.ravel())
    f, imax = o.refine_flags(LOHNER_SPLIT, 0, 5, 2, 0.1, 0.0125, EPS)

    def lohner(L, C, R):
        return abs(L - 2 * C + R) / (abs(R - C) + abs(C - L) + EPS * (abs(L) + 2 * abs(C) + abs(R)))
    # coarse leaves: every second difference is 0 -- inside a block, across blocks, and next to the
    # refined block, whose eight fine cells average to the coarse centre value -- except in the
    # brick's first and last blocks along the field's axis, whose border cells read themselves
    for k in range(o.nowned):
        if lv[k] == 0:
            assert (imax[k] == 0.0) == (bc[k, axis] in (1, 2)), (axis, bc[k])
    # fine block (2, 2, 2): its low faces read coarse parents.  Along the axis, f = 32 reads the
    # parent of f = 31 (coarse 15: 15.5), C = 16.25, R = 16.75.  Across it, at f_axis = 32, the
    # parent of the cell one step down on another axis has 16.5 along the axis and C = R = 16.25;
    # that value is the block's largest (larger f_axis have larger |C| at the same |L - C| = 0.25).
    along, across = lohner(15.5, 16.25, 16.75), lohner(16.5, 16.25, 16.25)
    assert across > along > 0.0
    kf = [k for k in range(o.nowned) if lv[k] == 1 and tuple(bc[k]) == (2, 2, 2)][0]
    assert imax[kf] == across
    # fine blocks (3, 3, 3) touch coarse leaves only on their high faces: f = 63 along the axis
    # reads coarse 32 (32.5) with C = 31.75 and L = 31.25; across it, f_axis = 48 reads a parent
    # with 24.5 and C = R = 24.25 (the larger of the two)
    hi_along, hi_across = lohner(31.25, 31.75, 32.5), lohner(24.25, 24.25, 24.5)
    kf3 = [k for k in range(o.nowned) if lv[k] == 1 and tuple(bc[k]) == (3, 3, 3)][0]
    assert imax[kf3] == max(hi_along, hi_across)
