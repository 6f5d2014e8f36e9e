"""Refine flags on the GPU (g5_refine_flags) against the oracle's (pinned by closed forms in
tests/test_oracle_refine.py): ComputeRefineFlags of scheme/ComputeRefineFlags.cpp with the Lohner
split, Lohner unsplit and simple-gradient indicators on the partial densities.

Bar: integer flags equal, and the per-block largest indicator value bit-identical (the kernel
evaluates the reference's expressions in the same order, FP64 with IEEE division and square root,
no contraction).  Cases: the 2-D air-helium shock-bubble at 512x128 after 50 steps (walls in y,
zero-gradient x: the outside-is-self rule at both kinds of border), the 3-D air-water THINC
shock-bubble at 64^3, a periodic 2-D box (wrap), the 2-D and 3-D two-level shock-bubbles after the
shock has entered the fine blocks (restriction and coarse-parent neighbours), and max-accumulation
into flags the caller set before.
"""
import sys

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

INDICATORS = [(0, "LOHNER_SPLIT"), (1, "LOHNER_UNSPLIT"), (2, "SIMPLE_GRADIENT")]


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def compare(ctx, o, label, base_level=2, level_min=1, level_max=4, flags=None, **th):
    out = []
    for ind, name in INDICATORS:
        fg, ig = ctx.refine_flags(ind, level_min, level_max, base_level, flags=flags, **th)
        fo, io = o.refine_flags(ind, level_min, level_max, base_level, th.get("refine_threshold", 0.01),
                                th.get("coarsen_threshold", 0.00125), th.get("epsilon_lohner", 0.02), flags)
        counts = np.bincount(fg, minlength=4)
        _announce(f"{label} {name}: flags (init, coarsen, nothing, refine) = {list(counts)} "
                  f"equal={np.array_equal(fg, fo)} indicator max bit-exact={np.array_equal(ig, io)}")
        assert np.array_equal(fg, fo), (name, np.nonzero(fg != fo))
        assert np.array_equal(ig, io), (name, np.nonzero(ig != io))
        out.append(fg)
    return out


def test_refine_flags_2d_shock_bubble(g5, orc):
    cfg, U0 = cases.air_helium_2d()
    with g5.Context(cfg) as ctx:
        ctx.upload(U0)
        ctx.run(50)
        U = ctx.download()
        o = orc.Oracle(cfg)
        o.set_state(U)
        split, unsplit, grad = compare(ctx, o, "2d air-helium 512x128 +50 steps")
        # the bubble and the shock refine; under the Lohner indicators the helium-free blocks vote
        # nothing (0/0 of the vanishing helium density), so none coarsens; the simple gradient, 0
        # below 0.001, lets the quiet air coarsen
        for f in (split, unsplit):
            assert np.any(f == 3) and np.any(f == 2) and not np.any(f == 1)
        assert np.any(grad == 3) and np.any(grad == 1)


def test_refine_flags_3d_air_water(g5, orc):
    cfg, U0 = cases.air_water_3d(nb=4)
    with g5.Context(cfg) as ctx:
        ctx.upload(U0)
        ctx.run(5)
        U = ctx.download()
        o = orc.Oracle(cfg)
        o.set_state(U)
        compare(ctx, o, "3d air-water 64^3 +5 steps", base_level=2, level_min=2, level_max=6)


def test_refine_flags_periodic_and_accumulation(g5, orc):
    cfg, U0, _ = cases.interface_advection_2d(nb=(4, 4))
    with g5.Context(cfg) as ctx:
        ctx.upload(U0)
        o = orc.Oracle(cfg)
        o.set_state(U0)
        compare(ctx, o, "2d periodic blob", refine_threshold=0.05, coarsen_threshold=0.01)
        pre = np.array([3, 0, 2, 1] * 4, dtype=np.int32)
        compare(ctx, o, "2d periodic blob, preset flags", flags=pre, refine_threshold=0.05,
                coarsen_threshold=0.01)


def test_refine_flags_two_level(g5, orc):
    cfg, Uc = cases.air_helium_2d(nbx=8, nby=4)
    _, Uf = cases.air_helium_2d(nbx=16, nby=8)
    refined = [[3, 1], [4, 1], [3, 2]]
    with g5.Context(cfg, refined=refined, prolongation=1) as ctx:
        lay = ctx.block_layout()
        Uc3, Uf3 = Uc.reshape(-1, 8, 256), Uf.reshape(-1, 8, 256)
        U0 = np.stack([Uc3[x + 8 * y] if lv == 0 else Uf3[x + 16 * y] for x, y, _, lv in lay]).ravel()
        ctx.upload(U0)
        ctx.run(150)
        U = ctx.download()
        o = orc.Oracle(cfg, refined=np.array([[x, y, 0] for x, y in refined], dtype=np.int32), prolongation=1)
        o.set_state(U)
        compare(ctx, o, "two-level shock-bubble +150 steps", base_level=2, level_min=2, level_max=3)


def test_refine_flags_two_level_3d(g5, orc):
    # the 3-D air-helium shock-bubble on a 4^3 brick with a corner of four refined blocks, after
    # the shock has entered them: coarse cells next to fine blocks read averages of eight fine
    # cells, fine cells next to coarse leaves read their parents on faces of all three orientations
    refined = [[1, 1, 1], [2, 1, 1], [1, 2, 1], [1, 1, 2]]
    cfg, Uc = cases.air_helium_3d(nb=4)
    _, Uf = cases.air_helium_3d(nb=8)
    with g5.Context(cfg, refined=refined, prolongation=1) as ctx:
        lay = ctx.block_layout()
        Uc3, Uf3 = Uc.reshape(-1, 9, 4096), Uf.reshape(-1, 9, 4096)
        U0 = np.stack([Uc3[x + 4 * (y + 4 * z)] if lv == 0 else Uf3[x + 8 * (y + 8 * z)]
                       for x, y, z, lv in lay]).ravel()
        ctx.upload(U0)
        ctx.run(45)
        U = ctx.download()
        o = orc.Oracle(cfg, refined=np.array(refined, dtype=np.int32), prolongation=1)
        o.set_state(U)
        compare(ctx, o, "3-D two-level air-helium +45 steps", base_level=2, level_min=2, level_max=3)
