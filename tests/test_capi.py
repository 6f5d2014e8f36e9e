"""The C ABI library loads, exports every entry point declared in include/*.h, and its host-only
entry points and error paths behave (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(headers=None):
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h") and (headers is None or fn in headers):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(g5_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol(g5):
    # libg5.so: godunov_five_eq.h; libg5h5.so: godunov_five_eq_io.h
    from godunov_five_eq_amd import hdf5io
    assert set(declared_functions()) == set(declared_functions(["godunov_five_eq.h"])) | set(
        declared_functions(["godunov_five_eq_io.h"]))
    L = g5.lib()
    declared = declared_functions(["godunov_five_eq.h"])
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(L, name), name
    assert set(g5.exported_symbols()) <= set(declared)
    assert b"gfx950" in L.g5_version()
    H = hdf5io.lib()
    declared_io = declared_functions(["godunov_five_eq_io.h"])
    assert len(declared_io) == 4
    for name in declared_io:
        assert hasattr(H, name), name
    assert set(hdf5io.exported_symbols()) == set(declared_io)


def test_invalid_configs_are_rejected(g5):
    bad = [
        dict(dim=4),
        dict(bsize=8),
        dict(bc=("NONE",) * 6),
        dict(gamma=(1.0, 1.4)),
    ]
    for kw in bad:
        base = dict(dim=2, nb=(2, 2), dx=0.1)
        base.update(kw)
        dim = base.pop("dim")
        nb = base.pop("nb")
        dx = base.pop("dx")
        cfg = g5.make_config(dim, nb, dx, **base)
        h = ctypes.c_void_p()
        rc = g5.lib().g5_create(ctypes.byref(cfg), ctypes.byref(h))
        assert rc in (g5.G5_ERR_INVALID, g5.G5_ERR_UNSUPPORTED), kw
        assert h.value is None
        assert len(g5.lib().g5_last_error(None)) > 0


def test_null_arguments(g5):
    L = g5.lib()
    assert L.g5_create(None, None) == g5.G5_ERR_INVALID
    assert L.g5_destroy(None) == g5.G5_OK
    assert L.g5_shape(None, None, None, None) == g5.G5_ERR_INVALID


def test_initial_condition_host_side(g5):
    import cases
    cfg, U0 = cases.air_water_3d(nb=2)
    assert U0.size == g5.state_size(cfg)
    V = U0.reshape(-1, 9, 4096)
    assert np.all(np.isfinite(V))
    assert V[:, 5].min() == 0.0 and V[:, 5].max() == 1.0  # air fraction
    assert np.allclose(V[:, 0], V[:, 6] + V[:, 8])


def test_triple_point_3d_host_geometry(g5):
    # 3-D triple point: region 1 is the cylinder sqrt(y^2 + z^2) <= yd right of xd
    # (InitTriplePoint.cpp:50-60), region 2 the rest right of xd; restated at the cell centres
    cfg = g5.make_config(3, (3, 2, 2), 1.0 / 16, gamma=(1.5, 1.4))
    cfg.xmin[1] = -1.0
    regs = [g5.region(1.0, (1.0, 0.0), 1.0), g5.region(0.0, (0.0, 1.0), 0.1), g5.region(1.0, (0.125, 0.0), 0.1)]
    U = g5.init_triple_point(cfg, regs, 1.0, 0.75).reshape(2, 2, 3, 9, 16, 16, 16)  # bz, by, bx, var, cz, cy, cx
    c = np.arange(16) + 0.5
    x = (np.arange(3)[:, None] * 16 + c[None, :]) / 16                 # (bx, cx)
    y = -1.0 + (np.arange(2)[:, None] * 16 + c[None, :]) / 16          # (by, cy)
    z = (np.arange(2)[:, None] * 16 + c[None, :]) / 16                 # (bz, cz)
    X = x[None, None, :, None, None, :]
    Y = y[None, :, None, None, :, None]
    Z = z[:, None, None, :, None, None]
    region = np.where(X <= 1.0, 0, np.where(np.sqrt(Y * Y + Z * Z) <= 0.75, 1, 2))
    region = np.broadcast_to(region, (2, 2, 3, 16, 16, 16))
    S = [g5.region_conservative(cfg, r) for r in regs]
    for v in range(9):
        assert np.array_equal(U[:, :, :, v], np.choose(region, [S[0][v], S[1][v], S[2][v]])), v
    assert (region == 1).sum() > 0 and (region == 2).sum() > 0


def _block_range(w, nblocks, wg_per_xcd, remap, grid):
    """Host restatement of the step kernels' block_range (blocks handled by workgroup w)."""
    per_xcd = (nblocks + 7) // 8
    if wg_per_xcd == 0:
        b = (w & 7) * per_xcd + (w >> 3) if remap else w
        return list(range(b, min(b + 1, nblocks)))
    if remap:
        x = w & 7
        b = x * per_xcd + (w >> 3)
        return list(range(b, min((x + 1) * per_xcd, nblocks), wg_per_xcd))
    return list(range(w, nblocks, grid))


@pytest.mark.parametrize("nblocks", [1, 7, 27, 256, 1920, 4097])
def test_launch_geometries_cover_every_block_once(nblocks):
    per_xcd = (nblocks + 7) // 8
    for wg_per_xcd, remap in [(0, 1), (0, 0), (32, 1), (64, 1), (128, 0), (3, 1)]:
        grid = 8 * wg_per_xcd if wg_per_xcd else 8 * per_xcd
        seen = []
        for w in range(grid):
            seen += _block_range(w, nblocks, wg_per_xcd, remap, grid)
        assert sorted(seen) == list(range(nblocks)), (nblocks, wg_per_xcd, remap)


def test_checkpoint_header_round_trip(g5, tmp_path):
    # the JSON header keeps every configuration field exactly (host side only)
    from godunov_five_eq_amd import checkpoint

    cfg = g5.make_config(3, (3, 4, 5), 1.0 / 48, thinc=True, thinc_beta=1.7, gamma=(1.4, 4.4), pinf=(0.0, 6e8),
                         precision=g5.PREC_MIXED, integrator="RK1", bc=("WALL",) * 6)
    d = checkpoint._cfg_to_dict(cfg)
    import json

    back = checkpoint._cfg_from_dict(json.loads(json.dumps(d)))
    assert bytes(back) == bytes(cfg)


def test_two_level_mesh_refusals(g5):
    """g5_create_two_level refuses, before allocating anything, what the two-level kernels do not
    cover: the mixed precision (its face records are star states, not summable fluxes), refined
    blocks on the brick's border in 2-D or 3-D (their fine ghosts would need border conditions at
    level 1), repeated blocks and unknown prolongations."""
    L = g5.lib()
    cases2 = [
        (g5.make_config(3, (4, 4, 4), 0.1), [[1, 1, 0]], 0, g5.G5_ERR_INVALID),
        (g5.make_config(2, (4, 4), 0.1, precision=g5.PREC_MIXED), [[1, 1]], 0, g5.G5_ERR_UNSUPPORTED),
        (g5.make_config(2, (4, 4), 0.1), [[0, 1]], 0, g5.G5_ERR_INVALID),
        (g5.make_config(2, (4, 4), 0.1), [[1, 3]], 0, g5.G5_ERR_INVALID),
        (g5.make_config(2, (4, 4), 0.1), [[1, 1], [1, 1]], 0, g5.G5_ERR_INVALID),
        (g5.make_config(2, (4, 4), 0.1), [[1, 1]], 4, g5.G5_ERR_INVALID),
        (g5.make_config(2, (4, 4), 0.1), [], 0, g5.G5_ERR_INVALID),
    ]
    for cfg, refined, prolong, code in cases2:
        r = np.ascontiguousarray(np.array(refined, dtype=np.int32).reshape(-1, cfg.dim))
        h = ctypes.c_void_p()
        rc = L.g5_create_two_level(ctypes.byref(cfg), r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                   r.shape[0], prolong, ctypes.byref(h))
        assert rc == code and h.value is None, (refined, prolong, rc)
        assert len(L.g5_last_error(None)) > 0
