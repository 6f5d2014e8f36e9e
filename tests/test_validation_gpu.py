"""Physics validation against the only measured numbers the reference holds: the wave-front
positions of its air-helium shock-bubble run (app/test_shock_bubble_air_helium_2d/plot.py:21-74,
case test_shock_bubble_air_helium_2d_measure.ini; fixture tests/golden/shock_bubble_air_helium_fronts.json).

The same case on a fixed mesh of 16x16 blocks at level 4 (dx = 0.35 mm, 1280 x 256 cells) runs to
each measurement time (0.06 ... 0.30 ms) through the device-dt loop; the incident shock, the
refracted and transmitted shocks, the upstream ("jet") and downstream interfaces are located as
in cases.wave_fronts.  Bar: every measured position within 1.2 mm (3.5 cells; the reference's own
points scatter by up to 0.8 mm about their linear fits), each front's mean deviation within 0.6 mm.
"""
import json
import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shock_bubble_air_helium_fronts.json")


def test_shock_bubble_wave_fronts_match_reference_measurements(g5):
    ref = json.load(open(GOLD))
    cfg, U0 = cases.haas_sturtevant(level=4)
    print(f"[gpu] air-helium shock-bubble fronts: level 4, {cfg.nb[0] * 16}x{cfg.nb[1] * 16} cells", flush=True)
    times = sorted({round(t, 3) for k in ref if isinstance(ref[k], dict) for t in ref[k]["t"]})
    fronts = {}
    t = 0.0
    with g5.Context(cfg) as ctx:
        ctx.upload(U0)
        for tk in times:
            t, _ = ctx.run(2 ** 31 - 1, t, tk)
            assert t == tk
            fronts[tk] = cases.wave_fronts(cfg, ctx.download())
    for name in ("shock", "refracted", "jet", "downstream_interface", "transmitted"):
        d = ref[name]
        err = np.array([fronts[round(tk, 3)][name] - xk for tk, xk in zip(d["t"], d["x"])])
        print(f"[gpu]   {name:22s} mean {err.mean():+.2f} max |err| {np.abs(err).max():.2f} mm", flush=True)
        assert np.all(np.isfinite(err)), name
        assert np.abs(err).max() <= 1.2, (name, err)
        assert abs(err.mean()) <= 0.6, (name, err)
