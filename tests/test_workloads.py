"""Seeded workload generators of the batching and serving forms (host side, no GPU): determinism,
the documented parameter ranges, the estimated size of the 4096-solve batch, Poisson arrivals,
the configuration/region mapping, and that downsized solves of the batch stay finite under the
oracle (the random physics range does not break the scheme)."""
import numpy as np
import pytest

SEED = 20251203


def test_batch_spec_is_deterministic(g5):
    a = g5.batch_spec(SEED, 17)
    b = g5.batch_spec(SEED, 17)
    c = g5.batch_spec(SEED + 1, 17)
    assert bytes(a) == bytes(b)
    assert bytes(a) != bytes(c)


def test_batch_spec_ranges_and_size(g5):
    n = 4096
    specs = [g5.batch_spec(SEED, k) for k in range(n)]
    cells = 0
    kinds = np.zeros(2)
    shapes = np.zeros(2)
    sides = {}
    water = 0
    for s in specs:
        assert s.dim == 2 and s.steps == 100
        side, lng = 16 * s.nb[1], 16 * s.nb[0]
        assert side in (64, 128, 256, 512)
        assert side <= lng <= 4 * side + 8
        assert 1.1 <= s.gamma[0] <= 4.4 and 1.1 <= s.gamma[1] <= 4.4
        assert s.pinf[0] == 0.0 and s.pinf[1] in (0.0, 6e8)
        assert 1.1 <= s.mach <= 3.0
        water += s.pinf[1] > 0
        kinds[s.kind] += 1
        shapes[s.shape] += 1
        sides[side] = sides.get(side, 0) + 1
        cells += lng * side
    # balanced categories
    assert all(abs(k / n - 0.5) < 0.05 for k in kinds)
    assert all(abs(k / n - 0.5) < 0.05 for k in shapes)
    assert abs(water / n - 0.5) < 0.05
    assert all(abs(v / n - 0.25) < 0.04 for v in sides.values())
    # E[side^2] * E[aspect] * 4096 = 87040 * 2.5 * 4096 = 0.89e9 cells
    assert 0.8e9 < cells < 1.0e9, cells


def test_serve_specs_poisson_and_mix(g5):
    n, rate = 4000, 200.0
    specs = g5.serve_specs(SEED, n, rate, 0.1)
    t = np.array([s.arrival for s in specs])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() * rate - 1.0) < 0.05
    # exponential gaps: coefficient of variation 1
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.08
    dims = np.array([s.dim for s in specs])
    assert abs((dims == 3).mean() - 0.1) < 0.02
    for s in specs:
        assert 50 <= s.steps <= 500
        lim = (64, 128) if s.dim == 3 else (64, 1024)
        for d in range(s.dim):
            assert lim[0] <= 16 * s.nb[d] <= lim[1]


def test_spec_config_and_regions(g5):
    for k in range(64):
        s = g5.batch_spec(SEED, k)
        cfg = g5.spec_config(s, thinc=True, precision=g5.PREC_MIXED)
        assert cfg.dim == 2 and cfg.thinc_enabled == 1 and cfg.precision == g5.PREC_MIXED
        assert cfg.dx == pytest.approx(1.0 / (16 * s.nb[1]))
        assert cfg.gamma[1] == s.gamma[1] and cfg.pinf[1] == s.pinf[1]
        names = ("NONE", "PERIODIC", "ZERO_GRADIENT", "WALL")
        bc = [names[b] for b in cfg.bc[:4]]
        if s.kind == 1:
            assert bc == ["ZERO_GRADIENT"] * 4 and cfg.slope_type == 2.0
        else:
            assert bc == ["ZERO_GRADIENT", "ZERO_GRADIENT", "WALL", "WALL"] and cfg.slope_type == 1.0
        r = g5.spec_regions(s)
        # post-shock state: Rankine-Hugoniot of material 1 at the spec's Mach number
        rho2, u2, p2, ws = g5.rankine_hugoniot(s.gamma[1], s.pinf[1], s.rho1, s.p1, mach=s.mach)
        assert (r[0].alpha_rho[1], r[0].u, r[0].p) == (rho2, u2, p2)
        # mass flux through the shock is continuous
        assert s.rho1 * ws == pytest.approx(rho2 * (ws - u2), rel=1e-12)
        assert r[1].alpha_rho[1] == s.rho1 and r[1].p == s.p1 and r[0].alpha0 == 0.0
        assert r[2].alpha0 == 1.0 and r[2].alpha_rho[0] == s.rho0 and r[2].p == s.p1


def test_downsized_batch_solves_stay_finite(g5, orc):
    # the circle shock-bubbles among the first 80 solves of the batch, shrunk to 8x4 blocks,
    # 60 steps of the FP64 oracle
    ran = 0
    for k in range(80):
        s = g5.batch_spec(SEED, k)
        s.nb[0], s.nb[1] = 8, 4
        cfg = g5.spec_config(s)
        r = g5.spec_regions(s)
        if s.kind == 1:
            continue  # the host IC generator covers the shock-bubble family only
        U0 = g5.init_shock_bubble(cfg, r, 0.2, (0.55, 0.5), 0.25) if s.shape == 0 else None
        if U0 is None:
            continue
        o = orc.Oracle(cfg)
        o.set_state(U0)
        o.run(60)
        U = o.get_state()
        assert np.all(np.isfinite(U)), k
        ran += 1
    assert ran >= 12
