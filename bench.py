"""Benchmark: cell-updates/s of the fused HIP step on one MI355X.

Workloads (--workload):
  air_water_3d  (default) 3-D air-water shock-bubble with THINC at 256^3 cells (4096 blocks of
                16^3), FP64, HANCOCK, HLLC, CFL 0.5, 50 steps: the plain FP64 THINC configuration.
                The default arithmetic is the reassociated FP64 path (--precision fp64_fast:
                relative L1 <= 1e-12 per conserved variable against the reference's operation
                order); the line also carries the bit-exact path's rate on the same steps and the
                relative L1 between the two final states (--precision fp64 times the bit-exact
                path alone).
  air_helium_2d 2-D air-helium shock-bubble, 512x128 cells, FP64, minmod, CFL 0.8, 200 steps (the
                reference's CPU-runnable case; the state fits in the last-level cache).
  mixed_3d      the air-water THINC case at 512x256x256 cells in mixed precision, 100 steps, with
                the per-field relative L1 error against the FP64 run of the same case.
  batch         4096 independent 2-D solves of the seeded batching workload (g5_batch_spec),
                mixed precision, one pool, 100 steps each.
  two_level_2d  the 2-D air-helium shock-bubble on a static two-level mesh: a 2048x512-cell brick
                (128x32 blocks, dx = 1/512) whose 24x24 blocks around the bubble are refined into
                96x96 blocks of half the cell length (5824 blocks, 1.49 M cells), reassociated
                FP64, 100 steps; the one-level brick of the same coarse size timed beside it.
  serve         the native serving loop (g5_serve): 400 requests with Poisson arrivals at 160/s
                (about 80% of the loop's capacity), 2-D and 10% 3-D requests, HBM budget 200 GB,
                mixed precision; reports p50/p99 latency.

A step is one godunov_unsplit iteration over every solve (device-resident dt, fused step kernel
+ dt finalise, one HIP graph).  value = owned cells * steps / wall time of the timed loop, state
resident in HBM.  roofline: algorithmic bytes per cell-update = 2 * nvar * 8 B (the FP64 state
read and written once: 128 B in 2-D, 144 B in 3-D; the mixed form keeps the FP64 state), divided
by the HIP-event time per step of the timed loop.  cpu_baseline: the CPU FP64 oracle (a
restatement of the reference V0 step) on a bounded sample of the same workload, on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s on 1 GPU vs plain FP64 at fixed L1 err; p50/p99 solve latency"
SEED = 20251203
DEFAULT_STEPS = {"air_water_3d": 50, "air_helium_2d": 200, "mixed_3d": 100, "batch": 100, "serve": 0, "two_level_2d": 100}
# arithmetic of the single-solve FP64 workloads: the reference's operation order (bit-exact) or the
# reassociated FP64 form (relative L1 <= 1e-12 per conserved variable against it)
PRECISIONS = {"fp64": 0, "fp64_fast": 2}


def traffic_from_profile(name):
    """HBM bytes per step-kernel launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE,
    scaled by a copy-kernel calibration); None when no measurement is committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def cpu_rate(cfg, U0, nsteps):
    """Cell-updates/s of the CPU FP64 oracle on one configuration (bounded sample)."""
    import oracle

    o = oracle.Oracle(cfg)
    o.set_state(U0)
    t0 = time.perf_counter()
    _, n = o.run(nsteps)
    dt = time.perf_counter() - t0
    return o.nowned * o.ncell * n, dt, oracle.max_threads()


def roofline(cells_per_step, nvar, ms_per_step, traffic):
    b_alg = 2 * nvar * 8 * cells_per_step
    achieved = b_alg / (ms_per_step * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic}


def run_context(g5, cfg, U0, steps, warmup):
    """Warm-up, then `steps` timed steps bracketed by synchronisations; returns
    (wall s, HIP-event ms per step, state, steps done)."""
    ctx = g5.Context(cfg)
    ctx.upload(U0)
    if warmup > 0:
        ctx.run(warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    _, n = ctx.run(steps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ctx.last_run_ms() / max(n, 1)
    U = ctx.download()
    ctx.close()
    return wall, ev_ms, U, n


def l1_vs(U, Uref, nvar, ncell, dim):
    """per-variable relative L1 of U against Uref; the momentum components against the momentum
    vector's L1 norm (a transverse component can be round-off by itself)"""
    A = U.reshape(-1, nvar, ncell)
    B = Uref.reshape(-1, nvar, ncell)
    mom = list(range(2, 2 + dim))
    mden = sum(float(np.abs(B[:, m]).sum()) for m in mom)
    out = []
    for v in range(nvar):
        den = mden if v in mom else float(np.abs(B[:, v]).sum())
        out.append(float(np.abs(A[:, v] - B[:, v]).sum()) / max(den, 1e-300))
    return out


def bench_single(g5, args, name):
    import cases

    if name == "air_water_3d":
        cfg, U0 = cases.air_water_3d(nb=args.nb, thinc=True)
        grid = [16 * args.nb] * 3
        wl = "shock_bubble_3d_air_water_thinc"
    else:
        cfg, U0 = cases.air_helium_2d()
        grid = [512, 128]
        wl = "shock_bubble_2d_air_helium"
    nvar = g5.nvar_of(cfg.dim)
    cells = int(np.prod(grid))
    cfg.precision = PRECISIONS[args.precision]
    wall, ev_ms, U, n = run_context(g5, cfg, U0, args.steps, args.warmup)
    assert n == args.steps and np.all(np.isfinite(U)), "non-finite state"
    out = base_line(args, cells * n / wall, wall * 1e3 / n, n, "f64")
    out["config"] = {"workload": wl, "cells": cells, "grid": grid, "blocks": int(np.prod(cfg.nb[:cfg.dim])),
                     "block": [16] * cfg.dim, "precision": args.precision, "integrator": "HANCOCK", "riemann": "HLLC",
                     "thinc": bool(cfg.thinc_enabled), "cfl": cfg.cfl}
    suffix = "" if args.precision == "fp64" else "_" + args.precision
    out["roofline"] = roofline(cells, nvar, ev_ms, traffic_from_profile(
        (f"step3d_air_water_thinc_{grid[0]}" if cfg.dim == 3 else "step2d_air_helium_512x128") + suffix))
    if args.precision != "fp64" and not args.no_exact:
        # the bit-exact path on the same steps: its rate, and the distance of the two final states
        cfg.precision = PRECISIONS["fp64"]
        wall_x, ev_x, U_x, n_x = run_context(g5, cfg, U0, args.steps, args.warmup)
        l1 = l1_vs(U, U_x, nvar, 16 ** cfg.dim, cfg.dim)
        out["fp64_exact"] = {"value": cells * n_x / wall_x, "ms_per_step": wall_x * 1e3 / n_x,
                             "kernel_ms_per_step": ev_x}
        out["config"]["rel_l1_vs_fp64_exact"] = {"max": max(l1), "per_variable": l1,
                                                 "momentum_norm": "momentum components against the "
                                                                  "momentum vector's L1 norm"}
    if not args.no_cpu:
        k = args.cpu_steps if cfg.dim == 3 else 200
        cu, dt, th = cpu_rate(cfg, U0, k)
        out["cpu_baseline"] = {"value": cu / dt, "unit": "cell-updates/s", "cores": th, "kind": "port",
                               "sample": f"{k} steps of the same {'x'.join(map(str, grid))} workload, CPU FP64 "
                                         f"oracle ({th} OpenMP threads, {dt:.1f} s)"}
    return out


def bench_two_level(g5, args):
    import cases

    nbx, nby = 128, 32
    refined = [[x, y] for y in range(4, 28) for x in range(20, 44)]
    cfg, Uc = cases.air_helium_2d(nbx=nbx, nby=nby)
    _, Uf = cases.air_helium_2d(nbx=2 * nbx, nby=2 * nby)
    cfg.precision = PRECISIONS[args.precision]
    ctx = g5.Context(cfg, refined=refined, prolongation=g5.PROLONG_EXTRAPOLATE_LINEAR_MINMOD)
    lay = ctx.block_layout()
    Uc3, Uf3 = Uc.reshape(-1, 8, 256), Uf.reshape(-1, 8, 256)
    U0 = np.stack([Uc3[x + nbx * y] if lv == 0 else Uf3[x + 2 * nbx * y] for x, y, _, lv in lay]).ravel()
    ctx.upload(U0)
    if args.warmup > 0:
        ctx.run(args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    _, n = ctx.run(args.steps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ctx.last_run_ms() / max(n, 1)
    U = ctx.download()
    ctx.close()
    assert n == args.steps and np.all(np.isfinite(U)), "non-finite state"
    cells = len(lay) * 256
    nfine = int((lay[:, 3] == 1).sum())
    # the one-level brick of the coarse size, same arithmetic and steps
    wall1, ev1, _, n1 = run_context(g5, cfg, Uc, args.steps, args.warmup)
    out = base_line(args, cells * n / wall, wall * 1e3 / n, n, "f64")
    out["config"] = {"workload": "shock_bubble_2d_air_helium_two_level", "coarse_grid": [16 * nbx, 16 * nby],
                     "refined_coarse_blocks": len(refined), "blocks": len(lay), "fine_blocks": nfine, "cells": cells,
                     "prolongation": "EXTRAPOLATE_LINEAR_MINMOD", "precision": args.precision,
                     "one_level_same_brick": {"cells": nbx * nby * 256, "value": nbx * nby * 256 * n1 / wall1,
                                              "ms_per_step": wall1 * 1e3 / n1, "kernel_ms_per_step": ev1}}
    out["roofline"] = roofline(cells, 8, ev_ms, None)
    return out


def bench_mixed3d(g5, args):
    import cases

    nbx = 2 * args.nb
    cfg64, U0 = cases.air_water_3d(nb=args.nb, thinc=True)
    # 512 x 256 x 256: brick 2x1x1 of the unit-cube case, same dx
    cfg64.nb[0] = nbx
    U0 = g5.init_shock_bubble(cfg64, air_water_regions(g5), 0.25, (0.5, 0.5, 0.5), 0.2)
    cfgm = g5.G5Config.from_buffer_copy(cfg64)
    cfgm.precision = g5.PREC_MIXED
    cells = nbx * args.nb * args.nb * 4096
    wall64, ev64, U64, n64 = run_context(g5, cfg64, U0, args.steps, args.warmup)
    wall, ev, Um, n = run_context(g5, cfgm, U0, args.steps, args.warmup)
    assert n == n64 == args.steps and np.all(np.isfinite(Um)) and np.all(np.isfinite(U64))
    # both runs include the warm-up steps: the states compared are after warmup + steps steps
    l1 = l1_vs(Um, U64, 9, 4096, 3)
    names = ["rho", "E", "rho_u", "rho_v", "rho_w", "alpha0", "alpha_rho0", "alpha1", "alpha_rho1"]
    out = base_line(args, cells * n / wall, wall * 1e3 / n, n, "mixed-f32-f64")
    out["config"] = {"workload": "shock_bubble_3d_air_water_thinc_mixed", "cells": cells,
                     "grid": [16 * nbx, 16 * args.nb, 16 * args.nb], "blocks": nbx * args.nb ** 2,
                     "precision": "mixed", "l1_vs_fp64_after_steps": args.steps + args.warmup,
                     "rel_l1_vs_fp64": dict(zip(names, l1)),
                     "momentum_norm": "momentum components against the momentum vector's L1 norm",
                     "fp64_value": cells * n64 / wall64, "speedup_vs_fp64": wall64 / wall}
    out["roofline"] = roofline(cells, 9, ev, traffic_from_profile(f"step3d_mixed_air_water_thinc_{16 * nbx}"))
    if not args.no_cpu:
        k = max(1, args.cpu_steps // 2)
        cu, dt, th = cpu_rate(cfg64, U0, k)
        out["cpu_baseline"] = {"value": cu / dt, "unit": "cell-updates/s", "cores": th, "kind": "port",
                               "sample": f"{k} steps of the same 512x256x256 case, CPU FP64 oracle "
                                         f"({th} OpenMP threads, {dt:.1f} s)"}
    return out


def air_water_regions(g5):
    rho2, u2, p2, _ = g5.rankine_hugoniot(4.4, 6e8, 1000.0, 1e5, p2=1e9)
    return [g5.region(0.0, (0.0, rho2), p2, u2), g5.region(0.0, (0.0, 1000.0), 1e5),
            g5.region(1.0, (1.0, 0.0), 1e5)]


def bench_batch(g5, args):
    prec = g5.PREC_FP64 if args.fp64 else g5.PREC_MIXED
    n = args.nsolves
    specs = [g5.batch_spec(SEED, k) for k in range(n)]
    cells = sum(256 * s.nb[0] * s.nb[1] for s in specs)
    bytes_needed = cells * 8 * 8 * 2
    budget = int(bytes_needed * 1.02) + (64 << 20)
    steps = args.steps
    pool = g5.Pool(2, thinc=False, precision=prec, budget_bytes=budget)
    t_adm = time.perf_counter()
    ids = [pool.admit_spec(s) for s in specs]
    t_adm = time.perf_counter() - t_adm
    pool.start(ids)
    # launch list, autotuning of the launch geometry and step-graph capture before the timed loop
    pool.prepare()
    if args.warmup > 0:
        pool.step(args.warmup)
    before = [pool.progress(i)[1] for i in ids]  # synchronises
    t0 = time.perf_counter()
    pool.step(steps)
    pool.progress(ids[0])  # synchronises
    wall = time.perf_counter() - t0
    ev_ms = pool.last_step_ms() / steps
    launch = pool.launch()
    # a solve whose state leaves the mixture EOS's domain (a cell with c^2 < 0, i.e. a negative
    # pressure) is stopped there in every precision: count the cell-updates done, per solve
    after = [pool.progress(i) for i in ids]
    failed = [k for k, a in enumerate(after) if a[3]]
    done_updates = sum(256 * s.nb[0] * s.nb[1] * (a[1] - b) for s, a, b in zip(specs, after, before))
    # every healthy solve advanced by warmup + steps (each has a 100-step budget: the timed loop
    # ends exactly at the budget with the default warm-up of 0)
    sample = [k for k in range(0, n, max(1, n // 64)) if k not in failed]
    results = {k: pool.download(ids[k]) for k in sample}
    finite = all(np.isfinite(u).all() for u in results.values())
    stepped = [after[k][1] for k in range(n) if k not in failed]
    pool.close()
    assert finite
    l1 = None
    if prec == g5.PREC_MIXED:
        # accuracy at this speed: the sampled solves rerun in FP64 (bit-identical to the oracle)
        num = np.zeros(8)
        den = np.zeros(8)
        with g5.Pool(2, precision=g5.PREC_FP64, budget_bytes=16 << 30) as p64:
            for k in sample:
                sid = p64.admit_spec(specs[k])
                p64.start([sid])
                p64.step(args.warmup + steps)
                ref = p64.download(sid).reshape(-1, 8, 256)
                p64.release(sid)
                got = results[k].reshape(-1, 8, 256)
                num += np.abs(got - ref).sum(axis=(0, 2))
                den += np.abs(ref).sum(axis=(0, 2))
        mom = den[2] + den[3]
        den[2] = den[3] = mom
        names = ["rho", "E", "rho_u", "rho_v", "alpha0", "alpha_rho0", "alpha1", "alpha_rho1"]
        l1 = dict(zip(names, (num / den).tolist()))
    out = base_line(args, done_updates / wall, wall * 1e3 / steps, steps, "mixed-f32-f64" if prec else "f64")
    out["config"] = {"workload": "batch_2d_seeded", "seed": SEED, "solves": n, "cells": cells,
                     "precision": "mixed" if prec else "fp64", "block": [16, 16], "blocks": cells // 256,
                     "steps_per_solve": min(stepped), "cell_updates": done_updates,
                     "failed_solves": len(failed), "failed_first": failed[:8],
                     "admission_s": t_adm, "launch": launch,
                     "hip_graph": True, "rel_l1_vs_fp64_sample": l1, "l1_sample_solves": len(sample)}
    out["roofline"] = roofline(done_updates / steps, 8, ev_ms,
                               traffic_from_profile("step2d_batch_mixed" if prec else "step2d_batch_fp64"))
    if not args.no_cpu:
        # bounded sample: the first solves of the same batch (device-generated initial states)
        cu = dt = 0.0
        th = 1
        with g5.Pool(2, precision=g5.PREC_FP64, budget_bytes=1 << 30) as p64:
            k = 0
            while dt < 10.0 and k < n:
                sid = p64.try_admit_spec(specs[k])
                if sid is None:
                    break
                U0 = p64.download(sid)
                c, d, th = cpu_rate(g5.spec_config(specs[k]), U0, 10)
                p64.release(sid)
                cu += c
                dt += d
                k += 1
        out["cpu_baseline"] = {"value": cu / dt, "unit": "cell-updates/s", "cores": th, "kind": "port",
                               "sample": f"10 steps of each of the first {k} solves of the same batch, CPU "
                                         f"FP64 oracle ({th} OpenMP threads, {dt:.1f} s)"}
    return out


def bench_serve(g5, args):
    prec = g5.PREC_FP64 if args.fp64 else g5.PREC_MIXED
    rep, lat = g5.serve(SEED, args.requests, args.rate, frac_3d=0.1, share_3d=0.3, budget_bytes=200 << 30,
                        precision=prec, steps_per_poll=args.poll)
    # a request whose state leaves the mixture EOS's domain (c^2 < 0) is stopped and counted as failed
    assert rep["completed"] + rep["failed"] == args.requests and rep["rejected"] == 0, rep
    out = base_line(args, rep["cell_updates_per_s"], None, None, "mixed-f32-f64" if prec else "f64")
    out["higher_is_better"] = True
    out["config"] = {"workload": "serve_poisson", "seed": SEED, "requests": args.requests, "rate_hz": args.rate,
                     "frac_3d": 0.1, "hbm_budget_gb": 200, "precision": "mixed" if prec else "fp64",
                     "steps_per_poll": args.poll, "p50_ms": rep["p50_ms"], "p99_ms": rep["p99_ms"],
                     "mean_ms": rep["mean_ms"], "max_ms": rep["max_ms"], "wall_s": rep["wall_s"],
                     "max_inflight": rep["max_inflight"], "cell_updates": rep["cell_updates"],
                     "completed": rep["completed"], "failed": rep["failed"],
                     "last_arrival_s": rep["last_arrival_s"],
                     # where the loop's time goes: HIP-event time of each pool's steps, host time
                     # per loop phase
                     "step_s_2d": rep["step_s_2d"], "step_s_3d": rep["step_s_3d"], "admit_s": rep["admit_s"],
                     "start_s": rep["start_s"], "wait_s": rep["wait_s"], "collect_s": rep["collect_s"]}
    out["roofline"] = None
    return out


def base_line(args, value, ms_per_step, steps, dtype):
    return {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="air_water_3d", choices=sorted(DEFAULT_STEPS))
    ap.add_argument("--nb", type=int, default=16, help="blocks per direction of the 3-D cases (16 -> 256)")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-exact", action="store_true",
                    help="fp64_fast: skip the bit-exact companion run (profiling runs of the fast kernel alone)")
    ap.add_argument("--fp64", action="store_true", help="batch/serve: FP64 instead of mixed precision")
    ap.add_argument("--precision", default="fp64_fast", choices=sorted(PRECISIONS),
                    help="air_water_3d / air_helium_2d: reassociated FP64 (default) or bit-exact FP64")
    ap.add_argument("--nsolves", type=int, default=4096)
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--rate", type=float, default=160.0)
    ap.add_argument("--poll", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the step runs on one GPU; sharding blocks over several GPUs (one halo "
              "exchange and one dt min-reduction per step) is not built", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = DEFAULT_STEPS[args.workload]
    if args.warmup is None:
        args.warmup = 0 if args.workload == "batch" else 5

    import godunov_five_eq_amd as g5

    if args.workload in ("air_water_3d", "air_helium_2d"):
        out = bench_single(g5, args, args.workload)
    elif args.workload == "mixed_3d":
        out = bench_mixed3d(g5, args)
    elif args.workload == "batch":
        out = bench_batch(g5, args)
    elif args.workload == "two_level_2d":
        out = bench_two_level(g5, args)
    else:
        out = bench_serve(g5, args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
