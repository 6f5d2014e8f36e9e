"""Benchmark: cell-updates/s of the fused FP64 HIP step on one MI355X.

Workload (default): 3-D air-water shock-bubble with THINC at 256^3 cells (4096 blocks of 16^3),
FP64, HANCOCK, HLLC, CFL 0.5 -- the plain FP64 THINC configuration of the reference's metric.
A step is one godunov_unsplit iteration (device-resident dt, fused step kernel + dt finalise).
value = owned cells * steps / wall time of the timed loop (state resident in HBM).

roofline: algorithmic bytes per cell-update = 2 * nvar * 8 B (read and write the FP64 state
once; 144 B in 3-D), divided by the HIP-event duration of the step loop per step.
cpu_baseline: the CPU FP64 oracle (a restatement of the reference V0 step) on a bounded sample of
the same workload, on the host cores of this machine.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--nb 16] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]

HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s on 1 GPU vs plain FP64 at fixed L1 err; p50/p99 solve latency"


def traffic_from_profile(nb):
    """HBM bytes per step-kernel launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE,
    scaled by a copy-kernel calibration); None when no measurement for this size is committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_step3d_air_water_thinc_{16 * nb}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def cpu_baseline(cfg, U0, nsteps):
    import oracle

    o = oracle.Oracle(cfg)
    o.set_state(U0)
    t0 = time.perf_counter()
    _, n = o.run(nsteps)
    dt = time.perf_counter() - t0
    cells = o.nowned * o.ncell
    return cells * n / dt, oracle.max_threads()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nb", type=int, default=16, help="blocks per direction (16 -> 256^3 cells)")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are out of scope "
              "(the halo-exchange design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import cases
    import godunov_five_eq_amd as g5

    cfg, U0 = cases.air_water_3d(nb=args.nb, thinc=True)
    cells = (16 * args.nb) ** 3
    ctx = g5.Context(cfg)
    ctx.upload(U0)
    if args.warmup > 0:
        ctx.run(args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    _, n = ctx.run(args.steps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ctx.last_run_ms() / max(n, 1)
    U = ctx.download()
    ctx.close()
    import numpy as np

    assert n == args.steps and np.all(np.isfinite(U)), "non-finite state"
    value = cells * n / wall
    b_alg = 2 * 9 * 8 * cells  # bytes per launch
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / n,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "shock_bubble_3d_air_water_thinc",
            "cells": cells,
            "grid": [16 * args.nb] * 3,
            "blocks": args.nb ** 3,
            "block": [16, 16, 16],
            "precision": "fp64",
            "integrator": "HANCOCK",
            "riemann": "HLLC",
            "thinc_beta": 1.6,
            "cfl": 0.5,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profile(args.nb),
        },
    }
    if not args.no_cpu:
        rate, threads = cpu_baseline(cfg, U0, args.cpu_steps)
        out["cpu_baseline"] = {
            "value": rate,
            "unit": "cell-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{args.cpu_steps} steps of the same {16 * args.nb}^3 workload, CPU FP64 oracle "
                      f"(OpenMP, {threads} threads)",
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
