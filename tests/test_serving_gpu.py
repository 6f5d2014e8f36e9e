"""Batching and serving forms on the GPU.

- Solves of the seeded batch (downsized), admitted with their device-generated initial
  condition into one FP64 pool, match the oracle run from the same initial state to the bit.
- Continuous batching: mixed-precision solves admitted while others are in flight, released, and
  their memory reused give exactly the bits they give alone.
- The native serving loop completes every request and reports consistent latencies.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20251203


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def small_batch_specs(g5, n, nbx=6, nby=4):
    out = []
    k = 0
    while len(out) < n:
        s = g5.batch_spec(SEED, k)
        k += 1
        s.nb[0], s.nb[1] = nbx + (k % 3), nby + (k % 2)
        out.append(s)
    return out


def test_batch_solves_match_oracle(g5, orc):
    specs = small_batch_specs(g5, 8)
    say("batch specs in one FP64 pool vs oracle: " +
        ", ".join(f"kind={s.kind} shape={s.shape} pinf={s.pinf[1]:.0e} M={s.mach:.2f}" for s in specs))
    nsteps = 20
    with g5.Pool(2, budget_bytes=1 << 28) as pool:
        ids = [pool.admit_spec(s) for s in specs]
        U0 = [pool.download(i) for i in ids]
        pool.start(ids)
        pool.step(nsteps)
        for s, sid, u0 in zip(specs, ids, U0):
            t, n, active = pool.status(sid)
            assert n == nsteps
            cfg = g5.spec_config(s)
            o = orc.Oracle(cfg)
            o.set_state(u0)
            t_o, _ = o.run(nsteps)
            assert t == t_o
            assert np.array_equal(pool.download(sid), o.get_state()), (s.kind, s.shape)


def run_alone(g5, spec, nsteps, prec):
    with g5.Pool(2, precision=prec, budget_bytes=1 << 27) as pool:
        sid = pool.admit_spec(spec)
        pool.start([sid])
        pool.step(nsteps)
        return pool.download(sid)


def test_continuous_batching_alone_equals_mix(g5):
    prec = g5.PREC_MIXED
    specs = small_batch_specs(g5, 6)
    for s in specs:
        s.steps = 9
    say("continuous batching (mixed precision): staggered admission, release, reuse")
    alone = [run_alone(g5, s, 9, prec) for s in specs]
    with g5.Pool(2, precision=prec, budget_bytes=1 << 22) as pool:  # about three solves at once
        live = {}
        results = {}
        # two solves start, a third joins after 4 steps, the rest as memory frees up
        for k in (0, 1):
            live[pool.admit_spec(specs[k])] = k
        pool.start(list(live))
        pool.step(4)
        sid = pool.admit_spec(specs[2])
        live[sid] = 2
        pool.start([sid])
        nxt = 3
        while live:
            pool.step(2)
            for sid in pool.finished():
                results[live.pop(sid)] = pool.download(sid)
                pool.release(sid)
            while nxt < len(specs):
                sid = pool.try_admit_spec(specs[nxt])
                if sid is None:
                    break
                live[sid] = nxt
                pool.start([sid])
                nxt += 1
        assert nxt == len(specs)
    for k in range(len(specs)):
        assert np.array_equal(results[k], alone[k]), k


def test_serve_loop_completes(g5):
    n = 24
    say(f"serving loop: {n} requests")
    rep, lat = g5.serve(SEED, n, 2000.0, budget_bytes=8 << 30, steps_per_poll=4)
    say(f"serving report {rep}")
    assert rep["completed"] == n
    assert np.all(lat > 0)
    assert rep["p50_ms"] <= rep["p99_ms"] <= rep["max_ms"]
    assert rep["cell_updates"] > 0 and rep["cell_updates_per_s"] > 0
    specs = g5.serve_specs(SEED, n, 2000.0, 0.1)
    cu = sum(s.steps * 16 ** s.dim * int(np.prod(list(s.nb)[:s.dim])) for s in specs)
    assert rep["cell_updates"] == cu
