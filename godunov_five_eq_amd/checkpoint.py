"""Per-solve checkpoint and restart of a pool solve.  A restored solve continues bit-identically
(g5_pool_set_clock).  This restates, for one solve, what the reference's HDF5 restart does
(SolverGodunovFiveEq.cpp:811-846, init/init.cpp:132-207), in two forms:
  save / load: the state as .npy (U[block][var][cell], FP64) and a JSON header (configuration,
      time, steps done, end time, step budget) -- self-contained;
  save_hdf5 / load_hdf5: the HDF5 + XDMF file of hdf5io (variables by name, run time and
      iteration); the configuration and the limits come from the caller, as the reference takes
      them from the .ini."""
import json

import numpy as np

from . import G5Config

FORMAT = "g5-solve-checkpoint-1"


def _cfg_to_dict(cfg):
    out = {}
    for name, ctype in G5Config._fields_:
        v = getattr(cfg, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


def _cfg_from_dict(d):
    cfg = G5Config()
    for name, _ in G5Config._fields_:
        v = d[name]
        if isinstance(v, list):
            arr = getattr(cfg, name)
            for k, x in enumerate(v):
                arr[k] = x
        else:
            setattr(cfg, name, v)
    return cfg


def save(prefix, pool, sid):
    """Write prefix.npy and prefix.json for solve `sid` of `pool` (waits for its steps)."""
    t, steps, active = pool.status(sid)
    t_end, max_steps = pool.limits(sid)
    np.save(prefix + ".npy", pool.download(sid), allow_pickle=False)
    with open(prefix + ".json", "w") as f:
        json.dump({"format": FORMAT, "layout": "U[block][var][cell], FP64, blocks in brick order",
                   "config": _cfg_to_dict(pool.config(sid)), "t": t, "steps": steps, "active": active,
                   "t_end": t_end, "max_steps": max_steps}, f, indent=1)


def load(prefix, pool, start=True):
    """Admit the saved solve into `pool`, restore its state and clock; start it unless asked not
    to.  Returns the new solve id."""
    with open(prefix + ".json") as f:
        h = json.load(f)
    if h.get("format") != FORMAT:
        raise ValueError(f"{prefix}.json: not a {FORMAT} file")
    cfg = _cfg_from_dict(h["config"])
    U = np.load(prefix + ".npy", allow_pickle=False)
    sid = pool.admit(cfg, h["t_end"], h["max_steps"])
    pool.upload(sid, U)
    pool.set_clock(sid, h["t"], h["steps"])
    if start and h["active"]:
        pool.start([sid])
    return sid


def save_hdf5(directory, basename, pool, sid):
    """Write directory/basename.h5 and .xmf for solve `sid` of `pool` (waits for its steps)."""
    from . import hdf5io
    t, steps, _ = pool.status(sid)
    return hdf5io.write(directory, basename, pool.config(sid), pool.download(sid), t, steps)


def load_hdf5(path, pool, cfg, t_end, max_steps, start=True):
    """Admit a solve of configuration `cfg` into `pool` and restore state, time and step count from
    an HDF5 checkpoint of the same mesh; start it unless asked not to.  Returns the solve id."""
    from . import hdf5io
    U, t, steps = hdf5io.read(path, cfg)
    sid = pool.admit(cfg, t_end, max_steps)
    pool.upload(sid, U)
    pool.set_clock(sid, t, steps)
    if start:
        pool.start([sid])
    return sid
