"""Launch-shape autotuner of the MPC-MD library.

Two knobs decide how the solvent kernels map onto the MI355X: particles per
thread of k_bin / k_perm (1, 2, 4, 8 -- fewer, fatter workgroups hide the
atomic round trip of the cell counts better, more workgroups fill the 256
CUs sooner) and cells per k_collide tile (64, 128, 256 -- the tile's
particles stay in registers between the three passes).  The best pair
depends on the box sizes, the density and the number of replicas, so it is
measured: every candidate runs a few whole collision intervals of the real
system (advancing it, like warm-up steps) and the fastest is kept.  Results
are cached per workload signature in a JSON file.

    from starpolymersmd_amd import MPCSystem, make_params
    from starpolymersmd_amd.autotune import autotune
    g = MPCSystem(make_params(L=(64, 64, 64), n_c=10, step_update=10))
    ...
    best = autotune(g, t0=1, mpc_dt=0.1)      # returns the chosen tuning and t after the sweep
"""
from __future__ import annotations

import json
import os
import time

PPT = (1, 2, 4, 8)
TILES = (64, 128, 256)
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "starpolymersmd_amd", "tuning.json")


def signature(g) -> str:
    """Workload key: per-replica box, density, monomers; the step cadence."""
    parts = []
    for r, p in enumerate(g.params):
        parts.append(f"{p.L[0]}x{p.L[1]}x{p.L[2]}:nc{p.n_c}:m{g.num_monomers(r)}")
    return f"S{g.params[0].step_update}|" + ",".join(parts)


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def autotune(g, t0: int, mpc_dt: float, intervals: int = 10, warm: int = 2, cache: str | None = DEFAULT_CACHE,
             candidates=None) -> dict:
    """Time every (ppt, tile_cells) candidate on `g` starting at MPC::step
    index t0 and keep the fastest.  Returns {"ppt", "tile_cells", "ms",
    "timings", "t"} where t is the step index after the sweep (the sweep
    advances the system by whole collision intervals)."""
    S = int(g.params[0].step_update)
    key = signature(g)
    if cache:
        hit = _load(cache).get(key)
        if hit:
            g.set_tuning(hit["ppt"], hit["tile_cells"])
            return dict(hit, t=t0, cached=True)
    t = t0
    # finish the current interval so every timed block starts at a collision boundary
    if t % S != 1 % S:
        n = (1 - t) % S
        g.run(t, n, mpc_dt)
        t += n
    timings = {}
    for ppt, tc in candidates or [(p, c) for p in PPT for c in TILES]:
        g.set_tuning(ppt, tc)
        g.run(t, warm * S, mpc_dt)
        t += warm * S
        g.sync()
        t_a = time.perf_counter()
        g.run(t, intervals * S, mpc_dt)
        g.sync()
        ms = (time.perf_counter() - t_a) * 1e3 / intervals
        t += intervals * S
        timings[f"{ppt}/{tc}"] = ms
    best = min(timings, key=timings.get)
    ppt, tc = (int(v) for v in best.split("/"))
    g.set_tuning(ppt, tc)
    out = {"ppt": ppt, "tile_cells": tc, "ms": timings[best], "timings": timings}
    if cache:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        db = _load(cache)
        db[key] = out
        with open(cache + ".tmp", "w") as f:
            json.dump(db, f, indent=1)
        os.replace(cache + ".tmp", cache)
    return dict(out, t=t, cached=False)
