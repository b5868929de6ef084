"""Long FP64 oracle runs that give the reference statistics the GPU
statistical tests compare against (test infrastructure; run on the CPU,
several seeds in parallel):

    python tests/golden/make_statistics.py [--workers 7] [--star-steps 1250000]

writes tests/golden/statistics_star_L20.json (the star under shear:
block means of the MPC temperature, energies, patches, gyration tensor and
the VelocityX shear rate) and tests/golden/statistics_fluid_L20.json (the
mpc-test fluid: transverse-mode autocorrelation and viscosity at zero shear,
shear rate and temperature under shear).  The sampling protocol is
tests/stat_protocol.py.  The oracle runs its Philox mode with seeds that
the GPU tests do not use, so the two estimates are independent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from multiprocessing import Pool

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

import stat_protocol as sp  # noqa: E402
from oracle_binding import Oracle  # noqa: E402
from starpolymersmd_amd.observables import FourierAutocorr, VelocityProfile  # noqa: E402

SEED0 = 1000


def star_run(args):
    seed, steps = args
    c = sp.STAR
    o = Oracle(L=(c["L"],) * 3, kT=c["kT"], n_c=c["n_c"], shear=c["shear"], step_update=c["step_update"],
               lam=c["lam"], md_dt=c["md_dt"], seed=seed, replica=0)
    o.initialize()
    o.add_star(c["A"], c["B"], c["arms"], c["mass"])
    o.md_init()
    blocks = sp.StarBlocks()
    n = o.n_mono
    t0 = time.time()
    for t in range(steps):
        if t > sp.STAR_EQUIL and t % sp.STAR_EVERY == 0:
            o.propagate(True)
            rg, G = o.gyration()
            w = o.rotation_frequency()
            npatch, psize = o.patches()
            ek = o.ekin_monomers()
            row = [o.epot() / n, ek, ek * 2.0 / (3.0 * (n - 1)), npatch, psize, rg] + list(G.ravel(order="F")) + list(w)
            blocks.add(t, sp.star_sample(row, o.temperature()), o.velocity_profile(sp.PROFILE_WIDTH))
        else:
            o.propagate(False)
        o.step(t, c["mpc_dt"])
    keys, vx, vy, cnt = blocks.prof_all.table(1)
    return dict(seed=seed, steps=steps, seconds=time.time() - t0, blocks=np.array(blocks.blocks).tolist(),
                slopes=blocks.slopes, profile=dict(keys=keys.tolist(), vx=vx.tolist(), count=cnt.tolist()))


def fluid_run(args):
    seed, collisions, shear = args
    c = sp.FLUID
    o = Oracle(L=(c["L"],) * 3, kT=c["kT"], n_c=10, shear=shear, step_update=1, seed=seed, replica=0)
    o.initialize()
    fa = FourierAutocorr(sp.FOURIER_ENTRIES, sp.FLUID_EVERY * c["mpc_dt"])
    prof = VelocityProfile(sp.SHEAR_WIDTH)
    temps = []
    t0 = time.time()
    for t in range(collisions):
        o.step(t, c["mpc_dt"])
        if t >= sp.FLUID_EQUIL and t % sp.FLUID_EVERY == 0:
            temps.append(o.temperature())
            if shear == 0.0:
                fa.add(o.fourier_modes())
            else:
                prof.add(o.velocity_profile(sp.SHEAR_WIDTH))
    out = dict(seed=seed, collisions=collisions, shear=shear, seconds=time.time() - t0,
               T_mean=float(np.mean(temps)), T_n=len(temps))
    if shear == 0.0:
        out["corr"] = fa.correlation().tolist()
        out["count"] = fa.count
        out["nu"] = fa.viscosity(c["L"])
    else:
        out["slope"] = prof.shear_rate()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=7)
    ap.add_argument("--star-runs", type=int, default=7)
    ap.add_argument("--star-steps", type=int, default=1_250_000)
    ap.add_argument("--fluid-runs", type=int, default=7)
    ap.add_argument("--fluid-collisions", type=int, default=40_000)
    ap.add_argument("--shear-collisions", type=int, default=5_000)
    ap.add_argument("--only", choices=["star", "fluid"], default=None)
    ap.add_argument("--resummarize", action="store_true",
                    help="recompute the star summary from the runs already in the fixture")
    a = ap.parse_args()
    if a.resummarize:
        write_star(json.load(open(os.path.join(HERE, "statistics_star_L20.json")))["runs"])
        return
    with Pool(a.workers) as pool:
        if a.only in (None, "fluid"):
            jobs = [(SEED0 + i, a.fluid_collisions, 0.0) for i in range(a.fluid_runs)]
            jobs += [(SEED0 + 100 + i, a.shear_collisions, sp.SHEAR_RATE) for i in range(a.fluid_runs)]
            res = pool.map(fluid_run, jobs)
            eq = [r for r in res if r["shear"] == 0.0]
            sh = [r for r in res if r["shear"] != 0.0]
            nu = [r["nu"] for r in eq]
            doc = dict(protocol=dict(FLUID=sp.FLUID, equil=sp.FLUID_EQUIL, every=sp.FLUID_EVERY,
                                     entries=sp.FOURIER_ENTRIES, shear=sp.SHEAR_RATE, width=sp.SHEAR_WIDTH),
                       nu_mean=float(np.mean(nu)), nu_se=float(np.std(nu, ddof=1) / np.sqrt(len(nu))),
                       slope_mean=float(np.mean([r["slope"] for r in sh])),
                       slope_se=float(np.std([r["slope"] for r in sh], ddof=1) / np.sqrt(len(sh))),
                       T_mean=float(np.mean([r["T_mean"] for r in res])), runs=res)
            with open(os.path.join(HERE, "statistics_fluid_L20.json"), "w") as f:
                json.dump(doc, f, indent=1)
            print("fluid: nu %.4f +- %.4f, slope %.5f +- %.5f" % (doc["nu_mean"], doc["nu_se"], doc["slope_mean"],
                                                                 doc["slope_se"]), flush=True)
        if a.only in (None, "star"):
            res = pool.map(star_run, [(SEED0 + 200 + i, a.star_steps) for i in range(a.star_runs)])
            write_star(res)


def write_star(res):
    doc = dict(protocol=dict(STAR=sp.STAR, equil=sp.STAR_EQUIL, every=sp.STAR_EVERY, block=sp.STAR_BLOCK,
                             batch=sp.STAR_SUPER, width=sp.PROFILE_WIDTH),
               keys=sp.STAR_KEYS, runs=res)
    m, se, sm, sse, nb = sp.star_reference(doc)
    doc.update(mean=m.tolist(), se=se.tolist(), n_batches=nb, slope_mean=sm, slope_se=sse)
    with open(os.path.join(HERE, "statistics_star_L20.json"), "w") as f:
        json.dump(doc, f, indent=1)
    for k, v, e in zip(sp.STAR_KEYS, m, se):
        print("star: %-10s %.5f +- %.5f" % (k, v, e), flush=True)
    print("star: slope %.5f +- %.5f  (%d batches)" % (sm, sse, nb), flush=True)


if __name__ == "__main__":
    main()
