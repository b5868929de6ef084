"""Benchmark of the MPC-MD hot path on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on):
plain MPC-MD, cubic box L=64 (262,144 unit cells), rho=10 solvent particles
per cell (2,621,440), FP32 solvent state with 64-bit fixed-point cell sums;
one star of f=20 arms x (15 A + 15 B) beads (601 monomers, mass 10); shear
rate 0.05, SRD angle 130 deg, collision time h=0.1, MD dt=0.01, 10
velocity-Verlet substeps per collision.  A "step" is one collision interval
(10 MD substeps + one collision), replayed from a HIP graph.

Prints one JSON line: particle-updates/s (solvent particles x collisions per
second), the roofline of the collision pipeline measured with HIP events on
the library's stream, the learned closure's star-steps/s against explicit
MPC on the same GPU, and the FP64 CPU oracle timed on a bounded sample.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 100]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MPC particle-updates/s (L=64,rho=10); % HBM peak; star-steps/s learned vs MPC"
HBM_PEAK_GBS = 8000.0
B_ALG = 48  # bytes per particle-update: FP32 position+velocity read once and written once

WORKLOAD = dict(L=(64, 64, 64), n_c=10, shear=0.05, step_update=10, mpc_dt=0.1, md_dt=0.01,
                arms=20, A=15, B=15, mass=10.0, seed=1)


def make_system(w):
    from starpolymersmd_amd import MPCSystem, make_params

    g = MPCSystem(make_params(L=w["L"], n_c=w["n_c"], shear=w["shear"], step_update=w["step_update"],
                              md_dt=w["md_dt"], seed=w["seed"], lam=1.0, kT=1.0))
    g.add_star(w["A"], w["B"], w["arms"], w["mass"], coil=True)
    g.md_init()
    g.initialize()
    return g


# rocprofv3 FETCH_SIZE/WRITE_SIZE summaries of the two collision paths
TRAFFIC_PROFILE = {"rows": "profiles/collision_pipeline_rows_L64_rho10.md",
                   "sort": "profiles/collision_pipeline_L64_rho10.md"}


def measured_traffic(path):
    """HBM bytes per collision of the pipeline of `path` ("rows" or "sort")
    from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (each kernel's
    counters scaled by the calibration of its access shape); None when the
    summary is absent."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_PROFILE[path])) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_collision:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(w, intervals, gpu_star, threads=1):
    """FP64 oracle (faithful restatement of the reference CPU path) on the
    same workload shape for a bounded number of intervals.  threads=1 is
    the reference's own single-threaded schedule; more threads run the
    oracle's solvent loops (streaming, shift, cell index, per-cell
    collision) with OpenMP, the sort's push order and the MD staying
    sequential."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as ob

    ob.set_threads(threads)
    o = ob.Oracle(L=w["L"], n_c=w["n_c"], shear=w["shear"], step_update=w["step_update"], md_dt=w["md_dt"],
                  rng_mode=ob.RNG_PHILOX, seed=w["seed"])
    o.add_star(w["A"], w["B"], w["arms"], w["mass"])  # topology; positions/velocities from the library
    pos, vel, _ = gpu_star
    o.set_monomers(pos, vel)
    o.md_init()
    o.initialize()
    S = w["step_update"]
    o.propagate()
    o.step(0, w["mpc_dt"])  # first collision outside the timed region
    t0 = time.perf_counter()
    t = 1
    for _ in range(intervals):
        for _ in range(S):
            o.propagate()
            o.step(t, w["mpc_dt"])
            t += 1
    el = time.perf_counter() - t0
    n = o.n_fluid
    used = ob.get_threads()
    return {"value": n * intervals / el, "unit": "particle-updates/s", "cores": used, "kind": "port",
            "sample": f"FP64 oracle, L=64 rho=10 f=20 15A+15B, {intervals} collision intervals "
                      f"(10 MD substeps each), {used} thread{'s' if used > 1 else ''}, {el:.1f} s"}


def batch_workload(n_rep=64, seed=2024):
    """64 replicas of uneven size: L in {20, 24, 32, 40}, f in 5..30 arms,
    N_A = N_B in 3..15 beads, shear rates log-spaced 1e-3 .. 1e-1, rho 10
    (seeded draws).  N_A + N_B is drawn again until the arm fits half the
    box less two cells: the Verlet grid spans the box around the star's
    centre (Box.cpp:129-142), so a longer arm stops the run there too."""
    import numpy as np

    rng = np.random.default_rng(seed)
    shears = np.logspace(-3, -1, n_rep)
    reps = []
    for r in range(n_rep):
        L = int(rng.choice([20, 24, 32, 40]))
        f = int(rng.integers(5, 31))
        nab = int(rng.integers(3, 16))
        while 2 * nab > L // 2 - 2:
            nab = int(rng.integers(3, 16))
        reps.append(dict(L=L, arms=f, A=nab, B=nab, shear=float(shears[r])))
    return reps


def batch_bench(intervals=100, seq_intervals=10):
    """Batched replicas: all 64 advance through one graph launch per
    collision interval with autotuned launch shapes, against the same
    replicas advanced one handle at a time on the same GPU."""
    from starpolymersmd_amd import MPCSystem, make_params
    from starpolymersmd_amd.autotune import autotune

    reps = batch_workload()
    S, h = 10, 0.1

    def params(c):
        return make_params(L=(c["L"],) * 3, n_c=10, shear=c["shear"], step_update=S, md_dt=0.01, seed=7, lam=1.0)

    g = MPCSystem([params(c) for c in reps])
    for r, c in enumerate(reps):
        g.add_star(c["A"], c["B"], c["arms"], 10.0, r=r, coil=True)
    g.md_init()
    g.initialize()
    g.run(0, 1, h)
    tune = autotune(g, 1, h, cache=None)
    t = tune["t"]
    g.run(t, 10 * S, h)
    t += 10 * S
    g.sync()
    n_f = sum(g.num_fluid(r) for r in range(len(reps)))
    t0 = time.perf_counter()
    g.run(t, intervals * S, h)
    g.sync()
    el = time.perf_counter() - t0
    batched = n_f * intervals / el
    g.close()
    # one handle per replica, the same intervals each
    seq_el = 0.0
    for c in reps:
        q = MPCSystem(params(c))
        q.add_star(c["A"], c["B"], c["arms"], 10.0, coil=True)
        q.md_init()
        q.initialize()
        q.run(0, 1 + 5 * S, h)
        q.sync()
        t1 = time.perf_counter()
        q.run(1 + 5 * S, seq_intervals * S, h)
        q.sync()
        seq_el += (time.perf_counter() - t1) * intervals / seq_intervals
        q.close()
    return {"workload": "64 replicas, L in {20,24,32,40}, f 5..30, N_A=N_B 3..15, shear 1e-3..1e-1 log-spaced, "
                        "rho 10, coil stars (seeded draws)",
            "solvent_particles": n_f, "monomers": sum(c["arms"] * 2 * c["A"] + 1 for c in reps),
            "tuning": {"ppt": tune["ppt"], "tile_cells": tune["tile_cells"]},
            "ms_per_interval": el * 1e3 / intervals, "particle_updates_per_s": batched,
            "one_handle_at_a_time_particle_updates_per_s": n_f * intervals / seq_el,
            "speedup": seq_el / el}


def closure_bench(replicas=256, explicit_replicas=16, check_replicas=64, check_shear=0.012, check_warmup=400,
                  check_intervals=800, check_box=80, train_stars=4, train_range=(4e-3, 1.6e-2), infer_range=None,
                  train_box=80):
    """Learned closure at batch (256 stars, f=30 arms of 20 A + 20 B beads,
    1,201 monomers, shear rates uniform in infer_range) against explicit
    MPC-MD of the same star at L=64, rho=10 on the same GPU: star-steps/s
    (MD steps x stars per second), the network trained here on explicit runs
    of the library (8 shear rates log-spaced over train_range, solvent
    started from the steady flow profile; train_stars independent stars per
    rate, boxes of train_box that hold the stretched star) with the
    headline inputs (closure.HEADLINE), fallbacks included.  The 256-star
    run is cross-checked every 100th interval by explicit MPC at L=48,
    rho=10 (ClosureRunner.cross_check_summary).  Accuracy: check_replicas
    replicas each way at check_shear in steady flow, Rg^2, alignment angle
    and rotation frequency averaged over the run after a warm-up, standard
    errors from the spread of replica means; each closure/explicit ratio
    carries the standard error propagated from both sides (Rg =
    sqrt(<Rg^2>), so its ratio is the square root of the Rg^2 ratio with
    half its relative error).  The same
    check runs with the full 24-feature network and with the network
    restricted to the flow-relative velocity and the shear rate
    (free_draining); the networks' likelihood of held-out collisions is
    compared per sample."""
    import numpy as np

    from starpolymersmd_amd import MPCSystem, make_params
    from starpolymersmd_amd import closure as cl

    arms, A, B, S, dt, L = 30, 20, 20, 10, 0.01, 64
    lo, hi = train_range
    ilo, ihi = infer_range if infer_range is not None else train_range
    shears_train = list(np.geomspace(lo, hi, 8))
    t0 = time.perf_counter()
    parts = []
    for k in range(train_stars):
        parts.append(cl.generate(shears_train, L=train_box, arms=arms, A=A, B=B, intervals=150, warmup=800,
                                 seed=11 + 7 * k))
        print(f"[closure] training runs: star {k + 1} of {train_stars} per rate", file=sys.stderr, flush=True)
    if len(parts) > 1:
        xs = [q.X.download((q.n, cl.N_FEAT)) for q in parts]
        ys = [q.Y.download((q.n, 3)) for q in parts]
        for q in parts:
            q.X.close()
            q.Y.close()
        data = cl.ClosureData.from_numpy(np.concatenate(xs), np.concatenate(ys))
        del xs, ys
    else:
        data = parts[0]
    t_gen = time.perf_counter() - t0
    print(f"[closure] training data: {data.n} samples in {t_gen:.1f} s", file=sys.stderr, flush=True)
    net = cl.ClosureNet(1)
    t0 = time.perf_counter()
    steps = 2000
    hist = net.fit(data, steps=steps, batch=65536, lr=2e-3, features=cl.HEADLINE)
    t_fit = time.perf_counter() - t0
    rng = np.random.default_rng(0)
    shears = list(rng.uniform(ilo, ihi, replicas))
    run = cl.ClosureRunner(net, data.fmin, data.fmax, shear=None, arms=arms, A=A, B=B, L=L, shears=shears)
    run.advance(2)
    run.sys.sync()
    n = 20
    t0 = time.perf_counter()
    run.advance(n, chunk=n)
    run.sys.sync()
    el = time.perf_counter() - t0
    closure_rate = replicas * n * S / el
    fallbacks = run.fallbacks
    del run
    # the same 256 stars with the explicit cross-check of every 100th interval
    xrun = cl.ClosureRunner(net, data.fmin, data.fmax, shear=None, arms=arms, A=A, B=B, L=48, shears=shears,
                            check_every=100, check_replicas=16, check_intervals=20, seed=700)
    xrun.advance(300, chunk=50)
    xrun.sys.sync()
    cross = xrun.cross_check_summary()
    print("[closure] throughput and cross-check done", file=sys.stderr, flush=True)
    cross["fallbacks"] = xrun.fallbacks
    xrun.close()
    # explicit MPC-MD of the same star, batched replicas
    g = MPCSystem([make_params(L=(L, L, L), n_c=10, shear=float(shears[r]), step_update=S, md_dt=dt, seed=3 + r)
                   for r in range(explicit_replicas)])
    for r in range(explicit_replicas):
        g.add_star(A, B, arms, 10.0, r=r, coil=True)
    g.md_init()
    g.initialize()
    cl.preshear(g)
    g.run(0, 1 + 5 * S, 0.1)
    g.sync()
    t0 = time.perf_counter()
    g.run(1 + 5 * S, n * S, 0.1)
    g.sync()
    el = time.perf_counter() - t0
    explicit_rate = explicit_replicas * n * S / el
    g.close()

    # accuracy in steady shear flow: ensembles of replicas, explicit vs closure
    def ensemble(advance, sysm, R):
        """per-replica time averages of Rg^2, alignment angle and the
        rotation frequency about z (Box::calculate_rotation_frequency);
        ensemble means and standard errors, the wall time, and the ensemble
        mean Rg^2 every 10 intervals from the coil start"""
        t0 = time.perf_counter()
        trace = []
        for _ in range(check_warmup // 10):
            advance(10)
            trace.append(np.mean([np.trace(sysm.gyration_tensor(r)) for r in range(R)]))
        acc = []
        for _ in range(check_intervals // 10):
            advance(10)
            acc.append([(*cl.shape_stats(sysm.gyration_tensor(r)), sysm.rotation_frequency(r)[2]) for r in range(R)])
            trace.append(np.mean([a[0] for a in acc[-1]]))
        sysm.sync()
        el = time.perf_counter() - t0
        pm = np.array(acc).mean(axis=0)  # per replica (Rg^2, chi, w_z)
        return pm.mean(axis=0), pm.std(axis=0, ddof=1) / np.sqrt(R), el, np.array(trace)

    R = check_replicas
    # the stars stretch over the check: a box whose Verlet grid holds them
    # (a monomer leaving it ends the run, as the reference's SIGSEGV does)
    Le = check_box
    e = MPCSystem([make_params(L=(Le, Le, Le), n_c=10, shear=check_shear, step_update=S, md_dt=dt, seed=100 + r)
                   for r in range(R)])
    for r in range(R):
        e.add_star(A, B, arms, 10.0, r=r, coil=True)
    e.md_init()
    e.initialize()
    cl.preshear(e)
    t = [0]

    def adv_e(k):
        e.run(t[0], k * S + (1 if t[0] == 0 else 0), 0.1)
        t[0] += k * S + (1 if t[0] == 0 else 0)

    em, es, et, etr = ensemble(adv_e, e, R)
    e.close()
    print(f"[closure] explicit ensemble done in {et:.1f} s", file=sys.stderr, flush=True)

    def check(network, seed=500):
        c = cl.ClosureRunner(network, data.fmin, data.fmax, shear=None, arms=arms, A=A, B=B, L=L,
                             shears=[check_shear] * R, seed=seed)
        cm, cs, ct, ctr = ensemble(lambda k: c.advance(k, chunk=k), c.sys, R)
        fb = c.fallbacks
        c.close()
        ratio = cm / em
        rse = np.abs(ratio) * np.sqrt((cs / cm) ** 2 + (es / em) ** 2)
        return {"Rg2": float(cm[0]), "Rg2_se": float(cs[0]), "chi": float(cm[1]), "chi_se": float(cs[1]),
                "wz": float(cm[2]), "wz_se": float(cs[2]),
                "Rg2_ratio": float(ratio[0]), "Rg2_ratio_se": float(rse[0]),
                "Rg_ratio": float(np.sqrt(ratio[0])), "Rg_ratio_se": float(0.5 * rse[0] / np.sqrt(ratio[0])),
                "chi_ratio": float(ratio[1]), "chi_ratio_se": float(rse[1]),
                "wz_ratio": float(ratio[2]), "wz_ratio_se": float(rse[2]),
                "seconds": ct, "fallbacks": fb,
                "time_to_equal_Rg2_error_speedup": float((et * es[0] ** 2) / (ct * cs[0] ** 2))}

    head = check(net)
    print("[closure] headline check done", file=sys.stderr, flush=True)
    # what the inputs gain: the full 24 features and the flow-relative
    # velocity with the shear rate alone (a learned free-draining Langevin
    # bath); the same ensemble check, and the likelihood of one collision's
    # velocity changes held out from training (an independent run at the
    # check's shear), compared per sample with standard errors from interval
    # blocks (the monomers of one interval are correlated)
    held = cl.generate([check_shear], L=80, arms=arms, A=A, B=B, intervals=50, warmup=800, seed=99)
    xh, yh = held.X.download((held.n, 24)), held.Y.download((held.n, 3))
    held.X.close()
    held.Y.close()
    nmono = 1 + arms * (A + B)
    nll = {"headline": net.nll(xh, yh)}

    def gain(a, b):
        gg = (nll[a] - nll[b]).reshape(-1, nmono).mean(axis=1)
        return float(gg.mean()), float(gg.std(ddof=1) / np.sqrt(len(gg)))

    ablations = {}
    for key, feats in (("full", None), ("free_draining", cl.FREE_DRAINING)):
        n2 = cl.ClosureNet(1)
        h2 = n2.fit(data, steps=steps, batch=65536, lr=2e-3, features=feats)
        ablations[key] = {"features": list(feats) if feats is not None else list(range(cl.N_FEAT)),
                          "loss_last": float(np.mean(h2[-10:])), **check(n2)}
        nll[key] = n2.nll(xh, yh)
        ablations[key]["heldout_nll"] = float(nll[key].mean())
    g_head, g_head_se = gain("free_draining", "headline")
    g_full, g_full_se = gain("free_draining", "full")
    return {
        "star_steps_per_s": closure_rate, "explicit_star_steps_per_s": explicit_rate,
        "speedup": closure_rate / explicit_rate,
        "workload": f"{replicas} stars f={arms} {A}A+{B}B (1,201 monomers) closure vs {explicit_replicas} stars "
                    f"explicit MPC L={L} rho=10, shear U[{ilo},{ihi}], {n} intervals of {S} MD steps",
        "fallbacks": fallbacks,
        "network_inputs": list(cl.HEADLINE),
        "train": {"samples": data.n, "shear_rates": 8, "shear_range": [lo, hi], "stars_per_rate": train_stars,
                  "box": train_box, "steps": steps, "batch": 65536,
                  "generate_seconds": t_gen, "fit_seconds": t_fit, "ms_per_adam_step": t_fit / steps * 1e3,
                  "loss_first": float(hist[0]), "loss_last": float(np.mean(hist[-10:]))},
        "cross_check": cross,
        "check": {"shear": check_shear, "replicas": R, "explicit_box": Le, "warmup_intervals": check_warmup,
                  "intervals": check_intervals, "steady_flow": True,
                  "explicit_Rg2": float(em[0]), "explicit_Rg2_se": float(es[0]),
                  "explicit_chi": float(em[1]), "explicit_chi_se": float(es[1]),
                  "explicit_wz": float(em[2]), "explicit_wz_se": float(es[2]),
                  "explicit_seconds": et, "closure": head},
        # the 1,200 intervals of the check are the start-up: the stars are
        # still stretching.  The steady state takes ~13,000 intervals and is
        # measured by --closure-converge, where the closure over-stretches
        # and does not converge to the explicit plateau.
        "time_to_converged_Rg_speedup": None,
        "ablations": ablations,
        "heldout": {"samples": int(len(yh)), "nll_headline": float(nll["headline"].mean()),
                    "nll_gain_headline_over_free_draining": g_head,
                    "nll_gain_headline_over_free_draining_se": g_head_se,
                    "nll_gain_full_over_free_draining": g_full, "nll_gain_full_over_free_draining_se": g_full_se},
    }


def closure_converge(intervals=20000, explicit_replicas=16, closure_replicas=64, shear=0.012, box=80,
                     train_range=(4e-3, 1.6e-2), train_stars=4, every=50, history=False):
    """Time to converged Rg from the coil start: explicit MPC-MD (L=box,
    rho=10) and the learned closure (headline inputs) each advance an
    ensemble of the f=30, 20A+20B star from the seeded coil in steady shear
    flow for `intervals` collision intervals, recording the ensemble mean
    Rg^2 every `every` intervals.  Plateau: the mean over the last quarter;
    converged: the first record from which the running mean of 5 records
    stays within 5% of the plateau.  Speedup: explicit over closure wall
    time per star to that point.  The plateaus are the steady-state
    accuracy check (Rg, alignment angle and w_z over the last quarter)."""
    import numpy as np

    from starpolymersmd_amd import MPCSystem, make_params
    from starpolymersmd_amd import closure as cl

    arms, A, B, S = 30, 20, 20, 10
    lo, hi = train_range
    t0 = time.perf_counter()
    if history:
        # one recorded interval every 150 over 15,000: the stretching, the
        # overshoot and the steady state
        data = cl.generate_batched(list(np.geomspace(lo, hi, 8)), stars_per_rate=2, L=box, arms=arms, A=A, B=B,
                                   intervals=100, stride=150,
                                   progress=lambda k, n: k % 10 == 0 and print(f"[converge] data {k}/{n}",
                                                                               file=sys.stderr, flush=True))
    else:
        parts = [cl.generate(list(np.geomspace(lo, hi, 8)), L=box, arms=arms, A=A, B=B, intervals=150,
                             warmup=800, seed=11 + 7 * k) for k in range(train_stars)]
        xs = [q.X.download((q.n, cl.N_FEAT)) for q in parts]
        ys = [q.Y.download((q.n, 3)) for q in parts]
        for q in parts:
            q.X.close()
            q.Y.close()
        data = cl.ClosureData.from_numpy(np.concatenate(xs), np.concatenate(ys))
        del xs, ys
    t_gen = time.perf_counter() - t0
    print(f"[converge] training data: {data.n} samples in {t_gen:.1f} s", file=sys.stderr, flush=True)
    net = cl.ClosureNet(1)
    net.fit(data, steps=2000, batch=65536, lr=2e-3, features=cl.HEADLINE)

    def trace(advance, sysm, R):
        t0 = time.perf_counter()
        tr, tail = [], []
        for k in range(intervals // every):
            advance(every)
            st = [(*cl.shape_stats(sysm.gyration_tensor(r)), sysm.rotation_frequency(r)[2]) for r in range(R)]
            tr.append(float(np.mean([x[0] for x in st])))
            if k % 40 == 0:
                print(f"[converge] {R} stars, interval {(k + 1) * every}: mean Rg^2 {tr[-1]:.2f}", file=sys.stderr,
                      flush=True)
            if k >= (3 * intervals // every) // 4:
                tail.append(st)
        sysm.sync()
        pm = np.array(tail).mean(axis=0)
        return np.array(tr), pm.mean(axis=0), pm.std(axis=0, ddof=1) / np.sqrt(R), time.perf_counter() - t0

    def converged(tr, plateau):
        sm = np.convolve(tr, np.ones(5) / 5, mode="valid")
        ok = np.abs(sm / plateau - 1.0) <= 0.05
        for k in range(len(ok)):
            if ok[k:].all():
                return every * (k + 5)
        return None

    e = MPCSystem([make_params(L=(box,) * 3, n_c=10, shear=shear, step_update=S, seed=100 + r)
                   for r in range(explicit_replicas)])
    for r in range(explicit_replicas):
        e.add_star(A, B, arms, 10.0, r=r, coil=True)
    e.md_init()
    e.initialize()
    cl.preshear(e)
    t = [0]

    def adv_e(k):
        n = k * S + (1 if t[0] == 0 else 0)
        e.run(t[0], n, 0.1)
        t[0] += n

    etr, em, es, et = trace(adv_e, e, explicit_replicas)
    e.close()
    c = cl.ClosureRunner(net, data.fmin, data.fmax, shear=None, arms=arms, A=A, B=B, L=64,
                         shears=[shear] * closure_replicas, seed=500)
    ctr, cm, cs, ct = trace(lambda k: c.advance(k, chunk=k), c.sys, closure_replicas)
    fb = c.fallbacks
    c.close()
    ke, kc = converged(etr, em[0]), converged(ctr, cm[0])
    per_e = et / (intervals * explicit_replicas)
    per_c = ct / (intervals * closure_replicas)
    ratio = cm / em
    rse = np.abs(ratio) * np.sqrt((cs / cm) ** 2 + (es / em) ** 2)
    return {"shear": shear, "explicit_box": box, "intervals": intervals, "record_every": every,
            "training_data": "history" if history else "window", "training_samples": data.n,
            "training_seconds": t_gen,
            "explicit_replicas": explicit_replicas, "closure_replicas": closure_replicas, "closure_fallbacks": fb,
            "explicit_converged_intervals": ke, "closure_converged_intervals": kc,
            "explicit_seconds_per_star_interval": per_e, "closure_seconds_per_star_interval": per_c,
            "time_to_converged_Rg_speedup": (ke * per_e) / (kc * per_c) if ke and kc else None,
            "steady_Rg_ratio": float(np.sqrt(ratio[0])), "steady_Rg_ratio_se": float(0.5 * rse[0] / np.sqrt(ratio[0])),
            "steady_chi_ratio": float(ratio[1]), "steady_chi_ratio_se": float(rse[1]),
            "steady_wz_ratio": float(ratio[2]), "steady_wz_ratio_se": float(rse[2]),
            "explicit_plateau": [float(x) for x in em], "closure_plateau": [float(x) for x in cm],
            "explicit_Rg2_trace": [round(float(x), 3) for x in etr],
            "closure_Rg2_trace": [round(float(x), 3) for x in ctr]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--cpu-intervals", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (ppt 4, 128-cell tiles)")
    ap.add_argument("--tuning", type=int, nargs=2, default=None, metavar=("PPT", "TILE_CELLS"),
                    help="use this launch shape instead of the warm-up sweep")
    ap.add_argument("--sort-path", action="store_true",
                    help="time the sort path (bin, scan, permutation, collide) instead of the row-tile path")
    ap.add_argument("--no-closure", action="store_true", help="skip the learned-closure throughput section")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-replicas section")
    ap.add_argument("--train-stars", type=int, default=4,
                    help="independent explicit stars per shear rate in the closure's training data")
    ap.add_argument("--closure-only", action="store_true",
                    help="only run the learned-closure section (for profiling its kernels)")
    ap.add_argument("--closure-train-range", type=float, nargs=2, default=(4e-3, 1.6e-2),
                    help="shear rates of the closure's training runs (8, log-spaced)")
    ap.add_argument("--closure-infer-range", type=float, nargs=2, default=None,
                    help="shear rates of the 256 closure stars (uniform; default: the training range)")
    ap.add_argument("--closure-train-box", type=int, default=80, help="explicit box of the training runs")
    ap.add_argument("--closure-check", type=float, nargs=2, default=(0.012, 80),
                    help="shear rate and explicit box of the closure's accuracy check")
    ap.add_argument("--closure-converge", type=int, default=0,
                    help="only run the time-to-converged-Rg comparison over N intervals")
    ap.add_argument("--closure-history", action="store_true",
                    help="train the converge run's network on whole-history data (closure.generate_batched)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run N MiB device-to-device copies (counter calibration)")
    ap.add_argument("--gather-probe", type=int, default=0,
                    help="only run gathers of N Mi floats through a permutation (counter calibration of the "
                         "collide's access shape): identity, random, then shuffled within 4,096-element blocks")
    a = ap.parse_args()
    if a.copy_probe:
        import torch

        x = torch.empty(a.copy_probe << 20, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_mib": a.copy_probe, "bytes_read": x.numel(), "bytes_written": y.numel()}))
        return
    if a.gather_probe:
        # dst[i] = src[perm[i]] with 4-byte elements: the payload is 4 B read
        # + 4 B of index + 4 B written per element, as the collide's gather
        # of one float through the sort permutation
        import torch

        n = a.gather_probe << 20
        src = torch.rand(n, device="cuda")
        g = torch.Generator(device="cuda")
        g.manual_seed(1)
        perms = {"identity": torch.arange(n, device="cuda", dtype=torch.int32),
                 "random": torch.randperm(n, device="cuda", generator=g, dtype=torch.int32)}
        blk = torch.arange(n, device="cuda", dtype=torch.int32).view(-1, 4096)
        perms["block4096"] = torch.gather(blk, 1, torch.argsort(torch.rand(blk.shape, device="cuda", generator=g),
                                                                dim=1).to(torch.int64)).view(-1).to(torch.int32)
        for name, perm in perms.items():
            idx = perm.to(torch.int64)  # index_select takes int64 indices: 8 B per element
            for _ in range(5):
                dst = torch.index_select(src, 0, idx)
            torch.cuda.synchronize()
            print(json.dumps({"gather_probe": name, "elements": n, "bytes_data_read": 4 * n,
                              "bytes_index_read": 8 * n, "bytes_written": 4 * dst.numel()}), flush=True)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (the MPC box would be "
              "slab-decomposed along z with per-face halo sums; not built). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if a.closure_converge:
        print(json.dumps({"closure_converge": closure_converge(
            intervals=a.closure_converge, shear=a.closure_check[0], box=int(a.closure_check[1]),
            train_range=tuple(a.closure_train_range), train_stars=a.train_stars,
            history=a.closure_history)}), flush=True)
        return
    if a.closure_only:
        print(json.dumps({"closure": closure_bench(train_stars=a.train_stars, train_range=tuple(a.closure_train_range),
                                                infer_range=a.closure_infer_range, train_box=a.closure_train_box,
                                                check_shear=a.closure_check[0], check_box=int(a.closure_check[1]))}), flush=True)
        return

    w = WORKLOAD
    S, h = w["step_update"], w["mpc_dt"]
    g = make_system(w)
    if a.sort_path:
        g.set_collision_path(g.PATH_SORT)
    n_f = g.num_fluid()
    star0 = g.get_monomers()
    t = 0
    g.run(t, 1, h)  # t=0: first collision
    t += 1
    tune = None
    if a.tuning:
        g.set_tuning(*a.tuning)
    elif not a.no_tune:  # launch-shape sweep inside the (untimed) warm-up
        from starpolymersmd_amd.autotune import autotune

        tune = autotune(g, t, h, cache=None)
        t = tune["t"]
    g.run(t, a.warmup * S, h)
    t += a.warmup * S
    g.sync()
    t0 = time.perf_counter()
    g.run(t, a.steps * S, h)
    g.sync()
    el = time.perf_counter() - t0
    t += a.steps * S
    ms_per_step = el * 1e3 / a.steps
    value = n_f * a.steps / el

    # roofline pass: HIP events around every launch on the library's stream
    g.profile(True)
    g.profile_reset()
    g.run(t, a.profile_steps * S, h)
    g.sync()
    t += a.profile_steps * S
    prof = {k: g.profile_read(k)[0]
            for k in ("md", "prologue", "bin", "scan", "perm", "advect", "gamma", "redo", "collide")}
    g.profile(False)
    rows = g.collision_path() == g.PATH_ROWS
    if rows:  # row-tile path: advect, the cells' Gamma draws, the redo kernel, row-window collide
        pipe_ms = prof["advect"] + prof["gamma"] + prof["redo"] + prof["collide"]
        pipe_name = "collision pipeline k_advect+k_cell_gamma+k_rows_redo+k_collide_rows"
    else:
        pipe_ms = prof["bin"] + prof["scan"] + prof["perm"] + prof["collide"]
        pipe_name = "collision pipeline k_bin+k_scan+k_perm+k_collide"
    traffic = measured_traffic("rows" if rows else "sort")
    achieved = B_ALG * n_f / (pipe_ms * 1e-3) / 1e9
    T = g.temperature()

    out = {
        "metric": METRIC, "value": value, "unit": "particle-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain MPC-MD L=64 rho=10 (2,621,440 solvent), star f=20 15A+15B "
                               "(601 monomers), shear 0.05, h=0.1, dt=0.01, 10 MD substeps/collision",
                   "solvent_particles": n_f, "step": "one collision interval", "parallelism": "1 GPU",
                   "collision_path": "row tiles" if rows else "sort",
                   "tuning": {"ppt": g.get_tuning()[0], "tile_cells": g.get_tuning()[1],
                              "sweep_ms": None if tune is None else tune["timings"]}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     # measured bytes (FETCH_SIZE + WRITE_SIZE of the pipeline's kernels) over the
                     # same kernel time: the bandwidth the hardware moved, amplification included
                     "measured_gbs": traffic / (pipe_ms * 1e-3) / 1e9 if traffic else None,
                     "measured_frac": traffic / (pipe_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
                     "amplification": traffic / (B_ALG * n_f) if traffic else None,
                     "traffic_source": TRAFFIC_PROFILE["rows" if rows else "sort"] if traffic else None,
                     "kernel": pipe_name,
                     "alg_bytes_per_launch": B_ALG * n_f, "avg_ms": pipe_ms,
                     "per_kernel_ms": prof},
        "md_us_per_interval": prof["md"] * 1e3,
        "temperature": T,
    }
    g.close()
    if not a.no_batch:
        out["batch"] = batch_bench()
    if not a.no_closure:
        out["closure"] = closure_bench(train_stars=a.train_stars, train_range=tuple(a.closure_train_range),
                                                infer_range=a.closure_infer_range, train_box=a.closure_train_box,
                                                check_shear=a.closure_check[0], check_box=int(a.closure_check[1]))
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_intervals, star0)
        # all the host threads this process may use (OMP_NUM_THREADS, else the CPUs it sees)
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
        out["cpu_baseline_threads"] = cpu_baseline(w, 3 * a.cpu_intervals, star0, threads=nthr)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
