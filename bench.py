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


TRAFFIC_PROFILE = "profiles/collision_pipeline_L64_rho10.md"


def measured_traffic():
    """HBM bytes per collision of the pipeline from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (scaled by a device copy of known size);
    None when the summary is absent."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_PROFILE)) as f:
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
                  check_intervals=800, check_box=80, train_stars=4):
    """Learned closure at batch (256 stars, f=30 arms of 20 A + 20 B beads,
    1,201 monomers, shear rates uniform in [4e-3, 1.6e-2]) against explicit
    MPC-MD of the same star at L=64, rho=10 on the same GPU (the smallest
    box that holds the star stretched by these rates): star-steps/s (MD
    steps x stars per second), the network trained here on explicit runs of
    the library (8 shear rates, solvent started from the steady flow
    profile; 4 independent stars per rate), fallbacks included.  Accuracy:
    64 replicas each way at shear
    0.012 in steady flow, Rg^2, alignment angle and rotation frequency
    averaged over the run after a warm-up, standard errors from the spread
    of replica means; each closure/explicit ratio carries the standard error
    propagated from both sides (Rg = sqrt(<Rg^2>), so its ratio is the
    square root of the Rg^2 ratio with half its relative error).  The same
    check runs with the network restricted to the flow-relative velocity and
    the shear rate (free_draining) and to those plus the neighbours' mean
    velocity, crowding and type (local_coupling); the networks' likelihood
    of held-out collisions is compared per sample."""
    import numpy as np

    from starpolymersmd_amd import MPCSystem, make_params
    from starpolymersmd_amd import closure as cl

    arms, A, B, S, dt, L = 30, 20, 20, 10, 0.01, 64
    lo, hi = 4e-3, 1.6e-2
    shears_train = list(np.geomspace(lo, hi, 8))
    t0 = time.perf_counter()
    data = cl.generate(shears_train, L=L, arms=arms, A=A, B=B, intervals=150, warmup=800)
    if train_stars > 1:
        # more independent stars per rate, in a box that holds them stretched
        parts = [data] + [cl.generate(shears_train, L=80, arms=arms, A=A, B=B, intervals=150, warmup=800,
                                      seed=11 + 7 * k) for k in range(1, train_stars)]
        xs = [q.X.download((q.n, cl.N_FEAT)) for q in parts]
        ys = [q.Y.download((q.n, 3)) for q in parts]
        for q in parts:
            q.X.close()
            q.Y.close()
        data = cl.ClosureData.from_numpy(np.concatenate(xs), np.concatenate(ys))
        del xs, ys
    t_gen = time.perf_counter() - t0
    net = cl.ClosureNet(1)
    t0 = time.perf_counter()
    steps = 2000
    hist = net.fit(data, steps=steps, batch=65536, lr=2e-3)
    t_fit = time.perf_counter() - t0
    rng = np.random.default_rng(0)
    shears = list(rng.uniform(lo, hi, replicas))
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
        ensemble means and standard errors, and the wall time"""
        t0 = time.perf_counter()
        advance(check_warmup)
        acc = []
        for _ in range(check_intervals // 10):
            advance(10)
            acc.append([(*cl.shape_stats(sysm.gyration_tensor(r)), sysm.rotation_frequency(r)[2]) for r in range(R)])
        sysm.sync()
        el = time.perf_counter() - t0
        pm = np.array(acc).mean(axis=0)  # per replica (Rg^2, chi, w_z)
        return pm.mean(axis=0), pm.std(axis=0, ddof=1) / np.sqrt(R), el

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

    em, es, et = ensemble(adv_e, e, R)
    e.close()
    c = cl.ClosureRunner(net, data.fmin, data.fmax, shear=None, arms=arms, A=A, B=B, L=L, shears=[check_shear] * R,
                         seed=500)
    cm, cs, ct = ensemble(lambda k: c.advance(k, chunk=k), c.sys, R)
    c.close()
    # time to the same standard error of Rg^2 (the error falls as 1/sqrt(time))
    ttc = (et * es[0] ** 2) / (ct * cs[0] ** 2)
    ratio = cm / em
    ratio_se = np.abs(ratio) * np.sqrt((cs / cm) ** 2 + (es / em) ** 2)
    # what the inputs gain: the same training restricted to the flow-relative
    # velocity and the shear rate (a learned free-draining Langevin bath) and
    # to those plus the neighbours' mean velocity, crowding and type (local
    # coupling); the same ensemble check, and the likelihood of one
    # collision's velocity changes held out from training (an independent
    # run at the check's shear), compared per sample with standard errors
    # from interval blocks (the monomers of one interval are correlated)
    held = cl.generate([check_shear], L=80, arms=arms, A=A, B=B, intervals=50, warmup=800, seed=99)
    xh, yh = held.X.download((held.n, 24)), held.Y.download((held.n, 3))
    held.X.close()
    held.Y.close()
    nmono = 1 + arms * (A + B)
    nll = {"full": net.nll(xh, yh)}

    def gain(a, b):
        g = (nll[a] - nll[b]).reshape(-1, nmono).mean(axis=1)
        return float(g.mean()), float(g.std(ddof=1) / np.sqrt(len(g)))

    restricted = {}
    for key, feats in (("free_draining", cl.FREE_DRAINING), ("local_coupling", cl.LOCAL_COUPLING)):
        n2 = cl.ClosureNet(1)
        h2 = n2.fit(data, steps=steps, batch=65536, lr=2e-3, features=feats)
        r2 = cl.ClosureRunner(n2, data.fmin, data.fmax, shear=None, arms=arms, A=A, B=B, L=L,
                              shears=[check_shear] * R, seed=500)
        m2, s2, t2 = ensemble(lambda k: r2.advance(k, chunk=k), r2.sys, R)
        fb2 = r2.fallbacks
        r2.close()
        nll[key] = n2.nll(xh, yh)
        rt = m2 / em
        rse = np.abs(rt) * np.sqrt((s2 / m2) ** 2 + (es / em) ** 2)
        restricted[key] = {"features": list(feats), "loss_last": float(np.mean(h2[-10:])),
                           "Rg2_ratio": float(rt[0]), "Rg2_ratio_se": float(rse[0]),
                           "Rg_ratio": float(np.sqrt(rt[0])), "Rg_ratio_se": float(0.5 * rse[0] / np.sqrt(rt[0])),
                           "chi_ratio": float(rt[1]), "chi_ratio_se": float(rse[1]),
                           "wz_ratio": float(rt[2]), "wz_ratio_se": float(rse[2]),
                           "fallbacks": fb2, "heldout_nll": float(nll[key].mean()),
                           "time_to_equal_Rg2_error_speedup": float((et * es[0] ** 2) / (t2 * s2[0] ** 2))}
    g_full, g_full_se = gain("free_draining", "full")
    g_loc, g_loc_se = gain("free_draining", "local_coupling")
    return {
        "star_steps_per_s": closure_rate, "explicit_star_steps_per_s": explicit_rate,
        "speedup": closure_rate / explicit_rate,
        "workload": f"{replicas} stars f={arms} {A}A+{B}B (1,201 monomers) closure vs {explicit_replicas} stars "
                    f"explicit MPC L={L} rho=10, shear U[{lo},{hi}], {n} intervals of {S} MD steps",
        "fallbacks": fallbacks,
        "train": {"samples": data.n, "shear_rates": 8, "stars_per_rate": train_stars, "steps": steps, "batch": 65536,
                  "generate_seconds": t_gen, "fit_seconds": t_fit,
                  "loss_first": float(hist[0]), "loss_last": float(np.mean(hist[-10:]))},
        "check": {"shear": check_shear, "replicas": R, "explicit_box": Le, "warmup_intervals": check_warmup,
                  "intervals": check_intervals, "steady_flow": True,
                  "explicit_Rg2": float(em[0]), "explicit_Rg2_se": float(es[0]),
                  "closure_Rg2": float(cm[0]), "closure_Rg2_se": float(cs[0]),
                  "explicit_chi": float(em[1]), "explicit_chi_se": float(es[1]),
                  "closure_chi": float(cm[1]), "closure_chi_se": float(cs[1]),
                  "explicit_wz": float(em[2]), "explicit_wz_se": float(es[2]),
                  "closure_wz": float(cm[2]), "closure_wz_se": float(cs[2]),
                  "Rg2_ratio": float(ratio[0]), "Rg2_ratio_se": float(ratio_se[0]),
                  "Rg_ratio": float(np.sqrt(ratio[0])), "Rg_ratio_se": float(0.5 * ratio_se[0] / np.sqrt(ratio[0])),
                  "chi_ratio": float(ratio[1]), "chi_ratio_se": float(ratio_se[1]),
                  "wz_ratio": float(ratio[2]), "wz_ratio_se": float(ratio_se[2]),
                  "explicit_seconds": et, "closure_seconds": ct,
                  "time_to_equal_Rg2_error_speedup": float(ttc),
                  "closure_fallbacks": c.fallbacks},
        "restricted_networks": restricted,
        "heldout": {"samples": int(len(yh)), "nll_full": float(nll["full"].mean()),
                    "nll_gain_full_over_free_draining": g_full, "nll_gain_full_over_free_draining_se": g_full_se,
                    "nll_gain_local_coupling_over_free_draining": g_loc,
                    "nll_gain_local_coupling_over_free_draining_se": g_loc_se},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--cpu-intervals", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (ppt 4, 128-cell tiles)")
    ap.add_argument("--no-closure", action="store_true", help="skip the learned-closure throughput section")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-replicas section")
    ap.add_argument("--train-stars", type=int, default=4,
                    help="independent explicit stars per shear rate in the closure's training data")
    ap.add_argument("--closure-only", action="store_true",
                    help="only run the learned-closure section (for profiling its kernels)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run N MiB device-to-device copies (counter calibration)")
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
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (the MPC box would be "
              "slab-decomposed along z with per-face halo sums; not built). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if a.closure_only:
        print(json.dumps({"closure": closure_bench(train_stars=a.train_stars)}), flush=True)
        return

    w = WORKLOAD
    S, h = w["step_update"], w["mpc_dt"]
    g = make_system(w)
    n_f = g.num_fluid()
    star0 = g.get_monomers()
    t = 0
    g.run(t, 1, h)  # t=0: first collision
    t += 1
    tune = None
    if not a.no_tune:  # launch-shape sweep inside the (untimed) warm-up
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
    prof = {k: g.profile_read(k)[0] for k in ("md", "bin", "scan", "perm", "collide")}
    g.profile(False)
    pipe_ms = prof["bin"] + prof["scan"] + prof["perm"] + prof["collide"]
    traffic = measured_traffic()
    achieved = B_ALG * n_f / (pipe_ms * 1e-3) / 1e9
    T = g.temperature()

    out = {
        "metric": METRIC, "value": value, "unit": "particle-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain MPC-MD L=64 rho=10 (2,621,440 solvent), star f=20 15A+15B "
                               "(601 monomers), shear 0.05, h=0.1, dt=0.01, 10 MD substeps/collision",
                   "solvent_particles": n_f, "step": "one collision interval", "parallelism": "1 GPU",
                   "tuning": {"ppt": g.get_tuning()[0], "tile_cells": g.get_tuning()[1],
                              "sweep_ms": None if tune is None else tune["timings"]}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     # measured bytes (FETCH_SIZE + WRITE_SIZE of the pipeline's kernels) over the
                     # same kernel time: the bandwidth the hardware moved, amplification included
                     "measured_gbs": traffic / (pipe_ms * 1e-3) / 1e9 if traffic else None,
                     "measured_frac": traffic / (pipe_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
                     "amplification": traffic / (B_ALG * n_f) if traffic else None,
                     "traffic_source": TRAFFIC_PROFILE if traffic else None,
                     "kernel": "collision pipeline k_bin+k_scan+k_perm+k_collide",
                     "alg_bytes_per_launch": B_ALG * n_f, "avg_ms": pipe_ms,
                     "per_kernel_ms": prof},
        "md_us_per_interval": prof["md"] * 1e3,
        "temperature": T,
    }
    g.close()
    if not a.no_batch:
        out["batch"] = batch_bench()
    if not a.no_closure:
        out["closure"] = closure_bench(train_stars=a.train_stars)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_intervals, star0)
        # all the host threads this process may use (OMP_NUM_THREADS, else the CPUs it sees)
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
        out["cpu_baseline_threads"] = cpu_baseline(w, 3 * a.cpu_intervals, star0, threads=nthr)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
