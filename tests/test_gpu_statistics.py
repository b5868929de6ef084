"""Statistical agreement of the device path with the FP64 oracle and with
the closed forms of the method (the sampling protocol is stat_protocol.py;
the oracle's long runs are tests/golden/statistics_*.json, made by
tests/golden/make_statistics.py).

The two sides run independent random streams, so they are compared through
means and standard errors: |mean_gpu - mean_oracle| <= Z * sqrt(se_gpu^2 +
se_oracle^2) with Z = 4 (a test of 15 quantities at Z = 4 has a false alarm
rate below 1e-3 for normal errors).  The device runs many replicas of the
same system in one handle; replicas are independent, so the device's error
is the spread of the replica means.  Closed forms: the MPC temperature is
kT (MBS thermostat), the shear profile slope is the imposed rate
(results/plot_MPC_tests.py:61), the transverse modes decay with the SRD
viscosity (results/plot_MPC_tests.py:107, 0.87002 at rho 10, h 0.1, 130 deg),
the speeds follow Maxwell's distribution (results/plot_MPC_tests.py:82)."""
import json
import math
import os

import numpy as np
import pytest

import stat_protocol as sp
from starpolymersmd_amd import MPCSystem, make_params
from starpolymersmd_amd.observables import FourierAutocorr, VelocityProfile, run_with_output, srd_viscosity

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
Z = 4.0


def _fixture(name):
    path = os.path.join(HERE, "golden", name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated")
    return json.load(open(path))


def _agree(name, mg, sg, mo, so):
    bound = Z * math.sqrt(sg * sg + so * so)
    print(f"  {name:10s} device {mg:.5f} +- {sg:.5f}  oracle {mo:.5f} +- {so:.5f}", flush=True)
    assert abs(mg - mo) <= bound, (name, mg, sg, mo, so)


def test_star_statistics_agree_with_oracle():
    ref = _fixture("statistics_star_L20.json")
    c = sp.STAR
    R = 64
    every = 500
    g = MPCSystem([make_params(L=(c["L"],) * 3, kT=c["kT"], n_c=c["n_c"], shear=c["shear"],
                               step_update=c["step_update"], md_dt=c["md_dt"], lam=c["lam"], seed=7)] * R)
    for r in range(R):
        g.add_star(c["A"], c["B"], c["arms"], c["mass"], r=r)
    g.md_init()
    g.initialize()
    blocks = [sp.StarBlocks() for _ in range(R)]

    def observe(t):
        for r in range(R):
            row = g.statistics(r)
            blocks[r].add(t, sp.star_sample(row, g.temperature(r)), g.velocity_profile(sp.PROFILE_WIDTH, r))

    steps = sp.STAR_EQUIL + 150_000 + 1
    run_with_output(g, 0, steps, c["mpc_dt"], every, observe, equil=sp.STAR_EQUIL)
    g.sync()
    per_rep = np.array([np.mean(b.blocks, axis=0) for b in blocks])
    slopes = np.array([np.mean(b.slopes) for b in blocks])
    mg, sg = sp.summary(per_rep)
    mo, so, smo, sso, _ = sp.star_reference(ref)
    for k, name in enumerate(sp.STAR_KEYS):
        _agree(name, mg[k], sg[k], mo[k], so[k])
    ms, ss = sp.summary(slopes)
    _agree("slope", ms, ss, smo, sso)
    # closed forms: MBS thermostat, imposed shear rate
    assert abs(mg[0] - c["kT"]) < 5e-3
    assert abs(ms - c["shear"]) < 0.1 * c["shear"]


def _pooled_nu(corrs, counts, L):
    fa = FourierAutocorr(sp.FOURIER_ENTRIES, sp.FLUID_EVERY * sp.FLUID["mpc_dt"])
    fa.sums = np.sum([np.asarray(cr) * n for cr, n in zip(corrs, counts)], axis=0)
    fa.count = int(np.sum(counts))
    return fa.viscosity(L)


def _jackknife_nu(corrs, counts, L, groups):
    idx = np.array_split(np.arange(len(corrs)), groups)
    full = _pooled_nu(corrs, counts, L)
    loo = []
    for g in idx:
        keep = [i for i in range(len(corrs)) if i not in set(g)]
        loo.append(_pooled_nu([corrs[i] for i in keep], [counts[i] for i in keep], L))
    loo = np.array(loo)
    se = math.sqrt((groups - 1) / groups * np.sum((loo - loo.mean()) ** 2))
    return full, se


def _fluid_batch(R, shear, seed):
    c = sp.FLUID
    g = MPCSystem([make_params(L=(c["L"],) * 3, kT=c["kT"], n_c=c["n_c"], shear=shear, step_update=1,
                               seed=seed)] * R)
    g.initialize()
    return g


def test_viscosity_and_maxwell_agree_with_oracle_and_closed_form():
    ref = _fixture("statistics_fluid_L20.json")
    L = sp.FLUID["L"]
    R = 32
    g = _fluid_batch(R, 0.0, seed=11)
    fas = [FourierAutocorr(sp.FOURIER_ENTRIES, sp.FLUID_EVERY * sp.FLUID["mpc_dt"]) for _ in range(R)]
    temps = []
    t = 0
    n_coll = 20_000
    g.run(0, sp.FLUID_EQUIL, sp.FLUID["mpc_dt"])
    t = sp.FLUID_EQUIL
    while t < n_coll:
        # the sample of step t is taken after MPC::step(t) as in mpc_test.cpp:119-131
        g.run(t, 1, sp.FLUID["mpc_dt"])
        for r in range(R):
            fas[r].sample(g, r)
        if t % 1000 == 0:
            temps.append(np.mean([g.temperature(r) for r in range(R)]))
        g.run(t + 1, sp.FLUID_EVERY - 1, sp.FLUID["mpc_dt"])
        t += sp.FLUID_EVERY
    g.sync()
    nu_g, se_g = _jackknife_nu([fa.correlation() for fa in fas], [fa.count for fa in fas], L, 8)
    eq = [r for r in ref["runs"] if r["shear"] == 0.0]
    nu_o, se_o = _jackknife_nu([r["corr"] for r in eq], [r["count"] for r in eq], L, len(eq))
    _agree("nu", nu_g, se_g, nu_o, se_o)
    nu_c = srd_viscosity()
    print(f"  nu closed form {nu_c:.5f}", flush=True)
    assert abs(nu_g - nu_c) < 0.05 * nu_c
    assert abs(np.mean(temps) - 1.0) < 2e-3
    # Maxwell's speed distribution at kT = 1 (unit mass)
    w, nb = 0.05, 160
    counts = np.sum([g.speed_histogram(w, nb, r) for r in range(R)], axis=0).astype(float)
    edges = np.arange(nb + 1) * w
    cdf = lambda v: math.erf(v / math.sqrt(2)) - math.sqrt(2 / math.pi) * v * math.exp(-v * v / 2)
    p = np.diff([cdf(e) for e in edges])
    n = counts.sum()
    exp = n * p
    sel = exp > 20
    chi2 = np.sum((counts[:nb][sel] - exp[sel]) ** 2 / exp[sel])
    dof = int(sel.sum()) - 1
    print(f"  maxwell chi2/dof {chi2 / dof:.3f} over {dof} bins", flush=True)
    assert chi2 / dof < 2.0


def test_shear_profile_agrees_with_oracle_and_imposed_rate():
    ref = _fixture("statistics_fluid_L20.json")
    R = 16
    g = _fluid_batch(R, sp.SHEAR_RATE, seed=13)
    profs = [VelocityProfile(sp.SHEAR_WIDTH) for _ in range(R)]
    n_coll = 5_000
    t = 0
    while t < n_coll:
        g.run(t, 1, sp.FLUID["mpc_dt"])
        if t >= sp.FLUID_EQUIL:
            for r in range(R):
                profs[r].sample(g, r)
        g.run(t + 1, sp.FLUID_EVERY - 1, sp.FLUID["mpc_dt"])
        t += sp.FLUID_EVERY
    g.sync()
    slopes = [p.shear_rate() for p in profs]
    ms, ss = sp.summary(slopes)
    _agree("slope", ms, ss, ref["slope_mean"], ref["slope_se"])
    assert abs(ms - sp.SHEAR_RATE) < 0.02 * sp.SHEAR_RATE
