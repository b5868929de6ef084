"""Parity of the HIP MPC step (bin/scan/perm/collide + prologue) with the
FP64 oracle on identical inputs.

Bar: cell indices bit-exact; positions and velocities after the step within
TOL = 1e-5 relative (floored at unit scale), see parity_util.TOL.  The FP32
state contributes ~1e-7; the oracle is FP64 throughout.
"""
import numpy as np
import pytest

import oracle_binding as ob
from parity_util import TOL, le_diff, random_injection, rel_err, seeded_fluid
from starpolymersmd_amd import MPCSystem, make_params

pytestmark = pytest.mark.gpu


def _pair(L, shear, n_c=10, rng_mode=ob.RNG_INJECTED, seed=7, step_update=1):
    g = MPCSystem(make_params(L=L, n_c=n_c, shear=shear, step_update=step_update, seed=seed))
    o = ob.Oracle(L=L, n_c=n_c, shear=shear, step_update=step_update, rng_mode=rng_mode, seed=seed)
    return g, o


def _compare_fluid(g, o, L, shear):
    pg, vg, ids = g.get_fluid(ids=True)
    po, vo = o.get_fluid()
    po, vo = po[ids], vo[ids]
    d, dv, ky = le_diff(pg, po, vg, vo, L, o.delrx, shear)
    ep, ev = rel_err(d, po).max(), rel_err(dv, vo).max()
    assert np.all(np.sort(ids) == np.arange(len(ids))), "ids are not a permutation"
    return ep, ev, int(np.count_nonzero(ky))


@pytest.mark.parametrize("L,shear,delrx", [((8, 8, 8), 0.0, 0.0), ((8, 8, 8), 0.05, 2.37),
                                           ((20, 20, 20), 0.01, -7.1), ((6, 10, 8), 0.1, 1.5)])
def test_one_collision_injected(L, shear, delrx):
    g, o = _pair(L, shear)
    pos, vel = seeded_fluid(L, 10, shear, seed=3)
    g.set_tracking(True)
    g.set_fluid(pos, vel)
    o.set_fluid(pos.astype(np.float64), vel.astype(np.float64))
    g.set_delrx(delrx)
    o.delrx = delrx
    ncell = int(np.prod(L))
    shift, axis, gamma = random_injection(ncell, seed=11)
    g.inject(shift, axis, gamma)
    o.inject(shift, axis, gamma)
    g.step(0, 0.1)
    assert o.step(0, 0.1) == 0
    g.sync()
    fc_g, _ = g.cell_index()
    fc_o, _ = o.cell_index()
    assert np.array_equal(fc_g, fc_o), f"{np.count_nonzero(fc_g != fc_o)} cell indices differ"
    ep, ev, nimg = _compare_fluid(g, o, L, shear)
    assert ep <= TOL and ev <= TOL, (ep, ev)
    assert abs(g.get_delrx() - o.delrx) == 0.0


@pytest.mark.parametrize("L,shear", [((8, 8, 8), 0.05), ((20, 20, 20), 0.01)])
def test_collisions_philox_streams(L, shear):
    """Same Philox streams on both sides: shift, axes and Gamma variates are
    drawn by the library itself; three collisions, state resynchronised
    before each."""
    g, o = _pair(L, shear, rng_mode=ob.RNG_PHILOX, seed=12345)
    pos, vel = seeded_fluid(L, 10, shear, seed=5)
    g.set_tracking(True)
    g.set_fluid(pos, vel)
    for t in range(3):
        pg, vg, ids = g.get_fluid(ids=True)
        inv = np.empty_like(ids)
        inv[ids] = np.arange(len(ids))
        o.set_fluid(pg[inv].astype(np.float64), vg[inv].astype(np.float64))
        o.delrx = g.get_delrx()
        g.step(t, 0.1)
        assert o.step(t, 0.1) == 0
        g.sync()
        assert np.allclose(g.last_collision()[0], o.last_shift(), rtol=0, atol=0)
        fc_g, _ = g.cell_index()
        fc_o, _ = o.cell_index()
        assert np.array_equal(fc_g, fc_o)
        ep, ev, _ = _compare_fluid(g, o, L, shear)
        assert ep <= TOL and ev <= TOL, (t, ep, ev)


def test_fluid_initialize_matches_oracle():
    """MPC::initialize (MPC.cpp:34-55) with the Philox streams."""
    L = (12, 12, 12)
    g, o = _pair(L, 0.05, rng_mode=ob.RNG_PHILOX, seed=99)
    g.set_tracking(True)
    g.initialize()
    o.initialize()
    pg, vg, ids = g.get_fluid(ids=True)
    po, vo = o.get_fluid()
    assert np.array_equal(ids, np.arange(len(ids)))
    assert np.abs(pg - po).max() <= 4e-6 * 6
    assert np.abs(vg - vo).max() <= 1e-7
    assert abs(vg.astype(np.float64).mean(axis=0)).max() < 1e-6


def test_step_update_cadence_and_delrx():
    """delrx advances on every MPC::step call, collisions only every
    step_update calls (MPC.cpp:116-119)."""
    L = (8, 8, 8)
    g, o = _pair(L, 0.05, rng_mode=ob.RNG_PHILOX, seed=3, step_update=4)
    pos, vel = seeded_fluid(L, 10, 0.05, seed=8)
    g.set_tracking(True)
    g.set_fluid(pos, vel)
    o.set_fluid(pos.astype(np.float64), vel.astype(np.float64))
    for t in range(9):
        g.step(t, 0.1)
        assert o.step(t, 0.1) == 0
    g.sync()
    assert g.get_delrx() == o.delrx
    ep, ev, _ = _compare_fluid(g, o, L, 0.05)
    # three collisions without resync: FP32 vs FP64 cell-membership flips
    # can appear; the bulk must still agree
    assert ep <= 1e-4


def test_temperature_matches_oracle():
    L = (10, 10, 10)
    g, o = _pair(L, 0.02, rng_mode=ob.RNG_PHILOX, seed=21)
    pos, vel = seeded_fluid(L, 10, 0.02, seed=9)
    g.set_tracking(True)
    g.set_fluid(pos, vel)
    o.set_fluid(pos.astype(np.float64), vel.astype(np.float64))
    g.step(0, 0.1)
    o.step(0, 0.1)
    tg, to = g.temperature(), o.temperature()
    assert abs(tg - to) / to < 1e-5, (tg, to)


def test_large_box_invariants():
    """L=64, rho=10 (2,621,440 particles): size-independent properties of one
    collision without shear -- ids remain a permutation, total momentum is
    conserved by the collision, every particle stays in the box."""
    L = (64, 64, 64)
    g = MPCSystem(make_params(L=L, n_c=10, shear=0.0, seed=5))
    g.set_tracking(True)
    g.initialize()
    p0, v0, _ = g.get_fluid(ids=True)
    P0 = v0.astype(np.float64).sum(axis=0)
    for t in range(3):
        g.step(t, 0.1)
    g.sync()
    p1, v1, ids = g.get_fluid(ids=True)
    assert np.array_equal(np.sort(ids), np.arange(len(ids)))
    P1 = v1.astype(np.float64).sum(axis=0)
    assert np.abs(P1 - P0).max() < 1e-6 * len(ids) ** 0.5 + 1e-2
    assert np.all(np.abs(p1) <= 32.0)
    T = g.temperature()
    assert 0.97 < T < 1.03


def test_determinism_two_runs():
    L = (16, 16, 16)

    def run():
        g = MPCSystem(make_params(L=L, n_c=10, shear=0.05, seed=77))
        g.set_tracking(True)
        g.initialize()
        for t in range(5):
            g.step(t, 0.1)
        p, v, ids = g.get_fluid(ids=True)
        inv = np.empty_like(ids)
        inv[ids] = np.arange(len(ids))
        return p[inv], v[inv]

    a, b = run(), run()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
