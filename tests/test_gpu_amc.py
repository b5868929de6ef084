"""The angular-momentum-conserving collision (MPC::collide with angular_mom,
MPC.cpp:258-269, 324-374) against the FP64 oracle, and its invariant on the
device.

Bars: cell indices bit-exact and state within parity_util.TOL of the oracle
fed the same randoms; per cell of more than two members, momentum kept and
angular momentum about the cell's centre of mass equal to s L_0 (the MBS
scale times the incoming one) to FP32 accuracy; replicas with and without
the flag in one batch each match their own oracle."""
import numpy as np
import pytest

import oracle_binding as ob
from parity_util import TOL, le_diff, random_injection, rel_err, seeded_fluid
from starpolymersmd_amd import MPCSystem, make_params

pytestmark = pytest.mark.gpu


def _fluid_state(g, r=0):
    p, v, ids = g.get_fluid(r, ids=True)
    o = np.argsort(ids)
    return p[o], v[o], ids


@pytest.mark.parametrize("L,shear,delrx,star", [((8, 8, 8), 0.0, 0.0, False), ((8, 8, 8), 0.05, 2.37, False),
                                                ((12, 10, 8), 0.02, -1.3, True)])
def test_amc_collision_matches_oracle(L, shear, delrx, star):
    g = MPCSystem(make_params(L=L, n_c=10, shear=shear, step_update=1, seed=7, angular_momentum=1))
    o = ob.Oracle(L=L, n_c=10, shear=shear, step_update=1, rng_mode=ob.RNG_INJECTED, seed=7, amc=True)
    pos, vel = seeded_fluid(L, 10, shear, seed=5)
    g.set_tracking(True)
    g.set_fluid(pos, vel)
    o.set_fluid(pos.astype(np.float64), vel.astype(np.float64))
    if star:
        g.add_star(2, 2, 4, 10.0)
        o.add_star(2, 2, 4, 10.0)
        pm, vm, _ = g.get_monomers()
        o.set_monomers(pm, vm)
        g.md_init()
        assert o.md_init() == 0
    g.set_delrx(delrx)
    o.delrx = delrx
    shift, axis, gamma = random_injection(int(np.prod(L)), seed=13)
    g.inject(shift, axis, gamma)
    o.inject(shift, axis, gamma)
    g.step(0, 0.1)
    assert o.step(0, 0.1) == 0
    g.sync()
    fc_g, mc_g = g.cell_index()
    fc_o, mc_o = o.cell_index()
    assert np.array_equal(fc_g, fc_o) and np.array_equal(mc_g, mc_o)
    pg, vg, ids = g.get_fluid(ids=True)
    po, vo = o.get_fluid()
    d, dv, _ = le_diff(pg, po[ids], vg, vo[ids], L, o.delrx, shear)
    assert rel_err(d, po[ids]).max() <= TOL and rel_err(dv, vo[ids]).max() <= TOL
    if star:
        _, vmg, _ = g.get_monomers()
        _, vmo, _ = o.get_monomers()
        assert np.abs(vmg - vmo).max() <= TOL * max(1.0, np.abs(vmo).max())


def test_amc_keeps_cell_angular_momentum_on_device():
    L, n_c = (8, 8, 8), 10
    g = MPCSystem(make_params(L=L, n_c=n_c, shear=0.0, step_update=1, seed=3, angular_momentum=1))
    pos, vel = seeded_fluid(L, n_c, 0.0, seed=21)
    g.set_tracking(True)
    g.set_fluid(pos, vel)
    shift, axis, gamma = random_injection(int(np.prod(L)), seed=17)
    g.inject(shift, axis, gamma)
    g.step(0, 0.1)
    g.sync()
    p2, v2, _ = _fluid_state(g)
    fc, _ = g.cell_index()
    p2 = p2.astype(np.float64)
    v2 = v2.astype(np.float64)
    v0 = vel.astype(np.float64)
    Lb = np.array(L, float)
    xs = p2 + shift
    xs -= Lb * np.round(xs / Lb)
    n_checked = 0
    for c in np.unique(fc):
        m = fc == c
        if m.sum() <= 2:
            continue
        n_checked += 1
        cmv = v0[m].mean(0)
        r = xs[m] - xs[m].mean(0)
        u0 = v0[m] - cmv
        s = np.sqrt(gamma[c] / (0.5 * (u0 ** 2).sum()))
        assert np.abs(v0[m].sum(0) - v2[m].sum(0)).max() < 1e-4
        l_want = s * np.cross(r, u0).sum(0)
        l_got = np.cross(r, v2[m] - cmv).sum(0)
        assert np.abs(l_want - l_got).max() < 1e-4 * max(1.0, np.abs(l_want).max()), c
    assert n_checked > 400


def test_amc_and_plain_replicas_in_one_batch():
    L = (8, 8, 8)
    ps = [make_params(L=L, n_c=10, shear=0.05, seed=2024, angular_momentum=a) for a in (1, 0)]
    g = MPCSystem(ps)
    g.set_tracking(True)
    ors = []
    for r, amc in enumerate((True, False)):
        pos, vel = seeded_fluid(L, 10, 0.05, seed=60 + r)
        g.set_fluid(pos, vel, r)
        o = ob.Oracle(L=L, n_c=10, shear=0.05, rng_mode=ob.RNG_PHILOX, seed=2024, replica=r, amc=amc)
        o.set_fluid(pos.astype(np.float64), vel.astype(np.float64))
        ors.append(o)
    for t in range(3):
        g.step(t, 0.1)
        g.sync()
        for r, o in enumerate(ors):
            assert o.step(t, 0.1) == 0
            pg, vg, ids = g.get_fluid(r, ids=True)
            po, vo = o.get_fluid()
            d, dv, _ = le_diff(pg, po[ids], vg, vo[ids], L, o.delrx, 0.05)
            assert rel_err(d, po[ids]).max() <= TOL and rel_err(dv, vo[ids]).max() <= TOL, (t, r)
            # continue both from the device state (FP32), as the batch test does
            p_id = np.empty_like(pg)
            v_id = np.empty_like(vg)
            p_id[ids] = pg
            v_id[ids] = vg
            o.set_fluid(p_id.astype(np.float64), v_id.astype(np.float64))
