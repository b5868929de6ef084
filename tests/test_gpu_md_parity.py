"""Parity of the device star builder, Verlet lists, forces and velocity
Verlet (one workgroup per star) with the FP64 oracle, and of a full
collision interval with the star coupled to the solvent.

Bar: star construction bit-identical (same Philox draws, same host FP64
operations); forces within 1e-10 of max|F| (the device sums each monomer's
forces in 64-bit fixed point with 2^-32 resolution, each pair evaluated once
with Newton's third law); MD state after 10 substeps within 1e-9; state
after a collision within parity_util.TOL."""
import numpy as np
import pytest

import oracle_binding as ob
from parity_util import TOL, le_diff, rel_err
from starpolymersmd_amd import MPCSystem, make_params

pytestmark = pytest.mark.gpu

CFG0 = dict(L=(20, 20, 20), n_c=10, shear=0.01, step_update=10, seed=1)


def _pair(**kw):
    c = dict(CFG0, **kw)
    g = MPCSystem(make_params(L=c["L"], n_c=c["n_c"], shear=c["shear"], step_update=c["step_update"],
                              seed=c["seed"], md_dt=0.01, lam=1.0))
    o = ob.Oracle(L=c["L"], n_c=c["n_c"], shear=c["shear"], step_update=c["step_update"],
                  rng_mode=ob.RNG_PHILOX, seed=c["seed"], md_dt=0.01, lam=1.0)
    return g, o


def test_star_build_forces_and_md():
    g, o = _pair()
    g.add_star(5, 5, 10, 10.0)
    o.add_star(5, 5, 10, 10.0)
    pg, vg, _ = g.get_monomers()
    po, vo, _ = o.get_monomers()
    assert pg.shape == (101, 3)
    assert np.array_equal(pg, po) and np.array_equal(vg, vo)
    g.md_init()
    assert o.md_init() == 0
    _, _, fg = g.get_monomers()
    _, _, fo = o.get_monomers()
    assert np.abs(fg - fo).max() <= 1e-10 * max(1.0, np.abs(fo).max())
    for s in range(10):
        g.propagate(calc_epot=(s == 9))
        assert o.propagate(calc_epot=(s == 9)) == 0
    pg, vg, fg = g.get_monomers()
    po, vo, fo = o.get_monomers()
    assert np.abs(pg - po).max() <= 1e-9
    assert np.abs(vg - vo).max() <= 1e-9
    assert abs(g.epot() - o.epot()) <= 1e-9 * abs(o.epot())


def test_interval_with_star_and_solvent():
    """propagate x10 + MPC::step every call (collision at t=0 and t=10):
    the collisions include the monomers with mass 10 (Q5) and the first
    one recentres (arms reach beyond L/2-0.5)."""
    g, o = _pair()
    g.add_star(5, 5, 10, 10.0)
    o.add_star(5, 5, 10, 10.0)
    g.md_init()
    o.md_init()
    g.set_tracking(True)
    g.initialize()
    o.initialize()
    pg, vg, ids = g.get_fluid(ids=True)
    o.set_fluid(pg.astype(np.float64), vg.astype(np.float64))
    for t in range(11):
        if t in (0, 10):  # resync the solvent before each collision
            pg, vg, ids = g.get_fluid(ids=True)
            inv = np.empty_like(ids)
            inv[ids] = np.arange(len(ids))
            o.set_fluid(pg[inv].astype(np.float64), vg[inv].astype(np.float64))
        g.propagate()
        assert o.propagate() == 0
        g.step(t, 0.1)
        assert o.step(t, 0.1) == 0
        if t == 0:
            g.sync()
            assert g.last_collision()[1] == o.last_recentered()
            fg_, mg = g.cell_index()
            fo_, mo = o.cell_index()
            assert np.array_equal(fg_, fo_)
            assert np.array_equal(mg, mo)
    g.sync()
    pg, vg, ids = g.get_fluid(ids=True)
    po, vo = o.get_fluid()
    d, dv, _ = le_diff(pg, po[ids], vg, vo[ids], (20, 20, 20), o.delrx, 0.01)
    assert rel_err(d, po[ids]).max() <= TOL and rel_err(dv, vo[ids]).max() <= TOL
    mpg, mvg, _ = g.get_monomers()
    mpo, mvo, _ = o.get_monomers()
    assert np.abs(mpg - mpo).max() <= TOL
    assert rel_err(mvg - mvo, mvo).max() <= TOL


def test_graph_run_equals_stepwise():
    """spmd_run (collision intervals replayed from a HIP graph) gives the
    same state as the explicit propagate/step call sequence."""
    def make():
        g = MPCSystem(make_params(L=(16, 16, 16), n_c=10, shear=0.02, step_update=10, seed=4, md_dt=0.01))
        g.add_star(4, 4, 6, 10.0)
        g.md_init()
        g.set_tracking(True)
        g.initialize()
        return g

    a, b = make(), make()
    a.run(0, 41, 0.1)
    for t in range(41):
        b.propagate()
        b.step(t, 0.1)
    a.sync()
    b.sync()

    def state(g):
        p, v, ids = g.get_fluid(ids=True)
        inv = np.empty_like(ids)
        inv[ids] = np.arange(len(ids))
        return p[inv], v[inv], g.get_monomers()

    sa, sb = state(a), state(b)
    assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1])
    for x, y in zip(sa[2], sb[2]):
        assert np.array_equal(x, y)
    assert a.get_delrx() == b.get_delrx()


def test_coil_star_builder_and_md():
    """Random-coil star of the L=64 workload (f=20, 15A+15B): topology of
    the reference builder, bonded distances exactly `bond` (anchor bonds
    2.0), non-bonded r >= 1 for almost all pairs, COM at 0, kinetic energy
    3N kT/2, and the device Verlet lists accept it; MD then matches the
    oracle started from the same coordinates."""
    g = MPCSystem(make_params(L=(64, 64, 64), n_c=0, shear=0.05, step_update=10, seed=1, md_dt=0.01))
    g.add_star(15, 15, 20, 10.0, coil=True)
    pos, vel, _ = g.get_monomers()
    o = ob.Oracle(L=(64, 64, 64), n_c=0, shear=0.05, step_update=10, rng_mode=ob.RNG_PHILOX, seed=1, md_dt=0.01)
    o.add_star(15, 15, 20, 10.0)
    b = o.bonds()
    blen = np.linalg.norm(pos[b[:, 0]] - pos[b[:, 1]], axis=1)
    anchor = b[:, 0] == 0
    assert np.allclose(blen[anchor], 2.0, atol=1e-12)
    assert np.all(blen[~anchor] >= 1.01 - 1e-12) and np.all(blen[~anchor] <= 1.01 * 1.3 + 1e-12)
    assert np.mean(np.abs(blen[~anchor] - 1.01) < 1e-12) > 0.95
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=2)
    bonded = np.zeros_like(d, bool)
    bonded[b[:, 0], b[:, 1]] = bonded[b[:, 1], b[:, 0]] = True
    nb = d[~bonded & ~np.eye(len(pos), dtype=bool)]
    assert nb.min() >= 1.0 - 1e-12
    assert np.abs(pos.mean(0)).max() < 1e-12
    v0 = vel - vel.mean(0)
    assert abs(0.5 * 10.0 * (v0 ** 2).sum() - 1.5 * len(pos)) < 1e-9 * len(pos)
    assert np.abs(pos).max() < 20
    g.md_init()
    o.set_monomers(pos, vel)
    assert o.md_init() == 0
    for _ in range(50):
        g.propagate()
        assert o.propagate() == 0
    pg, vg, _ = g.get_monomers()
    po, vo, _ = o.get_monomers()
    assert np.abs(pg - po).max() < 1e-9 and np.abs(vg - vo).max() < 1e-9


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_forces_of_compressed_coil_match_oracle(seed):
    """Forces of Box::calculate_forces_verlet on a coil star squeezed to 95%
    (many non-bonded pairs at every separation, B-B pairs on both sides of
    the 1.5 cutoff): within 1e-10 of max|F| of the FP64 oracle."""
    g = MPCSystem(make_params(L=(64, 64, 64), n_c=0, step_update=10, seed=seed, md_dt=0.01))
    g.add_star(15, 15, 20, 10.0, coil=True)
    pos, vel, _ = g.get_monomers()
    pos = pos * 0.95
    g.set_monomers(pos=pos)
    g.md_init()
    o = ob.Oracle(L=(64, 64, 64), n_c=0, step_update=10, rng_mode=ob.RNG_PHILOX, seed=seed, md_dt=0.01)
    o.add_star(15, 15, 20, 10.0)
    o.set_monomers(pos, vel)
    assert o.md_init() == 0
    _, _, fg = g.get_monomers()
    _, _, fo = o.get_monomers()
    d = np.linalg.norm(pos[:, None] - pos[None], axis=2)
    bb = np.outer(o.monomer_props()[1], o.monomer_props()[1]).astype(bool)
    assert np.sum(bb & (d > 1.5) & (d < 2.0)) > 0 and np.sum(bb & (d > 1.0) & (d < 1.5)) > 0
    assert np.abs(fg - fo).max() <= 1e-10 * np.abs(fo).max()


@pytest.mark.parametrize("arms,A,B", [(16, 10, 10), (21, 13, 12), (30, 20, 20)])
def test_md_large_stars_match_oracle(arms, A, B):
    """Stars of 321, 526 and 1,201 monomers run with one, two and four
    monomers per MD thread (the last is the cfg4 closure star): 50 substeps
    match the FP64 oracle from the same coil coordinates."""
    g = MPCSystem(make_params(L=(64, 64, 64), n_c=0, shear=0.05, step_update=10, seed=5, md_dt=0.01))
    g.add_star(A, B, arms, 10.0, coil=True)
    pos, vel, _ = g.get_monomers()
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=2) + 10 * np.eye(len(pos))
    assert d.min() >= 1.0 - 1e-12  # no overlap, so no blow-up
    o = ob.Oracle(L=(64, 64, 64), n_c=0, shear=0.05, step_update=10, rng_mode=ob.RNG_PHILOX, seed=5, md_dt=0.01)
    o.add_star(A, B, arms, 10.0)
    g.md_init()
    o.set_monomers(pos, vel)
    assert o.md_init() == 0
    _, _, fg = g.get_monomers()
    _, _, fo = o.get_monomers()
    assert np.abs(fg - fo).max() <= 1e-10 * max(1.0, np.abs(fo).max())
    for _ in range(50):
        g.propagate()
        assert o.propagate() == 0
    pg, vg, _ = g.get_monomers()
    po, vo, _ = o.get_monomers()
    assert np.abs(pg - po).max() < 1e-9 and np.abs(vg - vo).max() < 1e-9


def test_star_added_after_md_init_gets_lists_and_forces():
    """Adding a star to one replica after spmd_md_init rebuilds the device
    monomer buffers of all replicas; the lists and forces are made again at
    once (the Thermostat_None ctor's work), so the other replica's forces
    are unchanged and the new star's match the oracle before and after 20
    substeps."""
    prm = [make_params(L=(24, 24, 24), n_c=0, step_update=10, seed=s, md_dt=0.01, lam=1.0) for s in (1, 2)]
    g = MPCSystem(prm)
    g.add_star(5, 5, 10, 10.0, r=0)
    g.md_init()
    _, _, f0 = g.get_monomers(r=0)
    assert np.abs(f0).max() > 1.0
    g.add_star(4, 4, 8, 10.0, r=1, coil=True)
    _, _, f0b = g.get_monomers(r=0)
    assert np.abs(f0b - f0).max() <= 1e-10 * np.abs(f0).max()
    pos, vel, f1 = g.get_monomers(r=1)
    o = ob.Oracle(L=(24, 24, 24), n_c=0, step_update=10, rng_mode=ob.RNG_PHILOX, seed=2, md_dt=0.01, lam=1.0)
    o.add_star(4, 4, 8, 10.0)
    o.set_monomers(pos, vel)
    assert o.md_init() == 0
    _, _, fo = o.get_monomers()
    assert np.abs(fo).max() > 1.0
    assert np.abs(f1 - fo).max() <= 1e-10 * np.abs(fo).max()
    for _ in range(20):
        g.propagate()
        assert o.propagate() == 0
    pg, vg, _ = g.get_monomers(r=1)
    po, vo, _ = o.get_monomers()
    assert np.abs(pg - po).max() < 1e-9 and np.abs(vg - vo).max() < 1e-9
