"""The L=64 star workload as one system: L=64 (262,144 cells), rho=10
(2,621,440 solvent particles), shear 0.05, a 601-monomer coil star (f=20,
15 A + 15 B beads, mass 10), h = 0.1 with 10 velocity-Verlet substeps of
0.01 per collision.

At this box size FP32 positions carry 2^-19 of absolute resolution, so the
FP64 evaluation of x + s + L/2 (MPC.cpp:209) from the FP32 state is what
keeps cell indices exact; the Lees-Edwards images (MPC.cpp:421-429), the
monomers' shifted cells (MPC.cpp:221-235) and the check_bounds recentring
(MPC.cpp:160-190) all run here at their full size.

Bar (parity_util): cell indices of all 2,621,440 particles and the 601
monomers bit-exact against the FP64 oracle; solvent positions/velocities
within TOL = 1e-5 relative (floored at unit scale); monomers within TOL.
The long run checks size-independent properties of the reference's
dynamics over 10,000 collisions replayed from the interval graph.
"""
import numpy as np
import pytest

import oracle_binding as ob
from parity_util import TOL, le_diff, random_injection, rel_err, seeded_fluid
from starpolymersmd_amd import MPCSystem, make_params

pytestmark = pytest.mark.gpu

L = (64, 64, 64)
SHEAR = 0.05
ARMS, A, B, MASS = 20, 15, 15, 10.0


def _system(seed):
    g = MPCSystem(make_params(L=L, n_c=10, shear=SHEAR, step_update=10, seed=seed, md_dt=0.01, lam=1.0))
    g.add_star(A, B, ARMS, MASS, coil=True)
    return g


def _oracle(seed, rng_mode, mpos, mvel):
    o = ob.Oracle(L=L, n_c=10, shear=SHEAR, step_update=10, rng_mode=rng_mode, seed=seed, md_dt=0.01, lam=1.0)
    o.add_star(A, B, ARMS, MASS)
    o.set_monomers(mpos, mvel)
    return o


@pytest.mark.parametrize("rng_mode,recentre", [(ob.RNG_INJECTED, False), (ob.RNG_PHILOX, False),
                                               (ob.RNG_INJECTED, True)])
def test_interval_matches_oracle(rng_mode, recentre):
    """One collision interval (10 x propagate + MPC::step, the collision on
    the first call) from the same state on both sides: injected randoms
    (grid shift, 262,144 axes and Gamma variates) or the library's own
    Philox streams; with `recentre` the star starts with a monomer beyond
    L/2 - 0.5 so that check_bounds recentres solvent and star first."""
    seed = 11
    g = _system(seed)
    mpos, mvel, _ = g.get_monomers()
    assert mpos.shape == (601, 3)
    if recentre:
        # Offsets that are not short binary fractions: with -9.0 a solvent
        # y of exactly 23.0 sits exactly on the y = L/2 face after the
        # recentring, and which image it takes there depends on the last
        # ulp of the COM, while the two MD states agree to ~1e-13 only.
        mpos = mpos + np.array([31.7 - mpos[:, 0].max(), -8.6863, 4.2719])
        g.set_monomers(pos=mpos)
    o = _oracle(seed, rng_mode, mpos, mvel)
    g.md_init()
    assert o.md_init() == 0
    pos, vel = seeded_fluid(L, 10, SHEAR, seed=3)
    assert pos.shape[0] == 2_621_440
    g.set_tracking(True)
    g.set_fluid(pos, vel)
    o.set_fluid(pos.astype(np.float64), vel.astype(np.float64))
    delrx = 13.71
    g.set_delrx(delrx)
    o.delrx = delrx
    if rng_mode == ob.RNG_INJECTED:
        shift, axis, gamma = random_injection(o.ncells, seed=17)
        g.inject(shift, axis, gamma)
        o.inject(shift, axis, gamma)
    for t in range(10):
        g.propagate()
        assert o.propagate() == 0
        g.step(t, 0.1)
        assert o.step(t, 0.1) == 0
        if t == 0:
            g.sync()
            sh, rec, _, _ = g.last_collision()
            assert rec == o.last_recentered() == recentre
            if rng_mode == ob.RNG_PHILOX:
                assert np.array_equal(sh, o.last_shift())
            fc_g, mc_g = g.cell_index()
            fc_o, mc_o = o.cell_index()
            nbad = int(np.count_nonzero(fc_g != fc_o))
            assert nbad == 0, f"{nbad} of {len(fc_o)} solvent cell indices differ"
            assert np.array_equal(mc_g, mc_o), "monomer cell indices differ"
            assert len(np.unique(fc_o)) > 0.99 * o.ncells
    g.sync()
    assert g.get_delrx() == o.delrx
    pg, vg, ids = g.get_fluid(ids=True)
    assert np.array_equal(np.sort(ids), np.arange(len(ids))), "ids are not a permutation"
    po, vo = o.get_fluid()
    d, dv, ky = le_diff(pg, po[ids], vg, vo[ids], L, o.delrx, SHEAR)
    ep, ev = rel_err(d, po[ids]).max(), rel_err(dv, vo[ids]).max()
    assert ep <= TOL and ev <= TOL, (ep, ev)
    mpg, mvg, _ = g.get_monomers()
    mpo, mvo, _ = o.get_monomers()
    assert np.abs(mpg - mpo).max() <= TOL
    assert rel_err(mvg - mvo, mvo).max() <= TOL


def test_ten_thousand_collisions_replayed():
    """The full run length of the workload -- 10,000 collisions (100,000
    MD substeps) from MPC::initialize and the coil star, one interval graph
    replayed per collision -- with the properties the dynamics fixes:
    device status clean (no cell or Verlet-grid violation, no scan
    time-out), ids a permutation, every particle inside the box, y and z
    momentum of solvent + star conserved up to the recentrings' own
    subtraction (-(N_f + M_star) V_com each, MPC.cpp:177-188), the MPC
    temperature right after each checked collision equal to kT plus the
    Lees-Edwards layer term of the reference's estimator within five
    standard errors, the star's kinetic temperature bounded."""
    g = _system(seed=1)
    g.md_init()
    g.set_tracking(True)
    g.initialize()
    n_f = g.num_fluid()
    m_star = MASS * g.num_monomers()

    def momentum():
        _, v, ids = g.get_fluid(ids=True)
        _, mv, _ = g.get_monomers()
        return v.astype(np.float64).sum(axis=0) + MASS * mv.sum(axis=0), ids

    P0, _ = momentum()
    expect = P0.copy()
    n_rec = 0
    n_coll = 0

    def account():
        nonlocal n_rec
        _, rec, _, cv = g.last_collision()
        if rec:
            n_rec += 1
            expect[:] -= (n_f + m_star) * cv

    g.run(0, 1, 0.1)  # the collision of call 0; then whole intervals
    n_coll += 1
    account()
    t = 1
    for chunk in range(10):
        while n_coll < 1000 * (chunk + 1):
            g.run(t, 10, 0.1)  # calls t .. t+9, the collision on the last: one graph replay
            t += 10
            n_coll += 1
            account()
        g.sync()  # raises on any device status bit
        P, ids = momentum()
        assert np.array_equal(np.sort(ids), np.arange(n_f)), chunk
        pos, _ = g.get_fluid()
        assert np.all(np.abs(pos) <= 32.0), chunk
        # FP32 rounding of 2.6M velocities per collision is a random walk of
        # about 5e-5 per collision in each component; a recentring rounds
        # float(v - V_com) with one V_com for every particle, a bias that
        # does not average out (measured 1e-2 at |V_com| = 0.02; at most
        # N ulp/2 ~ 0.1), so each recentring gets that budget
        tol = 0.05 + 0.1 * n_rec
        assert np.abs(P[1:] - expect[1:]).max() < tol, (chunk, n_rec, P, expect)
        # MPC::calculateCurrentTemperature right after a collision uses that
        # collision's shifted cells with the velocities of the unshifted frame
        # (MPC.cpp:138-140 then 384-418): every cell's relative kinetic
        # energy is its Gamma variate (mean 1.5 (n-1) kT), except in the
        # layer of cells the y shift wraps across the Lees-Edwards face, where
        # a fraction p = |s_y| of the members carries vx offset by
        # D = shear * Ly.  A cell with k of n members offset adds D^2/2
        # k (n-k)/n, of mean D^2/2 (n-1) p (1-p); over Poisson(rho) occupancy
        # and Lx Lz cells that gives
        #   T = kT + (2/3) Lx Lz (D^2/2) (rho - 1 + e^-rho) p (1-p) / (N - filled).
        T = g.temperature()
        s_y = abs(g.last_collision()[0][1])
        dof = n_f - 64 ** 3  # N - filled: P(empty cell) = e^-10
        D = SHEAR * 64
        excess = (2.0 / 3.0) * 64 * 64 * 0.5 * D * D * (10 - 1 + np.exp(-10.0)) * s_y * (1 - s_y) / dof
        sig = (2.0 / (3.0 * dof)) ** 0.5 + excess / 64  # MBS estimator; 4,096 layer cells
        assert abs(T - 1.0 - excess) < 5 * sig, (chunk, T, excess)
        mp, mv, _ = g.get_monomers()
        assert np.all(np.isfinite(mp)) and np.all(np.isfinite(mv))
        vrel = mv - mv.mean(axis=0)
        t_star = MASS * (vrel ** 2).sum() / (3 * (len(mv) - 1))
        assert 0.7 < t_star < 1.3, (chunk, t_star)
        print(f"[cfg1] collisions {n_coll}: T={T:.5f} (kT + {excess:.5f} expected) T_star={t_star:.3f} "
              f"recentrings={n_rec}", flush=True)
