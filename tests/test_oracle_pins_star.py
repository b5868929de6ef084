"""Closed-form pins of the oracle's star observers, check_bounds recentring
and temperature estimator (CPU only).

Each case is built so that the quantity has a value fixed by the
mathematics or by the reference's stated semantics, and so that a
plausible slip (a transposed tensor, a missing mass, the monomers' kinetic
energy counted, the commented-out shear-frame correction restored, a wrong
wrap sign) changes the answer:

- Box::calculate_rotation_frequency (Box.cpp:355-379): a rigid rotation
  v = V + w0 x (r - c) gives omega = w0 exactly.
- Box::calculate_gyration_tensor (Box.cpp:330-353): six beads at +-a x,
  +-b y, +-c z around a core give diag(2a^2, 2b^2, 2c^2)/7, and a rotated,
  translated copy Q G Q^T.
- MPC::check_bounds (MPC.cpp:160-190): a monomer at L/2 - 0.4 recentres;
  monomers move by -COM and -V_com, the solvent by the same and is
  Lees-Edwards re-wrapped, its x velocity gets no -COM_y * shear term
  (MPC.cpp:185 is commented out).
- MPC::calculateCurrentTemperature (MPC.cpp:384-418): the thermostat sets a
  cell's relative kinetic energy, monomers included and mass-weighted, to
  the Gamma variate (MPC.cpp:273-299); the estimator counts the fluid part
  only, over N_fluid minus the cells holding fluid.
"""
import numpy as np
import pytest

import oracle_binding as ob


def _star_oracle(L, arms, A, B, mass=10.0, n_c=0, shear=0.0, rng_mode=ob.RNG_INJECTED):
    o = ob.Oracle(L=L, n_c=n_c, shear=shear, step_update=1, rng_mode=rng_mode, seed=1)
    o.add_star(A, B, arms, mass)
    return o


def _rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_rotation_frequency_of_rigid_rotation():
    o = _star_oracle((40, 40, 40), arms=5, A=2, B=2)
    n = o.n_mono
    rng = np.random.default_rng(4)
    r = rng.normal(scale=3.0, size=(n, 3)) + np.array([1.5, -2.0, 0.7])
    w0 = np.array([0.31, -0.72, 1.13])
    V = np.array([0.5, 0.2, -0.1])
    c = r.mean(axis=0)
    v = V + np.cross(w0, r - c)
    o.set_monomers(r, v)
    w = o.rotation_frequency()
    assert np.abs(w - w0).max() <= 1e-12 * np.abs(w0).max(), (w, w0)
    # a rotation about another axis is told apart
    o.set_monomers(r, V + np.cross(w0[[1, 2, 0]], r - c))
    assert np.abs(o.rotation_frequency() - w0).max() > 0.1


def test_gyration_tensor_closed_form():
    o = _star_oracle((40, 40, 40), arms=3, A=1, B=1)
    assert o.n_mono == 7
    a, b, c = 1.3, 0.7, 2.1
    pts = np.array([[0, 0, 0], [a, 0, 0], [-a, 0, 0], [0, b, 0], [0, -b, 0], [0, 0, c], [0, 0, -c]], float)
    G0 = np.diag([2 * a * a, 2 * b * b, 2 * c * c]) / 7.0
    o.set_monomers(pts, np.zeros_like(pts))
    rg, G = o.gyration()
    assert np.abs(G - G0).max() <= 1e-15
    assert abs(rg - np.sqrt(np.trace(G0))) <= 1e-15
    Q = _rotation(3)
    o.set_monomers(pts @ Q.T + np.array([3.0, -1.0, 2.5]), np.zeros_like(pts))
    rg, G = o.gyration()
    assert np.abs(G - Q @ G0 @ Q.T).max() <= 1e-13
    assert abs(rg - np.sqrt(np.trace(G0))) <= 1e-13
    # the off-diagonal entries are not symmetric by accident: G_xy of the
    # rotated shape is non-zero and equals G_yx
    assert abs(G[0, 1]) > 1e-2 and G[0, 1] == G[1, 0]


def test_check_bounds_recentring_closed_form():
    """L = 10, shear 0.05 (shear * Ly = 0.5), delrx 1.3 before the call;
    MPC::step first advances delrx by shear * Ly * dt (MPC.cpp:117), so
    the wraps of this collision use 1.35.  Grid shift 0; every particle and
    monomer sits alone in its cell, so the collision leaves velocities
    unchanged and the state after the step is fixed by the recentring and
    the streaming alone."""
    Lb, shear, dt = 10.0, 0.05, 0.1
    o = _star_oracle((10, 10, 10), arms=2, A=1, B=1, shear=shear)
    mp = np.array([[4.6, 0.5, 0.5], [2.5, 0.5, 0.5], [0.5, 0.5, 0.5], [-1.5, 2.5, 0.5], [0.5, -2.5, -1.5]])
    mv = np.array([[0.2, -0.1, 0.05], [0.1, 0.3, -0.2], [-0.4, 0.0, 0.1], [0.05, 0.02, 0.3], [0.3, -0.2, -0.1]])
    o.set_monomers(mp, mv)
    com, cv = mp.mean(axis=0), mv.mean(axis=0)   # equal masses
    fp = np.array([[-3.0, -4.0, -3.0],    # no wrap
                   [1.2, -4.9, 3.5],      # y - COM_y < -5: Lees-Edwards image
                   [-4.0, 3.0, -3.5]])    # x - COM_x < -5: plain x wrap
    fv = np.array([[0.3, 0.1, -0.2], [0.2, 0.3, 0.1], [-0.1, -0.2, 0.25]])
    o.set_fluid(fp, fv)
    o.delrx = 1.3
    ncell = 1000
    o.inject(np.zeros(3), np.full(2 * (ncell + 1), 0.37), np.full(ncell + 1, 5.0))
    assert o.step(0, dt) == 0
    assert o.last_recentered()
    delrx = 1.3 + shear * Lb * dt
    assert o.delrx == pytest.approx(delrx, abs=1e-15)
    pm, vm, _ = o.get_monomers()
    assert np.abs(pm - (mp - com)).max() <= 1e-15
    assert np.abs(vm - (mv - cv)).max() <= 1e-14
    # expected solvent state, each wrap written out
    u = fv - cv                               # -V_com, no shear-frame term in x
    x = fp - com
    x[1, 0] += delrx                          # cy = -1: x -= cy*delrx, y -= cy*Ly, vx -= cy*shear*Ly
    x[1, 1] += Lb
    u[1, 0] += shear * Lb
    x[2, 0] += Lb                             # plain x image
    x = x + u * dt                            # MPC::streaming, no further wrap
    assert np.all(np.abs(x) < Lb / 2)
    pf, vf = o.get_fluid()
    assert np.abs(pf - x).max() <= 1e-14, pf - x
    assert np.abs(vf - u).max() <= 1e-14, vf - u
    # the commented-out correction would have moved vx by -COM_y * shear
    assert abs(com[1] * shear) > 1e-3


def test_temperature_counts_fluid_energy_about_mass_weighted_cmv():
    """One cell holds two fluid particles and a monomer of mass 10; three
    more fluid particles and the other two monomers sit alone.  After the
    collision the cell's relative kinetic energy, monomer included, equals
    the injected Gamma variate G; the estimator is then
    T = 2 E_fluid / (3 (N_fluid - filled)) = (2/3) (G - 5 |v_m - CMV|^2),
    with CMV = (v_1 + v_2 + 10 v_m)/12 conserved by the collision."""
    L = (6, 6, 6)
    o = _star_oracle(L, arms=1, A=1, B=1)
    assert o.n_mono == 3
    mp = np.array([[0.5, 0.5, 0.5], [-1.5, 1.5, 0.5], [1.5, -1.5, -1.5]])
    mv = np.array([[0.3, -0.2, 0.1], [0.1, 0.1, 0.1], [-0.2, 0.3, 0.0]])
    o.set_monomers(mp, mv)
    fp = np.array([[0.3, 0.6, 0.4], [0.7, 0.3, 0.6],        # with monomer 0, cell (3,3,3)
                   [-2.5, -2.5, -2.5], [2.5, 0.5, -0.5], [-0.5, -1.5, 2.5]])
    fv = np.array([[1.0, -0.5, 0.2], [-0.8, 0.6, 0.4],
                   [0.1, 0.2, 0.3], [-0.3, 0.1, 0.2], [0.2, -0.1, -0.3]])
    o.set_fluid(fp, fv)
    ncell = 216
    G = 7.3
    o.inject(np.zeros(3), np.random.default_rng(2).random(2 * (ncell + 1)), np.full(ncell + 1, G))
    assert o.step(0, 0.1) == 0
    assert not o.last_recentered()
    fc, mc = o.cell_index()
    cell = 3 + 6 * 3 + 36 * 3
    assert fc[0] == fc[1] == mc[0] == cell
    assert len(set(fc[2:]) | set(mc[1:]) | {cell}) == 6
    cmv = (fv[0] + fv[1] + 10.0 * mv[0]) / 12.0
    _, vm, _ = o.get_monomers()
    _, vf = o.get_fluid()
    e_fluid = 0.5 * ((vf[:2] - cmv) ** 2).sum()
    e_mono = 5.0 * ((vm[0] - cmv) ** 2).sum()
    assert e_fluid + e_mono == pytest.approx(G, rel=1e-13)
    T = o.temperature()
    assert T == pytest.approx(2.0 * (G - e_mono) / 3.0, rel=1e-13)
    # the monomer's share is far above the tolerance, so counting it would show
    assert e_mono > 1e-3 * G
