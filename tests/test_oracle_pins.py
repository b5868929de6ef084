"""Pins of the FP64 CPU oracle (runs without a GPU).

Each check is chosen so that a plausible mistake in the restatement -- a
dropped term, a wrong sign, a transposed index -- fails it: published
known-answer vectors, the reference's own compiled potentials and RNG
(oracle/_ref, built from star-polymers/lib/Molecules/Potentials.h and
star-polymers/lib/Rand.cpp), closed forms and conservation laws.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_binding as ob

HERE = os.path.dirname(os.path.abspath(__file__))

# ------------------------------------------------------------------ Philox

# Random123 known-answer vectors for philox4x32 with 10 rounds
RANDOM123_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", RANDOM123_KAT)
def test_philox_random123_kat(ctr, key, out):
    assert tuple(int(x) for x in ob.philox(ctr, key)) == out


def test_philox_matches_rocrand_fixture():
    """tests/golden/philox_kat.json: rocRAND's ten_rounds (make_philox_kat.cpp)."""
    for c in json.load(open(os.path.join(HERE, "golden", "philox_kat.json"))):
        assert list(ob.philox(c["ctr"], c["key"])) == c["out"]


# -------------------------------------------------------------- potentials

R2 = np.concatenate([np.linspace(0.64, 2.3, 400), [1.0594631 - 1e-9, 1.0594631, 2.25 - 1e-12, 2.25]])


@pytest.mark.parametrize("lam", [0.5, 1.0, 1.15])
def test_potentials_bit_identical_to_reference_source(lam):
    ref = ob.ref_lib()
    if ref is None:
        pytest.skip("reference harness not built (no reference tree)")
    pot = ob.potentials()
    for r2 in R2:
        assert pot["aa_force"](r2) == ref.ref_aa_force(r2)
        assert pot["aa_potential"](r2) == ref.ref_aa_potential(r2)
        assert pot["bb_force"](r2, lam) == ref.ref_bb_force(r2, lam)
        assert pot["bb_potential"](r2, lam) == ref.ref_bb_potential(r2, lam)
    for r2 in np.linspace(0.05, 2.2, 200):
        assert pot["fene_force"](r2) == ref.ref_fene_force(r2)
        assert pot["fene_potential"](r2) == ref.ref_fene_potential(r2)
    for r2 in np.linspace(0.05, 20.0, 200):
        assert pot["fene_anchor_force"](r2) == ref.ref_fene_anchor_force(r2)
        assert pot["fene_anchor_potential"](r2) == ref.ref_fene_anchor_potential(r2)


def _fd_force(U, r, h=1e-6):
    # force magnitude over r: -(dU/dr)/r
    return -(U((r + h) ** 2) - U((r - h) ** 2)) / (2 * h) / r


def test_forces_are_minus_gradient_of_potentials():
    pot = ob.potentials()
    for r in np.linspace(0.9, 1.02, 25):
        assert math.isclose(pot["aa_force"](r * r), _fd_force(pot["aa_potential"], r), rel_tol=1e-6)
    for r in np.linspace(1.04, 1.48, 25):
        f = pot["bb_force"](r * r, 1.1)
        assert math.isclose(f, _fd_force(lambda x: pot["bb_potential"](x, 1.1), r), rel_tol=1e-5, abs_tol=1e-9)
    for r in np.linspace(0.5, 1.45, 25):
        assert math.isclose(pot["fene_force"](r * r), _fd_force(pot["fene_potential"], r), rel_tol=1e-7)
    for r in np.linspace(0.5, 4.4, 25):
        assert math.isclose(pot["fene_anchor_force"](r * r), _fd_force(pot["fene_anchor_potential"], r),
                            rel_tol=1e-7)


def test_potential_closed_forms():
    pot = ob.potentials()
    rc2 = 2 ** (1 / 12)  # minimum of the 48-24 core: r^-24 = 1/2
    # WCA-type core vanishes (with its +1 shift) and has zero force at 2^(1/24)
    assert abs(pot["aa_potential"](rc2 * (1 - 1e-12))) < 1e-9
    assert abs(pot["aa_force"](rc2 * (1 - 1e-12))) < 1e-8
    assert pot["aa_potential"](1.06) == 0.0 and pot["aa_force"](1.06) == 0.0
    # B-B: depth -lambda at the core minimum; tail lambda*4(r^-48-r^-24-5.93e-5)
    lam = 0.8
    assert math.isclose(pot["bb_potential"](rc2 * (1 - 1e-12), lam), -lam, abs_tol=1e-8)
    r2 = 1.8
    rm24 = r2 ** -12
    assert math.isclose(pot["bb_potential"](r2, lam), 4 * lam * (rm24 * rm24 - rm24 - 0.0000593), rel_tol=1e-14)
    assert pot["bb_potential"](2.26, lam) == 0.0 and pot["bb_force"](2.26, lam) == 0.0
    # FENE: k R0^2/2 = 15*2.25, R0^2 = 2.25 / anchor 20.25
    assert math.isclose(pot["fene_potential"](1.0), -15 * 2.25 * math.log(1 - 1 / 2.25), rel_tol=1e-15)
    assert math.isclose(pot["fene_anchor_force"](4.0), -30 / (1 - 4 / 20.25), rel_tol=1e-15)


# --------------------------------------------------------------- reference RNG


def test_reference_rng_stream_equals_reference_rand():
    """The oracle's reference mode draws exactly what Rand::real_uniform /
    Rand::real_gamma produce from mt19937_64(1) (Rand.cpp:6-43)."""
    ref = ob.ref_lib()
    if ref is None:
        pytest.skip("reference harness not built (no reference tree)")
    import ctypes as C
    rng = np.random.default_rng(0)
    ops = rng.integers(0, 2, 400).astype(np.int32)
    a = rng.uniform(1.5, 30.0, 400)
    b = rng.uniform(0.5, 2.0, 400)
    got = ob.reference_rng_sequence(ops, a, b)
    exp = np.empty(400)
    ref.ref_rng_sequence(ops.ctypes.data_as(C.POINTER(C.c_int)), a.ctypes.data_as(C.POINTER(C.c_double)),
                         b.ctypes.data_as(C.POINTER(C.c_double)), 400, exp.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.array_equal(got, exp)


# ------------------------------------------------------------- Gamma sampler


@pytest.mark.parametrize("shape", [1.5, 4.5, 13.5, 30.0])
def test_philox_gamma_distribution(shape):
    """Marsaglia-Tsang + Box-Muller on Philox: moments and a KS test against
    the closed-form Gamma(shape, scale) distribution."""
    from scipy import stats
    scale = 0.7
    x = np.array([ob.philox_gamma(42, 0, c, 3, shape, scale) for c in range(20000)])
    assert abs(x.mean() - shape * scale) < 4 * math.sqrt(shape) * scale / math.sqrt(len(x))
    assert abs(x.var() / (shape * scale * scale) - 1) < 0.05
    assert stats.kstest(x, "gamma", args=(shape, 0, scale)).pvalue > 1e-3


# ---------------------------------------------------------- MPC collision

def _cells(o):
    fc, _ = o.cell_index()
    return fc


def test_cell_index_known_positions():
    """Hand-placed particles in an L=10 box, zero shift, no shear: cell =
    floor(x+5) + 10 floor(y+5) + 100 floor(z+5)."""
    o = ob.Oracle(L=(10, 10, 10), n_c=0, shear=0.0, rng_mode=ob.RNG_INJECTED)
    pos = np.array([[0.3, -4.7, 2.2], [-5.0, -5.0, -5.0], [4.999, 4.999, 4.999], [0.0, 0.0, 0.0],
                    [-0.001, 1e-9, -2.5]])
    o.set_fluid(pos, np.zeros_like(pos))
    o.inject(np.zeros(3), np.full(2000, 0.5), np.ones(1000))
    assert o.step(0, 0.0) == 0
    exp = [5 + 0 * 10 + 7 * 100, 0, 999, 5 + 50 + 500, 4 + 50 + 200]
    assert list(_cells(o)) == exp


def test_lees_edwards_image_cell_and_velocity_jump():
    """A particle above the top face: LE wrap moves it to the bottom,
    shifts x by -delrx and vx by -shear*Ly (MPC.cpp:421-429), and the
    collision sees the image; with no rotation/scaling its velocity is
    unchanged after unshift."""
    L, sh, dl = (10, 10, 10), 0.1, 2.5
    o = ob.Oracle(L=L, n_c=0, shear=sh, rng_mode=ob.RNG_INJECTED, alpha_deg=0.0)
    pos = np.array([[1.2, 4.8, 0.3]])
    vel = np.array([[0.0, 0.5, 0.0]])  # streams across y = +5 in h=1
    o.set_fluid(pos, vel)
    o.delrx = dl
    o.inject(np.zeros(3), np.full(2000, 0.25), np.ones(1000))
    assert o.step(0, 1.0) == 0  # step also advances delrx by sh*Ly*dt first
    p, v = o.get_fluid()
    dl_now = dl + sh * 10 * 1.0
    assert math.isclose(p[0, 1], 4.8 + 0.5 - 10, abs_tol=1e-12)
    assert math.isclose(p[0, 0], 1.2 - dl_now, abs_tol=1e-12)
    assert math.isclose(v[0, 0], -sh * 10, abs_tol=1e-12)
    assert _cells(o)[0] == int(1.2 - dl_now + 5) + 10 * int(-4.7 + 5) + 100 * int(0.3 + 5)


def _cell_stats(o, pos, vel):
    fc = _cells(o)
    n = np.bincount(fc, minlength=o.ncells)
    P = np.stack([np.bincount(fc, vel[:, i], minlength=o.ncells) for i in range(3)], 1)
    return fc, n, P


def test_collision_conserves_cell_momentum_and_sets_relative_ke_to_gamma():
    """No shear, zero shift: per cell, the thermostat+rotation conserve the
    momentum and leave relative kinetic energy equal to the injected Gamma
    variate (MPC.cpp:273-299 rescales it to Gamma, the rotation preserves it)."""
    L = (6, 6, 6)
    rng = np.random.default_rng(3)
    o = ob.Oracle(L=L, n_c=10, shear=0.0, rng_mode=ob.RNG_INJECTED)
    n = o.n_fluid
    pos = (rng.random((n, 3)) - 0.5) * 6
    vel = rng.normal(size=(n, 3))
    o.set_fluid(pos, vel)
    g = rng.gamma(10.0, 1.0, o.ncells)
    o.inject(np.zeros(3), rng.random(2 * o.ncells), g)
    assert o.step(0, 0.0) == 0
    p2, v2 = o.get_fluid()
    fc, cnt, P0 = _cell_stats(o, pos, vel)
    P1 = np.stack([np.bincount(fc, v2[:, i], minlength=o.ncells) for i in range(3)], 1)
    assert np.abs(P1 - P0).max() < 1e-12
    cmv = P1 / np.maximum(cnt, 1)[:, None]
    ek = 0.5 * np.bincount(fc, ((v2 - cmv[fc]) ** 2).sum(1), minlength=o.ncells)
    m = cnt >= 2
    assert np.allclose(ek[m], g[m], rtol=1e-12)
    assert np.array_equal(p2, pos)  # zero dt, zero shift: positions untouched


def test_rotation_angle_and_axis():
    """Gamma variate = current relative KE (no rescaling): each relative
    velocity is rotated by 130 deg about the axis (u_phi, u_theta)."""
    L = (4, 4, 4)
    rng = np.random.default_rng(5)
    o = ob.Oracle(L=L, n_c=3, shear=0.0, rng_mode=ob.RNG_INJECTED)
    n = o.n_fluid
    pos = (rng.random((n, 3)) - 0.5) * 4
    vel = rng.normal(size=(n, 3))
    o.set_fluid(pos, vel)
    fc = np.floor(pos + 2).astype(int) @ np.array([1, 4, 16])
    cnt = np.bincount(fc, minlength=64)
    cmv = np.stack([np.bincount(fc, vel[:, i], minlength=64) for i in range(3)], 1) / np.maximum(cnt, 1)[:, None]
    ek = 0.5 * np.bincount(fc, ((vel - cmv[fc]) ** 2).sum(1), minlength=64)
    axis_u = rng.random(128)
    o.inject(np.zeros(3), axis_u, ek)
    o.step(0, 0.0)
    _, v2 = o.get_fluid()
    phi, cth = 2 * np.pi * axis_u[0::2], 2 * (axis_u[1::2] - 0.5)
    ax = np.stack([np.sqrt(1 - cth ** 2) * np.cos(phi), np.sqrt(1 - cth ** 2) * np.sin(phi), cth], 1)
    d0, d1 = vel - cmv[fc], v2 - cmv[fc]
    a = ax[fc]
    par0, par1 = (d0 * a).sum(1), (d1 * a).sum(1)
    p0, p1 = d0 - par0[:, None] * a, d1 - par1[:, None] * a
    ok = np.linalg.norm(p0, axis=1) > 1e-3
    assert np.allclose(par0, par1, atol=1e-12)
    cosang = (p0 * p1).sum(1)[ok] / (np.linalg.norm(p0, axis=1) * np.linalg.norm(p1, axis=1))[ok]
    assert np.allclose(cosang, math.cos(math.radians(130)), atol=1e-9)
    # right-handed rotation: (p0 x p1) is along +axis
    assert np.all(((np.cross(p0, p1) * a).sum(1))[ok] > 0)


def test_thermostat_drives_temperature_to_kT():
    """MBS cell thermostat: relative KE per cell ~ Gamma(3(n-1)/2, kT), so
    the temperature estimator (MPC.cpp:405-409) reads kT."""
    o = ob.Oracle(L=(8, 8, 8), n_c=10, shear=0.0, kT=1.7, rng_mode=ob.RNG_PHILOX, seed=9)
    o.initialize()
    Ts = []
    for t in range(30):
        o.step(t, 0.1)
        if t >= 5:
            Ts.append(o.temperature())
    assert abs(np.mean(Ts) - 1.7) < 0.02


def test_reference_mode_step_runs_and_draws_shift_in_reference_order():
    o = ob.Oracle(L=(6, 6, 6), n_c=10, shear=0.05, rng_mode=ob.RNG_REFERENCE)
    o.initialize()
    assert o.step(0, 0.1) == 0
    # 6 draws per particle for initialize(), then the shift: first draw -> z
    draws = ob.reference_rng_sequence(np.zeros(6 * o.n_fluid + 3, np.int32), np.zeros(6 * o.n_fluid + 3),
                                      np.zeros(6 * o.n_fluid + 3))
    sh = o.last_shift()
    assert sh[2] == draws[-3] - 0.5 and sh[1] == draws[-2] - 0.5 and sh[0] == draws[-1] - 0.5


# ------------------------------------------------------------------- MD


def test_star_topology_and_geometry():
    o = ob.Oracle(L=(30, 30, 30), n_c=0, rng_mode=ob.RNG_PHILOX)
    o.add_star(3, 4, 5, 10.0)
    pos, vel, _ = o.get_monomers()
    m, t, a = o.monomer_props()
    assert o.n_mono == 5 * 7 + 1
    b = o.bonds()
    assert len(b) == 5 * 7
    assert a[0] == 1 and a[1:].sum() == 0 and np.all(m == 10.0)
    # per arm: 3 A then 4 B beads
    assert list(t[1:8]) == [0, 0, 0, 1, 1, 1, 1]
    # bonds: core -> first beads (anchor), bead -> next within an arm
    core_bonds = sorted(int(x) for x in b[b[:, 0] == 0][:, 1])
    assert core_bonds == [1, 8, 15, 22, 29]
    # geometry: arm beads on a ray from the core at 2.0, 3.01, 4.02, ...
    rel = pos - pos[0]
    for arm in range(5):
        seg = rel[1 + 7 * arm: 8 + 7 * arm]
        r = np.linalg.norm(seg, axis=1)
        assert np.allclose(r, 2.0 + 1.01 * np.arange(7), atol=1e-12)
        u = seg / r[:, None]
        assert np.allclose(u, u[0], atol=1e-12)
    # COM at the origin, kinetic energy 3N kT/2 about the COM velocity
    assert np.abs(pos.mean(0)).max() < 1e-12
    assert math.isclose(o.ekin_monomers(), 1.5 * 36 * 1.0, rel_tol=1e-12)


def _md_star(dt=0.001, lam=1.0):
    o = ob.Oracle(L=(40, 40, 40), n_c=0, rng_mode=ob.RNG_PHILOX, md_dt=dt, lam=lam, seed=4)
    o.add_star(4, 4, 6, 1.0)
    assert o.md_init() == 0
    return o


def test_md_newton_third_law_and_energy_conservation():
    o = _md_star(dt=0.0005)
    _, _, f = o.get_monomers()
    assert np.abs(f.sum(0)).max() < 1e-9 * np.abs(f).max()
    E = []
    for s in range(2000):
        o.propagate(calc_epot=True)
        if s % 200 == 199:
            E.append(o.epot() + 0.5 * (o.get_monomers()[1] ** 2).sum())
    E = np.array(E)
    assert np.abs(E - E[0]).max() < 2e-3 * abs(E[0])


def test_md_time_reversible():
    o = _md_star(dt=0.001)
    p0, v0, _ = o.get_monomers()
    for _ in range(300):
        o.propagate()
    p1, v1, _ = o.get_monomers()
    o.set_monomers(p1, -v1)
    o.md_init()
    for _ in range(300):
        o.propagate()
    p2, v2, _ = o.get_monomers()
    assert np.abs(p2 - p0).max() < 1e-8
    assert np.abs(v2 + v0).max() < 1e-7


# ------------------------------------------- angular-momentum-conserving SRD

@pytest.mark.parametrize("amc", [True, False])
def test_amc_collision_conserves_cell_angular_momentum(amc):
    """MPC::collide with angular_mom (MPC.cpp:258-269, 344-374): the MBS
    thermostat scales a cell's relative velocities by s = sqrt(Gamma/E_rel),
    after which the rotation of every cell of more than two members must
    keep the cell's momentum and its angular momentum about the centre of
    mass, s L_0; without the flag the rotation changes it."""
    L, n_c, h = 4, 5, 0.1
    rng = np.random.default_rng(11)
    n = n_c * L ** 3
    pos = (rng.random((n, 3)) - 0.5) * L
    vel = rng.normal(size=(n, 3))
    o = ob.Oracle(L=(L, L, L), n_c=n_c, shear=0.0, rng_mode=ob.RNG_INJECTED, amc=amc)
    o.set_fluid(pos, vel)
    shift = rng.random(3) - 0.5
    gamma = rng.gamma(6.0, 1.0, L ** 3)
    o.inject(shift, rng.random(2 * L ** 3), gamma)
    assert o.step(0, h) == 0
    cells, _ = o.cell_index()
    p2, v2 = o.get_fluid()
    xs = p2 + shift                       # positions during the collision
    xs -= L * np.round(xs / L)
    dL_max, dp_max, n_checked = 0.0, 0.0, 0
    for c in np.unique(cells):
        m = cells == c
        if m.sum() <= 2:
            continue
        n_checked += 1
        cmv = vel[m].mean(0)
        r = xs[m] - xs[m].mean(0)
        u0 = vel[m] - cmv
        s_mbs = np.sqrt(gamma[c] / (0.5 * (u0 ** 2).sum()))
        dp_max = max(dp_max, np.abs(vel[m].sum(0) - v2[m].sum(0)).max())
        l_want = s_mbs * np.cross(r, u0).sum(0)
        l_got = np.cross(r, v2[m] - cmv).sum(0)
        dL_max = max(dL_max, np.abs(l_want - l_got).max())
    assert n_checked > 40
    assert dp_max < 1e-10
    if amc:
        assert dL_max < 1e-9
    else:
        assert dL_max > 1e-2


def test_oracle_threads_do_not_change_results():
    """The oracle's solvent loops run on OpenMP threads in the Philox mode;
    cells and particles are independent there, so one thread and many give
    the same state bit for bit (star, shear and the angular-momentum
    collision included)."""
    states = []
    for n in (1, 4):
        ob.set_threads(n)
        o = ob.Oracle(L=(14, 12, 16), n_c=10, shear=0.05, step_update=5, seed=9, amc=True)
        o.add_star(2, 2, 6, 10.0)
        assert o.md_init() == 0
        o.initialize()
        for t in range(26):
            o.propagate()
            assert o.step(t, 0.1) == 0
        p, v = o.get_fluid()
        pm, vm = o.get_monomers()[:2]
        states.append((p.copy(), v.copy(), pm.copy(), vm.copy()))
    ob.set_threads(0)
    assert all(np.array_equal(a, b) for a, b in zip(*states))
