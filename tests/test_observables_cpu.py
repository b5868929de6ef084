"""Pins of the oracle's observers (VelocityX, Fourier_Autocorr,
VelocityDistribution, Box::calculate_patches_new) by hand-worked cases, and
of the Python observers (starpolymersmd_amd.observables) by closed forms on
synthetic samples.  Runs without a GPU."""
import io
import math

import numpy as np
import pytest

import oracle_binding as ob
from starpolymersmd_amd import observables as obs


def fluid_oracle(pos, vel, L=8, shear=0.0, delrx=0.0):
    o = ob.Oracle(L=(L, L, L), n_c=1, shear=shear)
    o.set_fluid(np.asarray(pos, float), np.asarray(vel, float))
    o.delrx = delrx
    return o


def test_velocity_profile_hand_case():
    # three particles, no star (COM zero), width 0.5
    pos = [[0.2, -1.3, 0.0], [0.4, -1.1, 1.0], [-2.6, 3.9, 0.0]]
    vel = [[1.0, 2.0, 0.0], [3.0, -1.0, 0.0], [0.5, 0.25, 9.0]]
    o = fluid_oracle(pos, vel)
    (kx, sx), (ky, sy) = o.velocity_profile(0.5)
    # x keys: floor(0.2/.5)*.5 = 0 (two particles), floor(-2.6/.5)*.5 = -3.0
    assert list(kx) == [-3.0, 0.0]
    np.testing.assert_array_equal(sx, [[0.5, 0.25, 1], [4.0, 1.0, 2]])
    # y keys: -1.5 (both first particles: -1.3 and -1.1), 3.5
    assert list(ky) == [-1.5, 3.5]
    np.testing.assert_array_equal(sy, [[4.0, 1.0, 2], [0.5, 0.25, 1]])


def test_velocity_profile_lees_edwards_image():
    # a particle beyond +L/2 in y is wrapped down with vx -= shear * L and
    # x -= delrx (MPC::LEBC(pos, vel), MPC.cpp:431-439)
    L, shear, delrx = 8, 0.1, 1.25
    o = fluid_oracle([[0.3, 4.2, 0.0]], [[0.7, 0.0, 0.0]], L=L, shear=shear, delrx=delrx)
    (kx, sx), (ky, sy) = o.velocity_profile(0.25)
    x = 0.3 - delrx          # -0.95 -> key floor(-3.8) * 0.25 = -1.0
    y = 4.2 - L              # -3.8  -> key floor(-15.2) * 0.25 = -4.0
    assert kx[0] == math.floor(x / 0.25) * 0.25 == -1.0
    assert ky[0] == math.floor(y / 0.25) * 0.25 == -4.0
    assert sx[0][0] == pytest.approx(0.7 - shear * L)


def test_fourier_modes_single_particle():
    L = 8
    pos, vel = [[1.0, -2.0, 3.0]], [[0.3, -0.7, 1.1]]
    o = fluid_oracle(pos, vel, L=L)
    m = o.fourier_modes()
    for j in range(3):
        ph = np.exp(1j * 2 * np.pi / L * pos[0][j])
        want = np.array(vel[0]) * ph
        want[j] = 0.0
        np.testing.assert_allclose(m[j], want, atol=1e-15)


def test_fourier_modes_plane_wave():
    # v_x = cos(2 pi y / L) on a lattice: V_y has Re x-component 1/2
    L = 8
    g = (np.arange(L) - L / 2 + 0.25)
    pos = np.array([[x, y, z] for x in g for y in g for z in g])
    vel = np.zeros_like(pos)
    vel[:, 0] = np.cos(2 * np.pi * pos[:, 1] / L)
    o = fluid_oracle(pos, vel, L=L)
    m = o.fourier_modes()
    assert m[1][0].real == pytest.approx(0.5, abs=1e-12)
    assert abs(m[1][0].imag) < 1e-12
    assert np.abs(m[0]).max() < 1e-12 and np.abs(m[2]).max() < 1e-12


def test_speed_histogram_hand_case():
    vel = [[3.0, 4.0, 0.0], [0.0, 0.0, 0.05], [1.0, 1.0, 1.0], [30.0, 0.0, 0.0]]
    o = fluid_oracle(np.zeros((4, 3)), vel)
    c = o.speed_histogram(0.5, 20)
    # |v| = 5 -> bin 10, 0.05 -> 0, sqrt(3) = 1.732 -> 3, 30 -> beyond
    want = np.zeros(21, np.uint64)
    want[[10, 0, 3, 20]] = 1
    np.testing.assert_array_equal(c, want)


def patch_oracle(arm_pairs, arms=6, A=2, B=2, sep=1.2):
    """star with arms far apart (straight, radius ~ 3 + 10 k) except the
    listed arm pairs, whose last B beads are placed sep apart"""
    o = ob.Oracle(L=(64, 64, 64), n_c=0)
    o.add_star(A, B, arms, 1.0)
    n = o.n_mono
    per = A + B
    pos = np.zeros((n, 3))
    for a in range(arms):
        for k in range(1, per + 1):
            pos[a * per + k] = [40.0 * a, 0.0, 2.0 * k]
    for (i, k) in arm_pairs:
        pos[k * per + per] = pos[i * per + per] + [sep, 0.0, 0.0]
    o.set_monomers(pos=pos)
    return o


@pytest.mark.parametrize("pairs,want", [
    ([], (0.0, 0.0)),
    ([(0, 1)], (1.0, 2.0)),
    ([(0, 1), (2, 3)], (2.0, 2.0)),
    ([(0, 1), (1, 2), (4, 5)], (2.0, 2.5)),      # chains 0-1-2 and 4-5
    ([(5, 0), (4, 2)], (2.0, 2.0)),              # later arm bonded to an earlier one
])
def test_patches_hand_cases(pairs, want):
    # a chain i - k moves arm k's tip next to arm i's current tip
    o = patch_oracle(pairs)
    assert o.patches() == want


def test_patches_energy_sign():
    # B-B pair energy (lambda 1) is negative only for 1 < r^2 < 2.25
    for sep, bonded in [(0.99, False), (1.01, True), (1.49, True), (1.51, False)]:
        o = patch_oracle([(0, 1)], sep=sep)
        assert o.patches() == ((1.0, 2.0) if bonded else (0.0, 0.0)), sep


# ------------------------------------------------------ Python observers

def test_srd_viscosity_closed_form():
    # the value results/plot_MPC_tests.py:107 draws (0.8705) to 1e-3
    assert obs.srd_viscosity() == pytest.approx(0.8705, abs=1e-3)
    assert obs.srd_viscosity() == pytest.approx(0.87002, abs=2e-5)


def test_fourier_autocorr_window_semantics():
    # scalar series in the x-direction's y component; lag sums from the
    # sliding window equal the direct sum over origins
    rng = np.random.default_rng(3)
    x = rng.normal(size=40) + 1j * rng.normal(size=40)
    fa = obs.FourierAutocorr(entries=5, dt=1.0)
    for v in x:
        m = np.zeros((3, 3), complex)
        m[0, 1] = v
        fa.add(m)
    direct = [np.mean([np.real(np.conj(x[t]) * x[t + i]) for t in range(0, 36)]) for i in range(5)]
    np.testing.assert_allclose(fa.correlation()[0], direct, rtol=1e-12)
    assert fa.count == 36


def test_fourier_autocorr_recovers_decay_rate():
    # complex Ornstein-Uhlenbeck modes with decay nu k^2, L = 20, nu = 0.87
    L, nu, dt = 20.0, 0.87, 1.0
    lam = nu * (2 * np.pi / L) ** 2
    a = np.exp(-lam * dt)
    rng = np.random.default_rng(7)
    n = 60000
    fa = obs.FourierAutocorr(entries=50, dt=dt)
    z = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    s = np.sqrt(1 - a * a)
    for _ in range(n):
        z = a * z + s * (rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))
        fa.add(z)
    assert fa.viscosity(L) == pytest.approx(nu, rel=0.03)


def test_velocity_profile_shear_rate_and_format():
    vp = obs.VelocityProfile(width=0.5)
    keys = np.arange(-10, 10) * 0.5
    y = keys + 0.25
    sums = np.stack([0.05 * y * 100, np.zeros_like(y), np.full_like(y, 100)], axis=1)
    sums[3] = 0                        # an empty bin stays out of the map
    vp.add([(keys, sums), (keys, sums)])
    vp.add([(keys, sums), (keys, sums)])
    assert vp.shear_rate() == pytest.approx(0.05, rel=1e-12)
    k, vx, vy, n = vp.table(1)
    assert len(k) == 19 and n[0] == 200
    buf = io.StringIO()
    vp.print_result(buf)
    lines = buf.getvalue().split("\n")
    assert lines[0] == "%.8e %.8e %.8e %.8e" % (-5.0, 0.05 * -4.75, 0.0, 200.0)
    assert lines[19] == ""             # blank line between the maps


def test_statistics_row_format():
    # operator<< on double: 6 significant digits, shortest of %e/%f
    row = [20.2827401, 151.07, 1.0, 2.0, 2.5, 4.138329, 1e-5, -0.5] + [0.0] * 10
    s = obs.statistics_row(1500, row)
    assert s.startswith("1500 20.2827 151.07 1 2 2.5 4.13833 1e-05 -0.5 0 ")
    assert s.endswith(" \n") and len(s.split()) == 19
    assert obs.STAT_HEADER.split()[:7] == ["Step", "Epot", "Ekin", "Temp", "n_p", "s_p", "R_gyr"]


class _FakeSystem:
    def __init__(self):
        self.calls = []

    def run(self, t0, n, dt):
        self.calls.append(("run", t0, n))

    def propagate(self, e):
        self.calls.append(("prop", e))

    def step(self, t, dt):
        self.calls.append(("step", t))


def test_run_with_output_visits_the_reference_output_steps():
    fs = _FakeSystem()
    seen = []
    obs.run_with_output(fs, 0, 250, 0.1, 100, seen.append, equil=99)
    assert seen == [100, 200]
    # every step 0..249 is done exactly once, output steps by hand
    done = []
    for c in fs.calls:
        if c[0] == "run":
            done += list(range(c[1], c[1] + c[2]))
        elif c[0] == "step":
            done.append(c[1])
    assert done == list(range(250))
    assert ("prop", True) in fs.calls


# ---------------------------------------------------------------- fluid VACF
def _vacf_oracle():
    import os
    import sys

    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
    import vacf

    return vacf


def test_vacf_restatement_hand_worked():
    """One particle, 3 entries, velocities (1,0,0) .. (4,0,0): the window
    fills on the third call ([1,2,3]: products 1, 2, 3) and the fourth adds
    [2,3,4] (4, 6, 8); Counter = 2, so the average is [2.5, 4, 5.5] and
    value() = (2.5 + 4 + 5.5) dt / 2 / 3 = 2 for dt = 1."""
    vacf = _vacf_oracle()
    o = vacf.VelAutocorrFluid(1, 3, 1.0)
    for k in range(1, 5):
        o.sample(np.array([[k, 0.0, 0.0]]))
    assert o.counter == 2
    assert np.array_equal(o.lag_sums(), [5.0, 8.0, 11.0])
    o.calculate_average()
    assert np.allclose(o.average, [2.5, 4.0, 5.5])
    assert o.value() == pytest.approx(2.0)


def test_vacf_restatement_particles_apart_and_signs():
    """Two particles with constant and alternating velocities: the first's
    products are |v|^2 at every lag, the second's (-1)^lag |v|^2; nothing
    mixes between particles, and no sum is made before the windows fill."""
    vacf = _vacf_oracle()
    o = vacf.VelAutocorrFluid(2, 4, 0.5)
    a = np.array([1.0, 2.0, 2.0])
    for k in range(7):
        o.sample(np.array([a, (-1) ** k * np.array([0.0, 3.0, 4.0])]))
        if k < 3:
            assert o.counter == 0 and not o.autocorrelation.any()
    assert o.counter == 4
    assert np.allclose(o.autocorrelation[0], 4 * 9.0)
    assert np.allclose(o.autocorrelation[1], 4 * 25.0 * np.array([1, -1, 1, -1]))


def test_mpc_test_file_tag():
    from starpolymersmd_amd.observables import mpc_test_tag

    assert mpc_test_tag(1250, 0.5, 0.1, False) == "_N1250_Shear0.5_h0.1"
    assert mpc_test_tag(80000, 0.0, 0.05, True) == "_N80000_Shear0_h0.05_AMC"
