"""Device observers against the FP64 oracle on identical states: the
VelocityX profile, the Fourier_Autocorr modes, the VelocityDistribution
histogram, Box::calculate_patches_new and the statistics row.

Bars: bin keys and counts exact (both sides bin the same FP64 values with
the same operations); velocity sums within 1e-6 absolute (the device sums
in 2^-32 fixed point); Fourier modes within 1e-10 (normalised by N);
patches exact; statistics row within 1e-7 relative after one velocity-Verlet
step from the same state (MD parity is 1e-9)."""
import numpy as np
import pytest

import oracle_binding as ob
from starpolymersmd_amd import MPCSystem, SpmdError, make_params
from starpolymersmd_amd import observables as obs

pytestmark = pytest.mark.gpu


def _state(L=20, shear=0.02, star=True, seed=3, collisions=3):
    g = MPCSystem(make_params(L=(L,) * 3, n_c=10, shear=shear, step_update=10, seed=seed))
    if star:
        g.add_star(5, 5, 10, 10.0, coil=True)
    g.md_init()
    g.initialize()
    g.run(0, 10 * collisions + 1, 0.1)
    g.sync()
    pf, vf = g.get_fluid()
    o = ob.Oracle(L=(L,) * 3, n_c=10, shear=shear)
    o.set_fluid(pf.astype(np.float64), vf.astype(np.float64))
    o.delrx = g.get_delrx()
    if star:
        o.add_star(5, 5, 10, 10.0)
        pos, vel, _ = g.get_monomers()
        o.set_monomers(pos, vel)
    return g, o


@pytest.mark.parametrize("star,shear,width", [(True, 0.02, 0.5), (False, 0.0, 0.2), (True, 0.05, 0.25)])
def test_velocity_profile_matches_oracle(star, shear, width):
    g, o = _state(shear=shear, star=star)
    dev = g.velocity_profile(width)
    ref = o.velocity_profile(width)
    for a in range(2):
        keys, sums = dev[a]
        nz = sums[:, 2] > 0
        rk, rs = ref[a]
        assert np.array_equal(keys[nz], rk), a
        assert np.array_equal(sums[nz, 2], rs[:, 2]), a
        assert np.abs(sums[nz, :2] - rs[:, :2]).max() < 1e-6, a
        assert sums[:, 2].sum() == g.num_fluid()


def test_fourier_modes_match_oracle():
    g, o = _state(shear=0.0, star=False)
    m = g.fourier_modes()
    mo = o.fourier_modes()
    assert np.abs(m - mo).max() < 1e-10
    assert np.abs(mo).max() > 1e-4          # not trivially zero


@pytest.mark.parametrize("width,nbins", [(0.01, 1000), (0.1, 40)])
def test_speed_histogram_matches_oracle(width, nbins):
    g, o = _state(shear=0.05, star=True)
    c = g.speed_histogram(width, nbins)
    co = o.speed_histogram(width, nbins)
    assert int(c.sum()) == g.num_fluid()
    assert np.array_equal(c, co)


@pytest.mark.parametrize("side", [3.0, 5.0, 7.0, 10.0])
def test_patches_match_oracle(side):
    rng = np.random.default_rng(int(side * 10))
    g = MPCSystem(make_params(L=(64,) * 3, n_c=1, seed=1))
    g.add_star(5, 5, 10, 10.0)
    o = ob.Oracle(L=(64,) * 3, n_c=1)
    o.add_star(5, 5, 10, 10.0)
    for trial in range(5):
        pos = (rng.random((g.num_monomers(), 3)) - 0.5) * side
        g.set_monomers(pos=pos)
        o.set_monomers(pos=pos)
        assert g.patches() == o.patches(), (side, trial)


def test_patches_reject_more_than_128_arms():
    g = MPCSystem(make_params(L=(64,) * 3, n_c=1, seed=1))
    g.add_star(1, 1, 130, 1.0)
    with pytest.raises(SpmdError):
        g.patches()


def test_statistics_row_matches_oracle():
    g, o = _state(shear=0.02, star=True, collisions=5)
    o.md_init()
    g.propagate(True)
    assert o.propagate(True) == 0
    row = g.statistics()
    n = o.n_mono
    rg, G = o.gyration()
    ek = o.ekin_monomers()
    want = [o.epot() / n, ek, ek * 2 / (3 * (n - 1)), *o.patches(), rg, *G.ravel(order="F"),
            *o.rotation_frequency()]
    np.testing.assert_allclose(row, want, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(row[6:15].reshape(3, 3), g.gyration_tensor(), rtol=0, atol=0)


# ---------------------------------------------------------------- fluid VACF
def _vacf_oracle():
    import os
    import sys

    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
    import vacf

    return vacf


def _by_id(g, r=0):
    _, v, ids = g.get_fluid(r, ids=True)
    out = np.empty_like(v)
    out[ids] = v
    return out


def test_fluid_vacf_matches_restatement():
    """Device windows (by particle id) and lag sums against the reference's
    functor restated on the CPU, fed the same velocities in id order; 9
    samples of a 4-entry window, a collision between samples.  Bar: the
    calls counted equal, every lag sum within 1e-6 (the device rounds each
    product to 2^-32)."""
    vacf = _vacf_oracle()
    g = MPCSystem(make_params(L=(6, 6, 6), n_c=5, shear=0.1, step_update=1, seed=12))
    g.set_tracking(True)
    g.initialize()
    o = vacf.VelAutocorrFluid(g.num_fluid(), 4, 1.0)
    for t in range(9):
        g.step(t, 0.1)
        g.fluid_vacf_sample(4)
        o.sample(_by_id(g).astype(np.float64))
    sums, counter = g.fluid_vacf_read(4)
    assert counter == o.counter == 6
    assert np.allclose(sums, o.lag_sums(), rtol=0, atol=1e-6), (sums, o.lag_sums())
    with pytest.raises(SpmdError):
        g.fluid_vacf_sample(5)  # entries are fixed until a reset
    g.fluid_vacf_reset()
    g.fluid_vacf_sample(5)
    assert g.fluid_vacf_read(5)[1] == 0


def test_fluid_vacf_needs_tracking():
    g = MPCSystem(make_params(L=(6, 6, 6), n_c=5, seed=1))
    g.initialize()
    with pytest.raises(SpmdError):
        g.fluid_vacf_sample(4)


def test_run_mpc_test_writes_reference_files(tmp_path):
    """mpc_test.cpp's run on a 5^3 box (N=10, T=0.5, no shear, h=0.1):
    the five files under the reference's names and formats; the VACF at lag
    0 is <v^2> = 3 kT for the thermalised solvent."""
    g = MPCSystem(make_params(L=(5, 5, 5), n_c=10, kT=0.5, shear=0.0, step_update=1, seed=4))
    g.set_tracking(True)
    g.initialize()
    names = obs.run_mpc_test(g, 100, 1100, 100, 0.1, str(tmp_path))
    assert sorted(names) == ["fluid", "fourier", "maxwell", "temperature", "velocity"]
    assert names["velocity"].endswith("velocity_N1250_Shear0_h0.1.dat")
    temp = [l.split() for l in open(names["temperature"]).read().splitlines()]
    assert [int(a) for a, _ in temp] == list(range(200, 1100, 100))
    T = np.mean([float(b) for _, b in temp])
    assert abs(T - 0.5) < 0.05
    vel = open(names["velocity"]).read().splitlines()
    assert vel[0].startswith("#Diffusion coefficient = ") and len(vel) == 51
    lag = np.array([[float(x) for x in l.split()] for l in vel[1:]])
    assert np.allclose(lag[:, 0], np.arange(50) * 1.0)
    assert abs(lag[0, 1] - 3 * 0.5) < 0.1 * 1.5
    assert abs(lag[1:, 1]).max() < 0.05  # 10 collisions apart: decorrelated
    assert len(open(names["fourier"]).read().splitlines()) == 50
    assert open(names["maxwell"]).read().strip() and open(names["fluid"]).read().strip()


@pytest.mark.parametrize("star,shear", [(True, 0.05), (False, 0.0)])
def test_cell_fluid_velocity_matches_oracle(star, shear):
    """MPC::calculateFluidVelocity (MPC.cpp:334-342), as MPC::print_fluid
    reads it (MPC.cpp:533-553): after one collision from the same state
    with the same injected randoms, the mean solvent velocity of every
    cell of that collision's sort equals the oracle's mean over the
    particles its cell index puts there (monomers excluded) within 1e-6;
    every cell is empty before the first collision."""
    L = (12, 12, 12)
    g = MPCSystem(make_params(L=L, n_c=10, shear=shear, step_update=1, seed=5))
    o = ob.Oracle(L=L, n_c=10, shear=shear, step_update=1, rng_mode=ob.RNG_INJECTED, seed=5)
    if star:
        g.add_star(3, 3, 6, 10.0)
        o.add_star(3, 3, 6, 10.0)
        pos, vel, _ = g.get_monomers()
        o.set_monomers(pos, vel)
    from parity_util import random_injection, seeded_fluid
    p, v = seeded_fluid(L, 10, shear, seed=2)
    g.set_fluid(p, v)
    o.set_fluid(p.astype(np.float64), v.astype(np.float64))
    assert np.all(g.cell_fluid_velocity(0, L[2] - 1) == 0.0)
    shift, axis, gamma = random_injection(int(np.prod(L)), seed=9)
    g.inject(shift, axis, gamma)
    o.inject(shift, axis, gamma)
    g.step(0, 0.1)
    assert o.step(0, 0.1) == 0
    g.sync()
    fc, _ = o.cell_index()
    _, vo = o.get_fluid()
    nc = int(np.prod(L))
    cnt = np.bincount(fc, minlength=nc + 1)[:nc]
    want = np.zeros((nc, 3))
    for k in range(3):
        want[:, k] = np.bincount(fc, weights=vo[:, k], minlength=nc + 1)[:nc]
    want[cnt > 0] /= cnt[cnt > 0, None]
    got = g.cell_fluid_velocity(0, L[2] - 1).reshape(nc, 3)
    assert (cnt == 0).sum() < nc // 2
    assert np.abs(got - want).max() <= 1e-6, np.abs(got - want).max()
    # a z range is the matching slab of cells
    sub = g.cell_fluid_velocity(4, 6).reshape(-1, 3)
    assert np.array_equal(sub, got[4 * 144:7 * 144])
    with pytest.raises(SpmdError):
        g.cell_fluid_velocity(5, L[2])
    # the text of MPC::print_fluid for the middle plane
    txt = obs.print_fluid_text(g, 7, 6, 6)
    lines = txt.split("\n")
    assert lines[0] == "TIME 7 " and lines[1] == "Z 6 "
    assert len(lines[4].split()) == 2 * L[1]
    assert float(lines[4].split()[0]) == pytest.approx(want[0 + 12 * 0 + 144 * 6, 0], abs=1e-6)
