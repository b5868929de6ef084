"""The row-tile collision path (k_advect + k_rows_redo + k_collide_rows)
against the sort path (k_bin + k_scan + k_perm + k_collide) and the oracle.

Both paths evaluate every particle and cell with the same expressions and
sum the cells in 64-bit fixed point, so from the same state they must give
the same bits: particle positions and velocities (matched by id), monomer
state and cell indices, over many collisions -- including the recentring redo
after an MD check_bounds, the angular-momentum collision, batches of
uneven replicas and densities far from 10.  The oracle check runs one
row-tile collision of the L=64 star workload with injected randoms and
with the library's Philox streams.
"""
import numpy as np
import pytest

import oracle_binding as ob
from parity_util import TOL, le_diff, random_injection, rel_err, seeded_fluid
from starpolymersmd_amd import MPCSystem, make_params

pytestmark = pytest.mark.gpu


def _by_id(g, r=0):
    p, v, ids = g.get_fluid(ids=True, r=r)
    assert np.array_equal(np.sort(ids), np.arange(len(ids))), "ids are not a permutation"
    inv = np.empty_like(ids)
    inv[ids] = np.arange(len(ids))
    return p[inv], v[inv]


def _same_state(a, b, n_rep=1, star=False):
    for r in range(n_rep):
        pa, va = _by_id(a, r)
        pb, vb = _by_id(b, r)
        assert np.array_equal(pa, pb) and np.array_equal(va, vb), f"replica {r}: fluid differs"
        if star:
            for x, y in zip(a.get_monomers(r=r), b.get_monomers(r=r)):
                assert np.array_equal(x, y), f"replica {r}: monomers differ"
    fa, ma = a.cell_index()
    fb, mb = b.cell_index()
    assert np.array_equal(fa, fb)
    if star:
        assert np.array_equal(ma, mb)


def _twins(params, star=None):
    """Two handles from the same state: a on the automatic path, b on the
    sort path."""
    out = []
    for path in (MPCSystem.PATH_AUTO, MPCSystem.PATH_SORT):
        g = MPCSystem(params)
        g.set_collision_path(path)
        if star:
            g.add_star(*star)
            g.md_init()
        g.set_tracking(True)
        g.initialize()
        out.append(g)
    return out


@pytest.mark.parametrize("L,shear,su,n", [((20, 20, 20), 0.05, 1, 30), ((6, 10, 8), 0.1, 1, 40),
                                          ((32, 24, 20), 0.0, 1, 20), ((4, 4, 4), 0.2, 1, 40)])
def test_rows_equal_sort_path_solvent(L, shear, su, n):
    a, b = _twins(make_params(L=L, n_c=10, shear=shear, step_update=su, seed=21))
    assert a.collision_path() == MPCSystem.PATH_SORT  # fluid just set: not in cell order
    a.step(0, 0.1)
    b.step(0, 0.1)
    assert a.collision_path() == MPCSystem.PATH_ROWS and b.collision_path() == MPCSystem.PATH_SORT
    for t in range(1, n):
        a.step(t, 0.1)
        b.step(t, 0.1)
    a.sync()
    b.sync()
    _same_state(a, b)
    # per-cell FP64 sums in slot order (the order inside a cell may differ)
    assert a.temperature() == pytest.approx(b.temperature(), rel=1e-12)


def test_rows_equal_sort_path_star_graph():
    """L=20 star workload (f=10, 5A+5B, shear 0.01) replayed from graphs."""
    p = make_params(L=(20, 20, 20), n_c=10, shear=0.01, step_update=10, seed=5, md_dt=0.01, lam=1.0)
    a, b = _twins(p, star=(5, 5, 10, 10.0))
    a.run(0, 301, 0.1)
    b.run(0, 301, 0.1)
    a.sync()
    b.sync()
    _same_state(a, b, star=True)


def test_rows_recentring_equal_sort_path():
    """A check_bounds recentring inside a row-tile interval: the advect
    pass ran beside the MD assuming none, the redo kernel advects the
    replica again with the recentring into its own counters."""
    p = make_params(L=(20, 20, 20), n_c=10, shear=0.02, step_update=10, seed=3, md_dt=0.01, lam=1.0)
    a, b = _twins(p, star=(5, 5, 10, 10.0))
    a.run(0, 11, 0.1)
    b.run(0, 11, 0.1)
    recentred = 0
    for k in range(6):
        for g in (a, b):
            mpos, _, _ = g.get_monomers()
            # push the star so a monomer sits beyond L/2 - 0.5 in x, and in
            # y and z on alternate intervals
            off = np.array([9.6 - mpos[:, 0].max(), 0.0, 0.0])
            if k % 2:
                off[1] = -9.7 - mpos[:, 1].min()
                off[2] = 0.37
            g.set_monomers(pos=mpos + off)
        assert a.collision_path() == MPCSystem.PATH_ROWS
        a.run(11 + 10 * k, 10, 0.1)
        b.run(11 + 10 * k, 10, 0.1)
        a.sync()
        recentred += a.last_collision()[1]
    assert recentred >= 5
    b.sync()
    _same_state(a, b, star=True)


def test_rows_amc_equal_sort_path():
    p = make_params(L=(12, 12, 12), n_c=10, shear=0.05, step_update=10, seed=9, md_dt=0.01, lam=1.0,
                    angular_momentum=1)
    a, b = _twins(p, star=(2, 2, 6, 10.0))
    a.run(0, 121, 0.1)
    b.run(0, 121, 0.1)
    a.sync()
    b.sync()
    _same_state(a, b, star=True)


def test_rows_batch_uneven_replicas_equal_sort_path():
    """Replicas of different boxes, densities, shear rates and stars in one
    handle (the densities give tiles of one row up to twelve)."""
    boxes = [(20, 20, 20), (24, 16, 16), (12, 12, 12), (40, 20, 20)]
    dens = [10, 5, 30, 3]
    shears = [0.001, 0.01, 0.1, 0.05]
    ps = [make_params(L=L, n_c=nc, shear=sh, step_update=10, seed=30 + i, md_dt=0.01, lam=1.0)
          for i, (L, nc, sh) in enumerate(zip(boxes, dens, shears))]
    out = []
    for path in (MPCSystem.PATH_AUTO, MPCSystem.PATH_SORT):
        g = MPCSystem(ps)
        g.set_collision_path(path)
        for r, st in enumerate([(4, 4, 5, 10.0), (2, 2, 4, 10.0), (1, 2, 3, 10.0), (2, 3, 6, 10.0)]):
            g.add_star(*st, r=r, coil=True)
        g.md_init()
        g.set_tracking(True)
        g.initialize()
        out.append(g)
    a, b = out
    a.run(0, 201, 0.1)
    b.run(0, 201, 0.1)
    a.sync()
    b.sync()
    for r in range(4):
        pa, va = _by_id(a, r)
        pb, vb = _by_id(b, r)
        assert np.array_equal(pa, pb) and np.array_equal(va, vb), r
        for x, y in zip(a.get_monomers(r=r), b.get_monomers(r=r)):
            assert np.array_equal(x, y), r


def test_rows_dense_cells_equal_sort_path():
    """100 particles per cell: one row per tile, ranks far above a byte."""
    p = make_params(L=(12, 8, 6), n_c=100, shear=0.05, step_update=1, seed=12)
    a, b = _twins(p)
    for t in range(12):
        a.step(t, 0.1)
        b.step(t, 0.1)
    a.sync()
    b.sync()
    _same_state(a, b)


L64 = (64, 64, 64)


@pytest.mark.parametrize("rng_mode", [ob.RNG_INJECTED, ob.RNG_PHILOX])
def test_rows_interval_matches_oracle_l64_star(rng_mode):
    """The L=64 star workload (2,621,440 solvent, 601-monomer coil, shear
    0.05) on both sides from the same state: the first interval (sort path,
    it orders the buffer), the oracle reloaded with the device state, then
    one interval on the row-tile path: cell indices of all particles and
    monomers bit-exact, solvent and monomer state within TOL."""
    seed = 13
    g = MPCSystem(make_params(L=L64, n_c=10, shear=0.05, step_update=10, seed=seed, md_dt=0.01, lam=1.0))
    g.add_star(15, 15, 20, 10.0, coil=True)
    o = ob.Oracle(L=L64, n_c=10, shear=0.05, step_update=10, rng_mode=rng_mode, seed=seed, md_dt=0.01, lam=1.0)
    o.add_star(15, 15, 20, 10.0)
    mpos, mvel, _ = g.get_monomers()
    o.set_monomers(mpos, mvel)
    g.md_init()
    assert o.md_init() == 0
    pos, vel = seeded_fluid(L64, 10, 0.05, seed=4)
    g.set_tracking(True)
    g.set_fluid(pos, vel)
    o.set_fluid(pos.astype(np.float64), vel.astype(np.float64))
    g.set_delrx(5.25)
    o.delrx = 5.25
    for t in range(20):
        if t % 10 == 0 and rng_mode == ob.RNG_INJECTED:
            shift, axis, gamma = random_injection(o.ncells, seed=19 + t)
            g.inject(shift, axis, gamma)
            o.inject(shift, axis, gamma)
        if t == 10:
            # the row-tile interval starts from the device state on both sides
            g.sync()
            assert g.collision_path() == MPCSystem.PATH_ROWS
            pg, vg = _by_id(g)
            o.set_fluid(pg.astype(np.float64), vg.astype(np.float64))
            mpos, mvel, _ = g.get_monomers()
            o.set_monomers(mpos, mvel)
            assert o.md_init() == 0
        g.propagate()
        assert o.propagate() == 0
        g.step(t, 0.1)
        assert o.step(t, 0.1) == 0
    g.sync()
    assert abs(g.get_delrx() - o.delrx) == 0.0
    fg, mg = g.cell_index()
    fo, mo = o.cell_index()
    assert np.array_equal(mg, mo)
    assert np.array_equal(fg, fo), f"{np.count_nonzero(fg != fo)} cell indices differ"
    pg, vg, ids = g.get_fluid(ids=True)
    po, vo = o.get_fluid()
    d, dv, _ = le_diff(pg, po[ids], vg, vo[ids], L64, o.delrx, 0.05)
    assert rel_err(d, po[ids]).max() <= TOL and rel_err(dv, vo[ids]).max() <= TOL
    mpg, mvg, _ = g.get_monomers()
    mpo, mvo, _ = o.get_monomers()
    assert np.abs(mpg - mpo).max() <= TOL
    assert rel_err(mvg - mvo, mvo).max() <= TOL
