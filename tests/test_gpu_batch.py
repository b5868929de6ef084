"""Batched replicas (one handle, many independent boxes advanced by the
same launches) and launch-shape tuning.

Bars: every replica of a batch matches the FP64 oracle run with that
replica's Philox streams (cell indices bit-exact, state within TOL); the
launch shapes (particles per thread, cells per collide tile) do not change
any result bit (all sums are fixed point); the autotuner leaves the system
on the tuning it reports."""
import numpy as np
import pytest

import oracle_binding as ob
from parity_util import TOL, le_diff, rel_err, seeded_fluid
from starpolymersmd_amd import MPCSystem, make_params
from starpolymersmd_amd.autotune import autotune

pytestmark = pytest.mark.gpu

REPS = [dict(L=(8, 8, 8), shear=0.05, n_c=10), dict(L=(12, 10, 8), shear=0.0, n_c=5),
        dict(L=(6, 6, 6), shear=0.1, n_c=20)]


def _by_id(g, r):
    p, v, ids = g.get_fluid(r, ids=True)
    inv = np.empty_like(ids)
    inv[ids] = np.arange(len(ids))
    return p[inv], v[inv]


def test_batch_replicas_match_oracle():
    g = MPCSystem([make_params(L=c["L"], n_c=c["n_c"], shear=c["shear"], seed=2024) for c in REPS])
    assert g.n_replicas == len(REPS)
    g.set_tracking(True)
    fl = []
    for r, c in enumerate(REPS):
        pos, vel = seeded_fluid(c["L"], c["n_c"], c["shear"], seed=40 + r)
        g.set_fluid(pos, vel, r)
        fl.append((pos, vel))
    ors = []
    for r, c in enumerate(REPS):
        o = ob.Oracle(L=c["L"], n_c=c["n_c"], shear=c["shear"], rng_mode=ob.RNG_PHILOX, seed=2024, replica=r)
        ors.append(o)
    for t in range(3):
        for r, o in enumerate(ors):
            p, v = _by_id(g, r)
            o.set_fluid(p.astype(np.float64), v.astype(np.float64))
            o.delrx = g.get_delrx(r)
        g.step(t, 0.1)
        g.sync()
        for r, (o, c) in enumerate(zip(ors, REPS)):
            assert o.step(t, 0.1) == 0
            fc_g, _ = g.cell_index(r)
            fc_o, _ = o.cell_index()
            assert np.array_equal(fc_g, fc_o), (t, r)
            pg, vg, ids = g.get_fluid(r, ids=True)
            po, vo = o.get_fluid()
            d, dv, _ = le_diff(pg, po[ids], vg, vo[ids], c["L"], o.delrx, c["shear"])
            assert rel_err(d, po[ids]).max() <= TOL and rel_err(dv, vo[ids]).max() <= TOL, (t, r)


def test_replica_zero_of_batch_equals_single_system():
    c = REPS[0]
    single = MPCSystem(make_params(L=c["L"], n_c=c["n_c"], shear=c["shear"], seed=5))
    batch = MPCSystem([make_params(L=x["L"], n_c=x["n_c"], shear=x["shear"], seed=5) for x in REPS])
    for g in (single, batch):
        g.set_tracking(True)
        g.initialize()
        for t in range(4):
            g.step(t, 0.1)
    a, b = _by_id(single, 0), _by_id(batch, 0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("ppt,tc", [(1, 64), (2, 256), (8, 128)])
def test_launch_shapes_do_not_change_results(ppt, tc):
    def run(tuning):
        g = MPCSystem(make_params(L=(16, 16, 16), n_c=10, shear=0.03, step_update=5, seed=9, md_dt=0.01))
        g.add_star(3, 3, 6, 10.0)
        g.md_init()
        g.set_tracking(True)
        g.initialize()
        if tuning:
            g.set_tuning(*tuning)
        g.run(0, 21, 0.05)
        g.sync()
        return _by_id(g, 0), g.get_monomers()

    (pa, va), ma = run(None)
    (pb, vb), mb = run((ppt, tc))
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    for x, y in zip(ma, mb):
        assert np.array_equal(x, y)


def test_autotune_picks_and_applies(tmp_path):
    g = MPCSystem(make_params(L=(16, 16, 16), n_c=10, shear=0.02, step_update=10, seed=3))
    g.initialize()
    g.run(0, 1, 0.1)
    res = autotune(g, 1, 0.1, intervals=2, warm=1, cache=str(tmp_path / "t.json"),
                   candidates=[(2, 64), (4, 128)])
    assert (res["ppt"], res["tile_cells"]) in {(2, 64), (4, 128)}
    assert g.get_tuning() == (res["ppt"], res["tile_cells"])
    assert res["t"] == 1 + 2 * (1 + 2) * 10
    again = autotune(g, res["t"], 0.1, cache=str(tmp_path / "t.json"))
    assert again["cached"] and (again["ppt"], again["tile_cells"]) == (res["ppt"], res["tile_cells"])
    T = g.temperature()
    assert 0.5 < T < 1.5


def test_uneven_64_replica_batch_runs_in_one_graph():
    """The batched workload of bench.py (64 replicas, L in {20,24,32,40},
    f 5..30, N_A=N_B 3..15, shear 1e-3..1e-1): 60 graph-replayed intervals
    run clean, every replica's solvent keeps kT and its star its monomers'
    count; replica 0 equals the same system run alone."""
    import bench

    reps = bench.batch_workload()
    ps = [make_params(L=(c["L"],) * 3, n_c=10, shear=c["shear"], step_update=10, md_dt=0.01, seed=7, lam=1.0)
          for c in reps]
    g = MPCSystem(ps)
    for r, c in enumerate(reps):
        g.add_star(c["A"], c["B"], c["arms"], 10.0, r=r, coil=True)
    g.md_init()
    g.initialize()
    g.run(0, 601, 0.1)
    g.sync()
    for r, c in enumerate(reps):
        assert abs(g.temperature(r) - 1.0) < 0.05, r
        assert g.num_monomers(r) == c["arms"] * 2 * c["A"] + 1
    # replica 0 of the batch and a lone system share their Philox streams
    g0 = MPCSystem(ps[0])
    g0.add_star(reps[0]["A"], reps[0]["B"], reps[0]["arms"], 10.0, coil=True)
    g0.md_init()
    g0.initialize()
    g0.run(0, 601, 0.1)
    g0.sync()
    a = g.get_monomers(0)
    b = g0.get_monomers()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
