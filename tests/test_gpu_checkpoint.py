"""Restart and configuration files (MPC::print_fluid_complete /
MPC::initialize(file, Box_old), Box::print_PDB_with_velocity /
Box::add_star(file, ...)) and the binary snapshot.

Bars: the text files follow the reference's printf formats (checked field
by field, the CONECT block against a restatement of Box.cpp:575-597) and
read back within their printed precision (%f: 5e-7, %7.3f: 5e-4); a fluid
restart into a larger box keeps every particle of the file and puts the
rest in the region the box gained (MPC.cpp:73-101); a handle loaded from a
snapshot continues the run bit for bit."""
import re

import numpy as np
import pytest

from starpolymersmd_amd import MPCSystem, make_params

pytestmark = pytest.mark.gpu


def _system(L=8, shear=0.05, seed=3, star=None, tracking=False):
    g = MPCSystem(make_params(L=(L,) * 3, n_c=10, shear=shear, step_update=10, seed=seed))
    if tracking:
        g.set_tracking(True)
    if star:
        g.add_star(*star, 10.0, coil=True)
        g.md_init()
    g.initialize()
    return g


def test_fluid_file_format_and_round_trip(tmp_path):
    g = _system()
    g.run(0, 31, 0.1)
    g.sync()
    path = tmp_path / "end_fluid.dat"
    g.write_fluid(path)
    lines = path.read_text().split("\n")
    assert lines[-1] == "" and len(lines) - 1 == g.num_fluid()
    num = r"-?\d+\.\d{6}"
    assert all(re.fullmatch((num + " ") * 6, l) for l in lines[:-1][:2000])
    pos, vel = g.get_fluid()
    a = np.loadtxt(path)
    assert np.abs(a[:, :3] - pos).max() <= 5.1e-7 and np.abs(a[:, 3:] - vel).max() <= 5.1e-7
    # read back into a box of the same size: every particle from the file
    h = _system(seed=9)
    h.set_delrx(g.get_delrx())
    assert h.read_fluid(path) == g.num_fluid()
    p2, v2 = h.get_fluid()
    assert np.abs(p2 - a[:, :3].astype(np.float32)).max() <= 1e-6
    vmean = a[:, 3:].mean(axis=0)
    assert np.abs(v2 - (a[:, 3:] - vmean)).max() <= 1e-6
    assert np.abs(v2.astype(np.float64).mean(axis=0)).max() < 1e-6


def test_fluid_restart_into_larger_box(tmp_path):
    g = _system(L=8, shear=0.0)
    g.run(0, 11, 0.1)
    path = tmp_path / "fluid.dat"
    g.write_fluid(path)
    h = _system(L=10, shear=0.0, seed=5)
    n_old = g.num_fluid()
    assert h.read_fluid(path, old_box=(8, 8, 8)) == n_old
    p, v = h.get_fluid()
    a = np.loadtxt(path)
    assert np.abs(p[:n_old] - a[:, :3]).max() <= 1e-6
    added = p[n_old:]
    # every coordinate of an added particle lies in the slabs 4 <= |x| <= 5
    assert added.shape[0] == h.num_fluid() - n_old
    assert np.all(np.abs(added) >= 4.0 - 1e-6) and np.all(np.abs(added) <= 5.0 + 1e-6)
    assert np.abs(v.astype(np.float64).mean(axis=0)).max() < 1e-6


def _conect_reference(arms, per):
    """Box.cpp:575-597 restated"""
    out = []
    count, arm = 1, 0
    while arm < arms:
        s = "CONECT %4i " % 1
        while count % 5:
            s += "%4i " % (arm * per + 2)
            arm += 1
            count += 1
            if arm >= arms:
                break
        count = 1
        out.append(s)
    for i in range(arms):
        out.append("CONECT %4u %4u " % (i * per + 2, 1))
        for j in range(1, per):
            n = i * per + j
            out.append("CONECT %4u %4u %4u " % (n + 1, n, n + 2))
        out.append("CONECT %4u %4u " % ((i + 1) * per + 1, (i + 1) * per))
    return out


def test_pdb_format_and_star_from_file(tmp_path):
    arms, A, B = 7, 3, 4
    g = _system(L=32, star=(A, B, arms))
    g.run(0, 21, 0.1)
    pos, vel, _ = g.get_monomers()
    path = tmp_path / "end_config.pdb"
    g.write_pdb(path, 20)
    g.write_pdb(path, 21, append=True)
    lines = path.read_text().split("\n")
    n = len(pos)
    assert lines[0] == "MODEL     20 "
    atoms = lines[1:1 + n]
    for i, l in enumerate(atoms):
        bead_b = i > 0 and (i - 1) % (A + B) + 1 > A
        assert l.startswith("ATOM %6d  %s   GLY     1     " % (i, "C" if bead_b else "O")), l
        f = [float(l[c:c + 7]) for c in (31, 39, 47, 55, 63, 71)]
        assert np.allclose(f[:3], pos[i], atol=5.1e-4) and np.allclose(f[3:], vel[i], atol=5.1e-4)
        assert l == "ATOM %6d  %s   GLY    %2d     %7.3f %7.3f %7.3f %7.3f %7.3f %7.3f " % (
            i, "C" if bead_b else "O", 1, *pos[i], *vel[i])
    assert lines[1 + n] == "TER "
    conect = _conect_reference(arms, A + B)
    assert lines[2 + n:2 + n + len(conect)] == conect
    assert lines[2 + n + len(conect)] == "ENDMDL "
    assert lines[3 + n + len(conect)] == "MODEL     21 "
    # Box::add_star(file, ...): coordinates from the first MODEL
    h = MPCSystem(make_params(L=(32,) * 3, n_c=10, seed=4))
    h.add_star_pdb(path, A, B, arms, 10.0, set_zero=False)
    p2, v2, _ = h.get_monomers()
    want = np.array([[float(l[c:c + 7]) for c in (31, 39, 47, 55, 63, 71)] for l in atoms])
    assert np.array_equal(p2, want[:, :3]) and np.array_equal(v2, want[:, 3:])
    assert h.star_layout() == (n, arms, A, B)
    h2 = MPCSystem(make_params(L=(32,) * 3, n_c=10, seed=4))
    h2.add_star_pdb(path, A, B, arms, 10.0, set_zero=True)
    p3, v3, _ = h2.get_monomers()
    assert np.abs(p3.mean(axis=0)).max() < 1e-12 and np.abs(v3.mean(axis=0)).max() < 1e-12


def _by_id(g):
    p, v, ids = g.get_fluid(ids=True)
    o = np.argsort(ids)
    return p[o], v[o]


def test_snapshot_continues_bit_for_bit(tmp_path):
    g = _system(L=16, shear=0.05, star=(5, 5, 10), tracking=True)
    g.run(0, 251, 0.1)
    path = tmp_path / "run.snap"
    g.save_snapshot(path, step=251)
    g.run(251, 150, 0.1)
    g.sync()
    h, step = MPCSystem.load_snapshot(path)
    assert step == 251
    h.run(251, 150, 0.1)
    h.sync()
    pg, vg = _by_id(g)
    ph, vh = _by_id(h)
    assert np.array_equal(pg, ph) and np.array_equal(vg, vh)
    mg = g.get_monomers()
    mh = h.get_monomers()
    assert all(np.array_equal(a, b) for a, b in zip(mg, mh))
    assert g.get_delrx() == h.get_delrx()
    assert g.temperature() == h.temperature()
