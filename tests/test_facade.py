"""The C++ facade (reference class names over the C ABI): it compiles on
its own, the reference-style driver is built and linked against
libspmd.so, and on the GPU it produces exactly the trajectory the Python
binding produces for the same parameters (same library, same Philox
streams, same call order)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from starpolymersmd_amd import MPCSystem, make_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "starpolymersmd_amd", "facade")
DRIVER = os.path.join(FDIR, "star_polymers_run")


def test_facade_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "spmd_facade.h"\nint main() { return 0; }\n')
    gxx = shutil.which("g++")
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", FDIR, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


BASE_DRIVER = r"""
#include "spmd_facade.h"
#include <cstdio>
#include <iostream>
// the loop of star-polymers.cpp:351-536, integrator and solvent through
// their base classes (Thermostat.h:7-22, Hydrodynamics.h:7-22)
int run(int argc) {
    using namespace spmd;
    Box box(16, 16, 16, 1.0, 1.0);
    MPC hydrodynamics{box, 1.0, 10, 0.05, 100, false};
    VelocityX velocity_average_x{box, hydrodynamics, 0.2};
    box.add_star(3, 3, 6, 10.0);
    box.print_center_of_mass(std::cout);
    Thermostat* thermostat = new Thermostat_None{box, 0.01};
    Hydrodynamics* hyd = &hydrodynamics;
    hyd->initialize();
    FILE* fluid_file = std::fopen("/dev/null", "w");
    for (long n = 0; n <= argc; ++n) {
        if (n % 10 == 0) {
            thermostat->propagate(true);
            box.center_of_mass_reference_frame();
            hydrodynamics(velocity_average_x);
            hyd->print_fluid(fluid_file, (int)n, 8, 8);
            box.print_center_of_mass(std::cout);
            std::cout << hyd->calculateCurrentTemperature() << thermostat->name() << thermostat->dtime() << "\n";
        } else thermostat->propagate(false);
        hyd->step(n, 100 * 0.01);
    }
    FILE* restart = std::fopen("/dev/null", "w");
    box.print_PDB_with_velocity(restart, argc);
    hydrodynamics.print_fluid_complete(restart);
    std::fclose(restart);
    struct Count { unsigned n = 0; void operator()(const MPCParticle&) { ++n; } } count;
    hydrodynamics(count);
    double dt = 0.005;
    thermostat->dtime(dt);
    std::fclose(fluid_file);
    delete static_cast<Thermostat_None*>(thermostat);
    return (int)count.n;
}
int main(int argc, char**) { return argc > 5 ? run(argc) : 0; }
"""


def test_reference_loop_through_base_classes_compiles(tmp_path):
    """A driver holding Thermostat* and Hydrodynamics* as the reference's
    does (star-polymers.cpp:101, 382, 412-536) compiles against the facade
    with every warning an error."""
    src = tmp_path / "loop.cpp"
    src.write_text(BASE_DRIVER)
    gxx = shutil.which("g++")
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I", FDIR, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


HOST_ONLY = r"""
#include "spmd_facade.h"
#include <cassert>
#include <iostream>
#include <sstream>
int main() {
    using namespace spmd;
    // Eigen's default IOFormat: right-aligned to the widest coefficient
    std::ostringstream os;
    Vector3d a{1.5, -0.25, 10.125};
    os << a.transpose() << "|" << a;
    if (os.str() != "   1.5  -0.25 10.125|   1.5\n -0.25\n10.125") { std::cout << os.str(); return 1; }
    Vector3d b = (a + Vector3d{1, 1, 1}) * 2.0 - a / 2.0;
    if (b(0) != 4.25 || b[1] != 1.625 || std::abs(a.norm() * a.norm() - a.squaredNorm()) > 1e-12) return 2;
    // the mirror reads once per invalidation, writes back only after a non-const access
    int pulls = 0, pushes = 0;
    std::vector<int> dev{1, 2, 3};
    Mirror<int> m;
    m.bind([&](std::vector<int>& v) { ++pulls; v = dev; }, [&](const std::vector<int>& v) { ++pushes; dev = v; });
    const Mirror<int>& cm = m;
    int s = 0;
    for (int x : cm) s += x;
    if (s != 6 || pulls != 1 || cm.size() != 3 || pulls != 1) return 3;
    m.flush();
    if (pushes != 0) return 4;
    m[1] = 20;
    m.flush();
    if (pushes != 1 || dev[1] != 20) return 5;
    m.flush();
    if (pushes != 1) return 6;
    dev[0] = 7;
    m.invalidate();
    if (cm[0] != 7 || pulls != 2) return 7;
    return 0;
}
"""


def test_vector_format_and_mirror_semantics(tmp_path):
    """Host-only parts of the facade: Vector3d prints as Eigen does
    (Box::print_center_of_mass, Molecule::print), and the lazy mirrors of
    Box::Molecules / MPC::Fluid read the device once per change and write
    back only what the host modified."""
    src = tmp_path / "host.cpp"
    exe = tmp_path / "host"
    src.write_text(HOST_ONLY)
    gxx = shutil.which("g++")
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-I", FDIR, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=30)
    assert r.returncode == 0, (r.returncode, r.stdout)


def _close_text(a: str, b: str, tol: float) -> bool:
    """Same words; numbers equal within tol."""
    wa, wb = a.split(), b.split()
    if len(wa) != len(wb) or a.count("\n") != b.count("\n"):
        return False
    for x, y in zip(wa, wb):
        if x == y:
            continue
        try:
            if abs(float(x) - float(y)) > tol:
                return False
        except ValueError:
            return False
    return True


def test_driver_built_and_linked():
    assert os.path.exists(DRIVER), "run __graft_entry__.build()"
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr  # argument check only: no device use


@pytest.mark.gpu
def test_driver_matches_python_binding(tmp_path):
    """Same trajectory, same statistics.dat rows, VelocityX profile and
    restart files as the Python binding run with the same calls."""
    from starpolymersmd_amd import observables as ob

    L, A, B, arms, shear, steps, out, dt, nc = 16, 3, 3, 6, 0.02, 40, 20, 0.01, 10
    r = subprocess.run([DRIVER, str(L), str(A), str(B), str(arms), str(shear), str(steps), str(out), str(dt),
                        str(nc), "1", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l.split() for l in r.stdout.splitlines() if "T=" in l]
    assert [int(x[0]) for x in rows] == [20, 40]
    stat_lines = (tmp_path / "statistics.dat").read_text().split("\n")
    assert stat_lines[0] + "\n" == ob.STAT_HEADER

    g = MPCSystem(make_params(L=(L, L, L), kT=1.0, n_c=nc, shear=shear, step_update=10, md_dt=dt, seed=1, lam=1.0))
    g.add_star(A, B, arms, 10.0, coil=True)
    g.md_init()
    g.initialize()
    prof = ob.VelocityProfile(0.5)
    out_text, fluid_text = [], []
    for n in range(steps + 1):
        out_step = n % out == 0 and n > 0
        g.propagate(calc_epot=out_step)
        if out_step:
            k = n // out
            assert stat_lines[k] + "\n" == ob.statistics_row(n, g.statistics())
            prof.sample(g)
            out_text.append(f"{n} " + ob.center_of_mass_text(g) + "\n")
            fluid_text.append(ob.print_fluid_text(g, n, L // 2, L // 2))
            row = rows[k - 1]
            T = float(row[1].split("=")[1])
            E = float(row[2].split("=")[1])
            com = np.array([float(v) for v in row[3].split("=")[1].strip("()").split(",")])
            pos, _, _ = g.get_monomers()
            assert abs(g.temperature() - T) <= 1e-6
            assert abs(g.epot() - E) <= 1e-6 * max(1.0, abs(E))
            assert np.abs(pos.mean(axis=0) - com).max() <= 1e-4
        g.step(n, 10 * dt)
    g.sync()
    import io
    buf = io.StringIO()
    prof.print_result(buf)
    assert (tmp_path / "fluid_profile.dat").read_text() == buf.getvalue()
    # Box::print_center_of_mass rows (star-polymers.cpp:488-491), exact
    assert (tmp_path / "output.dat").read_text() == "".join(out_text)
    # MPC::print_fluid of the middle plane (:520-523): per-cell means over
    # the cell's particles, summed in the order the sort's atomics left them
    assert _close_text((tmp_path / "fluid.dat").read_text(), "".join(fluid_text), 2e-6)
    assert "TIME 20 \nZ 8 \n0.500000 " in "".join(fluid_text)
    g.write_fluid(tmp_path / "py_fluid.dat")
    # the same particles; their order inside a cell follows the atomics of the sort
    assert sorted((tmp_path / "end_fluid.dat").read_text().split("\n")) == \
        sorted((tmp_path / "py_fluid.dat").read_text().split("\n"))
    g.write_pdb(tmp_path / "py_config.pdb", steps)
    assert (tmp_path / "end_config.pdb").read_text() == (tmp_path / "py_config.pdb").read_text()


MPC_DRIVER = os.path.join(FDIR, "mpc_test_run")


@pytest.mark.gpu
def test_mpc_test_driver_matches_python_driver(tmp_path):
    """mpc_test.cpp over the facade (C++) and observables.run_mpc_test
    (Python) with the same parameters write the same five files byte for
    byte: temperature, VelocityX profile, speed distribution, fluid VACF
    and transverse-mode autocorrelation."""
    from starpolymersmd_amd import observables as ob

    cpp = tmp_path / "cpp"
    py = tmp_path / "py"
    cpp.mkdir()
    py.mkdir()
    args = ["10", "0.5", "6", "6", "6", "0.3", "0.1", "100", "1200", "100", "0"]
    r = subprocess.run([MPC_DRIVER, *args], capture_output=True, text=True, timeout=120, cwd=cpp)
    assert r.returncode == 0, r.stderr
    g = MPCSystem(make_params(L=(6, 6, 6), kT=0.5, n_c=10, shear=0.3, step_update=1, seed=1, lam=1.0, md_dt=0.01))
    g.set_tracking(True)
    g.initialize()
    names = ob.run_mpc_test(g, 100, 1200, 100, 0.1, str(py))
    assert sorted(os.listdir(cpp)) == sorted(os.listdir(py)) == sorted(os.path.basename(p) for p in names.values())
    for f in os.listdir(py):
        a, b = (cpp / f).read_text(), (py / f).read_text()
        assert a == b, (f, a[:300], b[:300])
    assert (py / "velocity_N2160_Shear0.3_h0.1.dat").read_text().startswith("#Diffusion coefficient = ")


@pytest.mark.gpu
def test_driver_signal_dump_and_continue(tmp_path):
    """SIGTERM stops a tagged run at the next step and writes its restart
    files with the step reached in the name (star-polymers.cpp:413-463);
    the driver then continues from that end_config file, reading the run's
    parameters from its name, the star from the PDB and the solvent from
    the matching end_fluid file (:118-133, 199-225, 361-404)."""
    import re
    import signal

    L, A, B, arms, shear, nc = 16, 3, 3, 6, 0.02, 10
    run1 = tmp_path / "run1"
    run2 = tmp_path / "run2"
    run1.mkdir()
    run2.mkdir()
    p = subprocess.Popen([DRIVER, str(L), str(A), str(B), str(arms), str(shear), "1000000", "20", "0.01", str(nc),
                          "1", str(run1), "1"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    try:
        first = p.stdout.readline()
        while first and "T=" not in first:  # the star's centre of mass comes first
            first = p.stdout.readline()
        assert "T=" in first, first  # the loop has reached its first output step
        p.send_signal(signal.SIGTERM)
        out, err = p.communicate(timeout=120)
    finally:
        if p.poll() is None:
            p.kill()
            p.wait()
    assert p.returncode == signal.SIGTERM, err
    assert "signal 15 caught, data saved" in out
    ends = sorted(run1.glob("end_config_*.pdb"))
    assert len(ends) == 1
    m = re.fullmatch(r"end_config_Star_A3_B3_Arms6_Mass10_Lx16_Ly16_Lz16_Lambda1\.00_T1\.00_run(\d+)"
                     r"_MPCON_Shear0\.02000\.pdb", ends[0].name)
    assert m, ends[0].name
    n = int(m.group(1))
    assert 20 <= n < 1000000
    tag = ends[0].name[len("end_config"):-len(".pdb")]
    fluid = (run1 / f"end_fluid{tag}.dat").read_text().splitlines()
    assert len([l for l in fluid if l.strip()]) >= L ** 3 * nc
    stats = (run1 / f"statistics{tag}.dat").read_text().splitlines()
    steps = [int(l.split()[0]) for l in stats[1:] if l.strip()]
    assert steps and all(s % 20 == 0 and s < n + 1 for s in steps)
    assert (run1 / f"fluid_profile{tag}.dat").exists()
    atoms = [l for l in ends[0].read_text().splitlines() if l.startswith(("ATOM", "HETATM"))]
    assert len(atoms) == 1 + arms * (A + B)

    r = subprocess.run([DRIVER, str(ends[0]), str(n + 40), "20", "0.01", str(nc), str(run2)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l.split() for l in r.stdout.splitlines() if "T=" in l]
    assert [int(x[0]) for x in rows] == [s for s in range(n + 1, n + 41) if s % 20 == 0]
    for x in rows:
        T = float(x[1].split("=")[1])
        if int(x[0]) <= (n + 1 + 9) // 10 * 10:
            # an output step before the continued run's first collision:
            # the cell lists are still empty, as in the reference, T = 0
            assert T == 0.0 or abs(T - 1.0) < 0.1, x
        else:  # the continued run stays at kT
            assert abs(T - 1.0) < 0.1, x
    tag2 = tag.replace(f"_run{n}_", f"_run{n + 40}_")
    assert (run2 / f"end_config{tag2}.pdb").exists()
    assert (run2 / f"end_fluid{tag2}.dat").exists()
