"""File-name encoded run parameters of the C++ facade (spmd::run_tag,
find_parameter, parse_restart_name) against the reference's own naming.

The expected names follow the pattern the reference's job script globs for
(params.py:103-108: "end_config_Star_A%i_B%i_Arms%i_Mass%i_Lx%i_Ly%i_Lz%i_
Lambda%.2f_T%.2f_run*_" + "MPCON_Shear%.5f.pdb" or "MPCOFF.pdb"), which is
how star-polymers.cpp:270-288 writes them; the chain form drops _Arms and
starts with _Chain (star-polymers.cpp:271-275)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "run_tag_check.cpp")

CASES = [
    # chain A B Arms Mass Lx Ly Lz Lambda T Steps mpc Shear
    (0, 20, 20, 30, 10, 64, 64, 64, 1.0, 1.0, 10000, 1, 0.05),
    (0, 5, 5, 10, 10, 20, 20, 20, 0.5, 0.5, 0, 1, 0.01),
    (0, 15, 3, 6, 1, 32, 24, 40, 0.25, 1.25, 123456789, 0, 0.0),
    (1, 7, 9, 0, 5, 16, 16, 16, 1.0, 1.0, 42, 1, 0.001),
]


def expected(c):
    chain, A, B, arms, mass, lx, ly, lz, lam, T, steps, mpc, shear = c
    s = f"_Chain_A{A}_B{B}" if chain else "_Star_A%i_B%i_Arms%i" % (A, B, arms)
    s += "_Mass%i_Lx%i_Ly%i_Lz%i_Lambda%.2f_T%.2f_run%i_" % (mass, lx, ly, lz, lam, T, steps)
    s += ("MPCON_Shear%.5f" % shear) if mpc else "MPCOFF"
    return s


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("tag") / "run_tag_check"
    subprocess.run([gxx, "-std=c++17", "-O1", SRC, "-o", str(exe)], check=True, capture_output=True, text=True)
    return str(exe)


def test_run_tag_matches_reference_pattern_and_parses_back(prog):
    stdin = "\n".join(" ".join(str(v) for v in c) for c in CASES) + "\n"
    out = subprocess.run([prog], input=stdin, capture_output=True, text=True, check=True).stdout.splitlines()
    assert len(out) == len(CASES) + 2
    for c, line in zip(CASES, out):
        f = line.split()
        assert f[0] == expected(c)
        chain, A, B, arms, mass, lx, ly, lz, lam, T, steps, mpc, shear = c
        assert [int(v) for v in f[1:9]] == [chain, A, B, arms, mass, lx, ly, lz]
        assert float(f[9]) == pytest.approx(lam) and float(f[10]) == pytest.approx(T)
        assert int(f[11]) == steps and int(f[12]) == mpc
        assert float(f[13]) == pytest.approx(shear if mpc else 0.0)
    assert out[-2] == "0.0"          # find_parameter of an absent name
    assert out[-1] == "error 1"      # SPMD_ERR_ARG: no end_config in the name
