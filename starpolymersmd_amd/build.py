"""Build recipe: libspmd.so (HIP, gfx950) in-tree, plus the CPU oracle.

    python -m starpolymersmd_amd.build          # library, C++ facade driver, oracle
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LIB_SOURCES = ["spmd_kernels.hip", "spmd_api.hip", "spmd_mlp.hip", "spmd_io.hip"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_library():
    out = os.path.join(HERE, "libspmd.so")
    tmp = out + ".tmp"
    _run([hipcc(), *HIP_FLAGS, *[os.path.join(CSRC, s) for s in LIB_SOURCES], "-o", tmp])
    os.replace(tmp, out)
    return out


def build_facade():
    """The reference-style C++ drivers over the facade (star-polymers and
    mpc-test), linked to libspmd.so."""
    fdir = os.path.join(HERE, "facade")
    gxx = shutil.which("g++") or "g++"
    outs = []
    for name in ("star_polymers_run", "mpc_test_run"):
        out = os.path.join(fdir, name)
        _run([gxx, "-O2", "-std=c++17", "-Wall", os.path.join(fdir, name + ".cpp"), "-o", out + ".tmp",
              "-L" + HERE, "-lspmd", "-Wl,-rpath,$ORIGIN/.."])
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_oracle():
    """The FP64 oracle and, when the reference tree is present, its harness."""
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "all"])
    if os.path.isdir("/root/reference/star-polymers/lib"):
        _run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build_all():
    build_library()
    build_facade()
    build_oracle()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
