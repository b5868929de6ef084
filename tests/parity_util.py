"""Shared helpers of the GPU parity tests: seeded inputs and comparisons
against the oracle, with the tolerances the tests assert."""
from __future__ import annotations

import numpy as np

# positions and velocities after one step: |a-b| <= TOL * max(1, |b|)
# (relative error, floored at unit scale = one MPC cell / one thermal speed)
TOL = 1e-5


def seeded_fluid(L, n_c, shear, seed=0, maxwell=True):
    rng = np.random.default_rng(seed)
    L = np.asarray(L, np.float64)
    n = int(n_c * np.prod(L))
    pos = ((rng.random((n, 3)) - 0.5) * L).astype(np.float32)
    if maxwell:
        vel = rng.normal(size=(n, 3))
    else:
        vel = rng.random((n, 3)) - 0.5
    vel[:, 0] += shear * pos[:, 1]
    return pos, vel.astype(np.float32)


def random_injection(ncells, seed=1):
    rng = np.random.default_rng(seed)
    shift = rng.random(3) - 0.5
    axis = rng.random(2 * ncells)
    gamma = rng.gamma(15.0, 1.0, ncells)
    return shift, axis, gamma


def le_diff(pg, po, vg, vo, L, delrx, shear):
    """Differences between two particle states modulo the Lees-Edwards
    lattice (a particle wrapped to the other face in one of them differs by
    a lattice vector, and by shear*Ly in vx)."""
    d = pg.astype(np.float64) - po
    dv = vg.astype(np.float64) - vo
    Lx, Ly, Lz = L
    ky = np.round(d[:, 1] / Ly)
    d[:, 1] -= ky * Ly
    if shear != 0.0:
        d[:, 0] -= ky * delrx
        dv[:, 0] -= ky * shear * Ly
    d[:, 0] -= Lx * np.round(d[:, 0] / Lx)
    d[:, 2] -= Lz * np.round(d[:, 2] / Lz)
    return d, dv, ky


def rel_err(d, ref):
    return np.abs(d) / np.maximum(1.0, np.abs(ref))
