"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so) and of the
reference harness (oracle/_ref/libref.so) -- test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

RNG_REFERENCE, RNG_PHILOX, RNG_INJECTED = 0, 1, 2


class OParams(C.Structure):
    _fields_ = [
        ("L", C.c_int32 * 3), ("kT", C.c_double), ("n_c", C.c_int32), ("shear", C.c_double),
        ("step_update", C.c_uint32), ("alpha_deg", C.c_double), ("lambda_", C.c_double),
        ("md_dt", C.c_double), ("rng_mode", C.c_int32), ("seed", C.c_uint64), ("replica", C.c_uint32),
        ("amc", C.c_int32),
    ]


_d = C.POINTER(C.c_double)
_u32 = C.POINTER(C.c_uint32)
_u8 = C.POINTER(C.c_uint8)
_SIGS = {
    "orc_create": (C.c_void_p, [C.POINTER(OParams)]),
    "orc_destroy": (None, [C.c_void_p]),
    "orc_fluid_initialize": (None, [C.c_void_p]),
    "orc_num_fluid": (C.c_uint64, [C.c_void_p]),
    "orc_set_fluid": (None, [C.c_void_p, _d, _d, C.c_uint64]),
    "orc_get_fluid": (None, [C.c_void_p, _d, _d]),
    "orc_add_star": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double]),
    "orc_num_monomers": (C.c_uint32, [C.c_void_p]),
    "orc_set_monomers": (None, [C.c_void_p, _d, _d]),
    "orc_get_monomers": (None, [C.c_void_p, _d, _d, _d]),
    "orc_get_monomer_props": (None, [C.c_void_p, _d, _u8, _u8]),
    "orc_get_bonds": (C.c_uint32, [C.c_void_p, _u32, C.c_uint32]),
    "orc_md_init": (C.c_int, [C.c_void_p]),
    "orc_md_propagate": (C.c_int, [C.c_void_p, C.c_int]),
    "orc_epot": (C.c_double, [C.c_void_p]),
    "orc_mpc_step": (C.c_int, [C.c_void_p, C.c_int64, C.c_double]),
    "orc_set_threads": (None, [C.c_int]),
    "orc_get_threads": (C.c_int, []),
    "orc_inject": (None, [C.c_void_p, _d, _d, _d]),
    "orc_get_delrx": (C.c_double, [C.c_void_p]),
    "orc_set_delrx": (None, [C.c_void_p, C.c_double]),
    "orc_last_shift": (None, [C.c_void_p, _d]),
    "orc_last_recentered": (C.c_int, [C.c_void_p]),
    "orc_get_cell_index": (None, [C.c_void_p, _u32, _u32]),
    "orc_temperature": (C.c_double, [C.c_void_p]),
    "orc_com_frame": (None, [C.c_void_p]),
    "orc_gyration_tensor": (C.c_double, [C.c_void_p, _d]),
    "orc_rotation_frequency": (None, [C.c_void_p, _d]),
    "orc_ekin_monomers": (C.c_double, [C.c_void_p]),
    "orc_velocity_profile": (None, [C.c_void_p, C.c_double, C.c_uint32, _u32, _d, _d]),
    "orc_fourier_modes": (None, [C.c_void_p, _d]),
    "orc_speed_histogram": (None, [C.c_void_p, C.c_double, C.c_uint32, C.POINTER(C.c_uint64)]),
    "orc_patches": (None, [C.c_void_p, _d, _d]),
    "orc_fene_force": (C.c_double, [C.c_double]),
    "orc_fene_anchor_force": (C.c_double, [C.c_double]),
    "orc_aa_force": (C.c_double, [C.c_double]),
    "orc_bb_force": (C.c_double, [C.c_double, C.c_double]),
    "orc_fene_potential": (C.c_double, [C.c_double]),
    "orc_fene_anchor_potential": (C.c_double, [C.c_double]),
    "orc_aa_potential": (C.c_double, [C.c_double]),
    "orc_bb_potential": (C.c_double, [C.c_double, C.c_double]),
    "orc_philox4x32_10": (None, [_u32, _u32, _u32]),
    "orc_philox_gamma": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double]),
    "orc_reference_rng_sequence": (None, [C.POINTER(C.c_int), _d, _d, C.c_int, _d]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        l = C.CDLL(ORACLE_SO)
        for n, (res, args) in _SIGS.items():
            f = getattr(l, n)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def p(a, t=C.c_double):
    if a is None:
        return C.cast(None, C.POINTER(t))
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    def __init__(self, L=(8, 8, 8), kT=1.0, n_c=10, shear=0.0, step_update=1, alpha_deg=130.0, lam=1.0,
                 md_dt=0.01, rng_mode=RNG_PHILOX, seed=1, replica=0, amc=False):
        prm = OParams()
        for i in range(3):
            prm.L[i] = int(L[i])
        prm.kT, prm.n_c, prm.shear, prm.step_update = kT, n_c, shear, step_update
        prm.alpha_deg, prm.lambda_, prm.md_dt = alpha_deg, lam, md_dt
        prm.rng_mode, prm.seed, prm.replica = rng_mode, seed, replica
        prm.amc = int(amc)
        self.L = tuple(int(l) + (int(l) % 2) for l in L)
        self.ncells = self.L[0] * self.L[1] * self.L[2]
        self.h = lib().orc_create(C.byref(prm))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def initialize(self):
        lib().orc_fluid_initialize(self.h)

    @property
    def n_fluid(self):
        return lib().orc_num_fluid(self.h)

    def set_fluid(self, pos, vel):
        pos = np.ascontiguousarray(pos, np.float64)
        vel = np.ascontiguousarray(vel, np.float64)
        lib().orc_set_fluid(self.h, p(pos), p(vel), pos.shape[0])

    def get_fluid(self):
        n = self.n_fluid
        pos, vel = np.empty((n, 3)), np.empty((n, 3))
        lib().orc_get_fluid(self.h, p(pos), p(vel))
        return pos, vel

    def add_star(self, A, B, arms, mass, bond=1.01, anchor_bond=2.0):
        rc = lib().orc_add_star(self.h, A, B, arms, mass, bond, anchor_bond)
        assert rc == 0

    @property
    def n_mono(self):
        return lib().orc_num_monomers(self.h)

    def get_monomers(self):
        n = self.n_mono
        a, b, c = np.empty((n, 3)), np.empty((n, 3)), np.empty((n, 3))
        lib().orc_get_monomers(self.h, p(a), p(b), p(c))
        return a, b, c

    def set_monomers(self, pos=None, vel=None):
        pp = None if pos is None else np.ascontiguousarray(pos, np.float64)
        vv = None if vel is None else np.ascontiguousarray(vel, np.float64)
        lib().orc_set_monomers(self.h, p(pp), p(vv))

    def monomer_props(self):
        n = self.n_mono
        m, t, a = np.empty(n), np.empty(n, np.uint8), np.empty(n, np.uint8)
        lib().orc_get_monomer_props(self.h, p(m), p(t, C.c_uint8), p(a, C.c_uint8))
        return m, t, a

    def bonds(self):
        nb = lib().orc_get_bonds(self.h, None, 0)
        arr = np.empty((nb, 2), np.uint32)
        lib().orc_get_bonds(self.h, p(arr, C.c_uint32), nb)
        return arr

    def md_init(self):
        return lib().orc_md_init(self.h)

    def propagate(self, calc_epot=False):
        return lib().orc_md_propagate(self.h, int(calc_epot))

    def epot(self):
        return lib().orc_epot(self.h)

    def step(self, t, dt):
        return lib().orc_mpc_step(self.h, t, dt)

    def inject(self, shift, axis_u, gamma):
        s = np.ascontiguousarray(shift, np.float64)
        a = np.ascontiguousarray(axis_u, np.float64)
        g = np.ascontiguousarray(gamma, np.float64)
        lib().orc_inject(self.h, p(s), p(a), p(g))

    @property
    def delrx(self):
        return lib().orc_get_delrx(self.h)

    @delrx.setter
    def delrx(self, v):
        lib().orc_set_delrx(self.h, v)

    def last_shift(self):
        s = np.empty(3)
        lib().orc_last_shift(self.h, p(s))
        return s

    def last_recentered(self):
        return bool(lib().orc_last_recentered(self.h))

    def cell_index(self):
        f = np.empty(self.n_fluid, np.uint32)
        m = np.empty(max(self.n_mono, 1), np.uint32)
        lib().orc_get_cell_index(self.h, p(f, C.c_uint32), p(m, C.c_uint32))
        return f, m[: self.n_mono]

    def temperature(self):
        return lib().orc_temperature(self.h)

    def gyration(self):
        lib().orc_com_frame(self.h)
        g = np.empty(9)
        rg = lib().orc_gyration_tensor(self.h, p(g))
        return rg, g.reshape(3, 3, order="F")

    def rotation_frequency(self):
        lib().orc_com_frame(self.h)
        o = np.empty(3)
        lib().orc_rotation_frequency(self.h, p(o))
        return o

    def ekin_monomers(self):
        return lib().orc_ekin_monomers(self.h)

    def velocity_profile(self, width, cap=4096):
        n = np.zeros(2, np.uint32)
        keys = np.zeros(2 * cap)
        sums = np.zeros(2 * cap * 3)
        lib().orc_velocity_profile(self.h, width, cap, p(n, C.c_uint32), p(keys), p(sums))
        out = []
        for a in range(2):
            k = int(n[a])
            out.append((keys[a * cap:a * cap + k].copy(), sums.reshape(2, cap, 3)[a, :k].copy()))
        return out

    def fourier_modes(self):
        m = np.empty(18)
        lib().orc_fourier_modes(self.h, p(m))
        return m[0::2].reshape(3, 3) + 1j * m[1::2].reshape(3, 3)

    def speed_histogram(self, width, nbins):
        c = np.empty(nbins + 1, np.uint64)
        lib().orc_speed_histogram(self.h, width, nbins, p(c, C.c_uint64))
        return c

    def patches(self):
        n, sz = C.c_double(), C.c_double()
        lib().orc_patches(self.h, C.byref(n), C.byref(sz))
        return n.value, sz.value


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.empty(4, np.uint32)
    lib().orc_philox4x32_10(p(c, C.c_uint32), p(k, C.c_uint32), p(o, C.c_uint32))
    return o


def philox_gamma(seed, replica, cell, k, shape, scale):
    return lib().orc_philox_gamma(seed, replica, cell, k, shape, scale)


def reference_rng_sequence(ops, a, b):
    ops = np.ascontiguousarray(ops, np.int32)
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    out = np.empty(len(ops))
    lib().orc_reference_rng_sequence(p(ops, C.c_int), p(a), p(b), len(ops), p(out))
    return out


def potentials():
    l = lib()
    return {
        "fene_force": l.orc_fene_force, "fene_anchor_force": l.orc_fene_anchor_force,
        "aa_force": l.orc_aa_force, "bb_force": l.orc_bb_force,
        "fene_potential": l.orc_fene_potential, "fene_anchor_potential": l.orc_fene_anchor_potential,
        "aa_potential": l.orc_aa_potential, "bb_potential": l.orc_bb_potential,
    }


def ref_lib():
    """The reference's own Potentials.h / Rand.cpp compiled in oracle/_ref,
    or None where it was not built (no reference tree)."""
    if not os.path.exists(REF_SO):
        return None
    l = C.CDLL(REF_SO)
    for n in ("ref_fene_force", "ref_fene_anchor_force", "ref_aa_force", "ref_fene_potential",
              "ref_fene_anchor_potential", "ref_aa_potential"):
        getattr(l, n).restype = C.c_double
        getattr(l, n).argtypes = [C.c_double]
    for n in ("ref_bb_force", "ref_bb_potential"):
        getattr(l, n).restype = C.c_double
        getattr(l, n).argtypes = [C.c_double, C.c_double]
    l.ref_rng_sequence.restype = None
    l.ref_rng_sequence.argtypes = [C.POINTER(C.c_int), _d, _d, C.c_int, _d]
    return l


def set_threads(n: int):
    """OpenMP threads of the oracle's solvent loops (Philox and injected
    modes; results do not depend on it)."""
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())
