"""Python mirror of the reference's MPC / Thermostat_None / Box hot-path API.

`MPCSystem` owns one handle of libspmd.so (one or more replicas) and exposes
the reference's call names (MPC::initialize, MPC::step, MPC::Fluid,
Thermostat_None::propagate, Box::add_star, Box::Molecules ...) with numpy
arrays at the boundary.  Every method calls the HIP library; nothing is
computed on the host.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Params, SpmdError, load, make_params, ptr


class MPCSystem:
    def __init__(self, params: Params | list[Params]):
        lib = load()
        self._lib = lib
        h = C.c_void_p()
        if isinstance(params, Params):
            rc = lib.spmd_create(C.byref(params), C.byref(h))
            self.n_replicas = 1
            self.params = [params]
        else:
            arr = (Params * len(params))(*params)
            rc = lib.spmd_create_batch(arr, len(params), C.byref(h))
            self.n_replicas = len(params)
            self.params = list(params)
        if rc:
            raise SpmdError(rc, lib.spmd_error_string(rc).decode())
        self.h = h

    @classmethod
    def load_snapshot(cls, path: str):
        """A handle restored from spmd_snapshot_save; returns (system, step)."""
        lib = load()
        h = C.c_void_p()
        step = C.c_int64()
        rc = lib.spmd_snapshot_load(str(path).encode(), C.byref(h), C.byref(step))
        if rc:
            raise SpmdError(rc, lib.spmd_error_string(rc).decode())
        self = cls.__new__(cls)
        self._lib = lib
        self.h = h
        n = C.c_uint32()
        lib.spmd_num_replicas(h, C.byref(n))
        self.n_replicas = n.value
        self.params = []
        for r in range(self.n_replicas):
            p = Params()
            lib.spmd_get_params(h, r, C.byref(p))
            self.params.append(p)
        return self, step.value

    def save_snapshot(self, path: str, step: int = 0):
        """Lossless binary snapshot (parameters, solvent, stars, delrx,
        Philox counters); a loaded copy continues bit for bit."""
        self._chk(self._lib.spmd_snapshot_save(self.h, str(path).encode(), step))

    # ------------------------------------------------------- restart files
    def write_fluid(self, path: str, r: int = 0):
        """MPC::print_fluid_complete (MPC.cpp:570-579)."""
        self._chk(self._lib.spmd_write_fluid(self.h, r, str(path).encode()))

    def read_fluid(self, path: str, old_box=None, r: int = 0) -> int:
        """MPC::initialize(filename, Box_old) (MPC.cpp:57-113); returns the
        number of particles taken from the file."""
        L = [int(l) for l in self.params[r].L] if old_box is None else list(old_box)
        ob = (C.c_int32 * 3)(*L)
        n = C.c_uint64()
        self._chk(self._lib.spmd_read_fluid(self.h, r, str(path).encode(), ob, C.byref(n)))
        return n.value

    def write_pdb(self, path: str, step: int, r: int = 0, append: bool = False):
        """Box::print_PDB_with_velocity (Box.cpp:563-602)."""
        self._chk(self._lib.spmd_write_pdb(self.h, r, str(path).encode(), step, int(append)))

    def add_star_pdb(self, path: str, A: int, B: int, arms: int, mass: float, set_zero: bool = True, r: int = 0):
        """Box::add_star(filename, A, B, Arms, Mass, set_zero) (Box.cpp:48-54)."""
        self._chk(self._lib.spmd_add_star_pdb(self.h, r, str(path).encode(), A, B, arms, mass, int(set_zero)))

    def star_layout(self, r: int = 0):
        """(monomers, arms, A, B) of a replica's star."""
        v = [C.c_uint32() for _ in range(4)]
        self._chk(self._lib.spmd_star_layout(self.h, r, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    # ------------------------------------------------------------ helpers
    def _chk(self, rc: int):
        if rc:
            raise SpmdError(rc, self._lib.spmd_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self._lib.spmd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ solvent
    def num_fluid(self, r: int = 0) -> int:
        n = C.c_uint64()
        self._chk(self._lib.spmd_num_fluid(self.h, r, C.byref(n)))
        return n.value

    def initialize(self):
        """MPC::initialize() (MPC.cpp:34-55) for every replica."""
        self._chk(self._lib.spmd_fluid_initialize(self.h))

    def set_fluid(self, pos: np.ndarray, vel: np.ndarray, r: int = 0):
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        vel = np.ascontiguousarray(vel, dtype=np.float32)
        self._chk(self._lib.spmd_set_fluid(self.h, r, ptr(pos, C.c_float), ptr(vel, C.c_float), pos.shape[0]))

    def get_fluid(self, r: int = 0, ids: bool = False):
        n = self.num_fluid(r)
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        idv = np.empty(n, np.uint32) if ids else None
        self._chk(self._lib.spmd_get_fluid(self.h, r, ptr(pos, C.c_float), ptr(vel, C.c_float),
                                           ptr(idv, C.c_uint32)))
        return (pos, vel, idv) if ids else (pos, vel)

    def set_tracking(self, on: bool = True):
        self._chk(self._lib.spmd_set_tracking(self.h, int(on)))

    # ------------------------------------------------------------ star
    def add_star(self, A, B, arms, mass, bond=1.01, anchor_bond=2.0, r: int = 0, coil: bool = False):
        """Box::add_star (Box.cpp:39-46); coil=True places the arms as seeded
        random coils instead of straight rays."""
        fn = self._lib.spmd_add_star_coil if coil else self._lib.spmd_add_star
        self._chk(fn(self.h, r, A, B, arms, mass, bond, anchor_bond))

    def num_monomers(self, r: int = 0) -> int:
        n = C.c_uint32()
        self._chk(self._lib.spmd_num_monomers(self.h, r, C.byref(n)))
        return n.value

    def get_monomers(self, r: int = 0):
        n = self.num_monomers(r)
        pos, vel, frc = (np.empty((n, 3)) for _ in range(3))
        self._chk(self._lib.spmd_get_monomers(self.h, r, ptr(pos, C.c_double), ptr(vel, C.c_double),
                                              ptr(frc, C.c_double)))
        return pos, vel, frc

    def set_monomers(self, pos=None, vel=None, r: int = 0):
        p = None if pos is None else np.ascontiguousarray(pos, dtype=np.float64)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
        self._chk(self._lib.spmd_set_monomers(self.h, r, ptr(p, C.c_double), ptr(v, C.c_double)))

    # ------------------------------------------------------------ stepping
    def md_init(self):
        """Thermostat_None ctor (Thermostat_None.cpp:5-12)."""
        self._chk(self._lib.spmd_md_init(self.h))

    def set_md_dt(self, dt: float):
        self._chk(self._lib.spmd_set_md_dt(self.h, dt))

    def propagate(self, calc_epot: bool = False):
        """Thermostat_None::propagate (Thermostat_None.cpp:21-38)."""
        self._chk(self._lib.spmd_md_propagate(self.h, int(calc_epot)))

    def step(self, t: int, dt: float):
        """MPC::step (MPC.cpp:116-142)."""
        self._chk(self._lib.spmd_mpc_step(self.h, t, dt))

    def run(self, t0: int, n: int, mpc_dt: float):
        """star-polymers.cpp:412-536 loop body, n calls from t0."""
        self._chk(self._lib.spmd_run(self.h, t0, n, mpc_dt))

    def sync(self):
        self._chk(self._lib.spmd_sync(self.h))

    # ------------------------------------------------------------ hooks
    def set_delrx(self, v: float, r: int = 0):
        self._chk(self._lib.spmd_set_delrx(self.h, r, v))

    def get_delrx(self, r: int = 0) -> float:
        v = C.c_double()
        self._chk(self._lib.spmd_get_delrx(self.h, r, C.byref(v)))
        return v.value

    def inject(self, shift, axis_u, gamma, r: int = 0):
        s = np.ascontiguousarray(shift, dtype=np.float64)
        a = np.ascontiguousarray(axis_u, dtype=np.float64)
        g = np.ascontiguousarray(gamma, dtype=np.float64)
        self._chk(self._lib.spmd_inject_collision_randoms(self.h, r, ptr(s, C.c_double), ptr(a, C.c_double),
                                                          ptr(g, C.c_double)))

    def last_collision(self, r: int = 0):
        sh, cp, cv = np.empty(3), np.empty(3), np.empty(3)
        rec = C.c_int()
        self._chk(self._lib.spmd_last_collision(self.h, r, ptr(sh, C.c_double), C.byref(rec),
                                                ptr(cp, C.c_double), ptr(cv, C.c_double)))
        return sh, bool(rec.value), cp, cv

    def cell_index(self, r: int = 0):
        n = self.num_fluid(r)
        fc = np.empty(n, np.uint32)
        mc = np.empty(max(self.num_monomers(r), 1), np.uint32)
        self._chk(self._lib.spmd_get_cell_index(self.h, r, ptr(fc, C.c_uint32), ptr(mc, C.c_uint32)))
        return fc, mc[: self.num_monomers(r)]

    # ------------------------------------------------------------ observables
    def temperature(self, r: int = 0) -> float:
        v = C.c_double()
        self._chk(self._lib.spmd_temperature(self.h, r, C.byref(v)))
        return v.value

    def cell_fluid_velocity(self, z_start: int, z_stop: int, r: int = 0) -> np.ndarray:
        """MPC::calculateFluidVelocity (MPC.cpp:334-342) of the cells with
        z_start <= z <= z_stop: array [z, y, x, 3] of the mean solvent
        velocity per cell of the last collision (0 where empty)."""
        L = [int(l) + int(l) % 2 for l in self.params[r].L]   # even-rounded box (Box.cpp:14-19)
        out = np.zeros(((z_stop - z_start + 1) * L[0] * L[1] * 3,), dtype=np.float64)
        self._chk(self._lib.spmd_cell_fluid_velocity(self.h, r, z_start, z_stop, ptr(out, C.c_double)))
        return out.reshape(z_stop - z_start + 1, L[1], L[0], 3)

    def epot(self, r: int = 0) -> float:
        v = C.c_double()
        self._chk(self._lib.spmd_epot(self.h, r, C.byref(v)))
        return v.value

    def gyration_tensor(self, r: int = 0) -> np.ndarray:
        """Box::calculate_gyration_tensor (Box.cpp:330-352), 3x3."""
        g = np.empty(9)
        self._chk(self._lib.spmd_gyration_tensor(self.h, r, ptr(g, C.c_double)))
        return g.reshape(3, 3)

    def rotation_frequency(self, r: int = 0) -> np.ndarray:
        """Box::calculate_rotation_frequency (Box.cpp:354-379)."""
        w = np.empty(3)
        self._chk(self._lib.spmd_rotation_frequency(self.h, r, ptr(w, C.c_double)))
        return w

    def patches(self, r: int = 0):
        """Box::calculate_patches_new (Box.cpp:434-507): (number, mean size)."""
        n, sz = C.c_double(), C.c_double()
        self._chk(self._lib.spmd_patches(self.h, r, C.byref(n), C.byref(sz)))
        return n.value, sz.value

    def statistics(self, r: int = 0) -> np.ndarray:
        """The 18 values of a statistics*.dat row after the step number
        (star-polymers.cpp:470-482): Epot, Ekin, Temp, n_p, s_p, R_gyr,
        G (9), omega (3)."""
        row = np.empty(18)
        self._chk(self._lib.spmd_statistics(self.h, r, ptr(row, C.c_double)))
        return row

    def velocity_profile(self, width: float, r: int = 0):
        """VelocityX sample (Velocity_x.h:76-96) of the current solvent:
        [(keys, sums)] for the x- and the y-binned map, sums[:, 0:3] = sum
        vx, sum vy, count."""
        L = [int(l) + int(l) % 2 for l in self.params[r].L]   # even-rounded box (Box.cpp:14-19)
        nmax = max(int(np.floor(L[a] / 2 / width)) - int(np.floor(-L[a] / 2 / width)) + 1 for a in range(2))
        b0 = np.empty(2, np.int32)
        nb = np.empty(2, np.uint32)
        sums = np.zeros((2, nmax, 3))
        self._chk(self._lib.spmd_velocity_profile(self.h, r, width, nmax, ptr(b0, C.c_int32), ptr(nb, C.c_uint32),
                                                  ptr(sums, C.c_double)))
        out = []
        for a in range(2):
            keys = (b0[a] + np.arange(nb[a])) * width
            out.append((keys, sums[a, : nb[a]].copy()))
        return out

    def fourier_modes(self, r: int = 0) -> np.ndarray:
        """Fourier_Autocorr sample (Fourier_Autocorr.h:55-72): complex (3, 3),
        row j = transverse velocity mode at k = 2 pi / L e_j."""
        m = np.empty(18)
        self._chk(self._lib.spmd_fourier_modes(self.h, r, ptr(m, C.c_double)))
        return m[0::2].reshape(3, 3) + 1j * m[1::2].reshape(3, 3)

    def speed_histogram(self, width: float, nbins: int, r: int = 0) -> np.ndarray:
        """VelocityDistribution counts (Vel_distribution.h:35-40); the last
        entry counts the speeds beyond nbins * width."""
        c = np.empty(nbins + 1, np.uint64)
        self._chk(self._lib.spmd_speed_histogram(self.h, r, width, nbins, ptr(c, C.c_uint64)))
        return c

    def fluid_vacf_sample(self, entries: int, r: int = 0):
        """Vel_Autocorr_Fluid::operator() (Vel_Autocorr_Fluid.h:44-66) over
        the solvent; the windows stay on the device (tracking must be on)."""
        self._chk(self._lib.spmd_fluid_vacf_sample(self.h, r, entries))

    def fluid_vacf_read(self, entries: int, r: int = 0):
        """(lag sums over particles and calls, calls that made them)."""
        s = np.empty(entries, np.float64)
        n = C.c_uint64()
        self._chk(self._lib.spmd_fluid_vacf_read(self.h, r, ptr(s, C.c_double), C.byref(n)))
        return s, int(n.value)

    def fluid_vacf_reset(self, r: int = 0):
        self._chk(self._lib.spmd_fluid_vacf_reset(self.h, r))

    def closure_features(self, feat_dev_ptr: int):
        """24 closure features of every monomer into device memory
        (float32, total monomers x 24)."""
        self._chk(self._lib.spmd_closure_features(self.h, C.c_void_p(feat_dev_ptr)))

    def closure_run(self, mlp, n_intervals: int, fmin=None, fmax=None):
        """n_intervals of MD + the network's solvent kick; returns the number
        of monomers per replica whose features left [fmin, fmax]."""
        oor = np.zeros(self.n_replicas, np.uint32)
        lo = None if fmin is None else np.ascontiguousarray(fmin, np.float32)
        hi = None if fmax is None else np.ascontiguousarray(fmax, np.float32)
        self._chk(self._lib.spmd_closure_run(self.h, mlp.h, int(n_intervals), ptr(lo, C.c_float),
                                             ptr(hi, C.c_float), ptr(oor, C.c_uint32)))
        return oor

    def md_stats(self, r: int = 0):
        """(Verlet rebuilds so far, entries of the current half lists)."""
        a, b = C.c_uint32(), C.c_uint32()
        self._chk(self._lib.spmd_md_stats(self.h, r, C.byref(a), C.byref(b)))
        return a.value, b.value

    # ------------------------------------------------------------ tuning/profiling
    def set_tuning(self, ppt: int, tile_cells: int):
        self._chk(self._lib.spmd_set_tuning(self.h, ppt, tile_cells))

    def get_tuning(self):
        a, b = C.c_int(), C.c_int()
        self._chk(self._lib.spmd_get_tuning(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    PATH_AUTO, PATH_SORT, PATH_ROWS = 0, 1, 2

    def set_collision_path(self, path: int):
        """PATH_AUTO (row tiles once the buffer is in cell order) or PATH_SORT."""
        self._chk(self._lib.spmd_set_collision_path(self.h, path))

    def collision_path(self) -> int:
        """The path the next collision takes: PATH_SORT or PATH_ROWS."""
        a = C.c_int()
        self._chk(self._lib.spmd_get_collision_path(self.h, C.byref(a)))
        return a.value

    def profile(self, on: bool = True):
        self._chk(self._lib.spmd_profile_enable(self.h, int(on)))

    def profile_reset(self):
        self._chk(self._lib.spmd_profile_reset(self.h))

    def profile_read(self, name: str):
        ms = C.c_double()
        n = C.c_int64()
        self._chk(self._lib.spmd_profile_read(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value


__all__ = ["MPCSystem", "make_params", "Params", "SpmdError"]
