"""ctypes binding of libspmd.so, the C ABI declared in include/spmd.h.

The library is the only compute path: importing this module does not need a
GPU, but every call that touches device state goes through the HIP kernels
and raises SpmdError when the library is missing or the device is unusable.
There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspmd.so")

SPMD_OK = 0
ERRORS = {
    1: "SPMD_ERR_ARG", 2: "SPMD_ERR_HIP", 3: "SPMD_ERR_OOM", 4: "SPMD_ERR_CELL_OOB",
    5: "SPMD_ERR_VERLET_OOB", 6: "SPMD_ERR_VERLET_OVERFLOW", 7: "SPMD_ERR_UNSUPPORTED",
    8: "SPMD_ERR_STATE", 9: "SPMD_ERR_RANK_OVERFLOW", 10: "SPMD_ERR_NO_DEVICE", 11: "SPMD_ERR_SCAN",
}


class SpmdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    """spmd_params (include/spmd.h)."""
    _fields_ = [
        ("L", C.c_int32 * 3), ("kT", C.c_double), ("n_c", C.c_int32), ("shear", C.c_double),
        ("step_update", C.c_uint32), ("alpha_deg", C.c_double), ("lambda_", C.c_double),
        ("md_dt", C.c_double), ("seed", C.c_uint64), ("angular_momentum", C.c_int32),
    ]


def make_params(L=(20, 20, 20), kT=1.0, n_c=10, shear=0.0, step_update=1, alpha_deg=130.0,
                lam=1.0, md_dt=0.01, seed=1, angular_momentum=0) -> Params:
    p = Params()
    for i in range(3):
        p.L[i] = int(L[i])
    p.kT, p.n_c, p.shear, p.step_update = float(kT), int(n_c), float(shear), int(step_update)
    p.alpha_deg, p.lambda_, p.md_dt, p.seed = float(alpha_deg), float(lam), float(md_dt), int(seed)
    p.angular_momentum = int(angular_momentum)
    return p


_P = C.POINTER
_d = _P(C.c_double)
_f = _P(C.c_float)
_u32 = _P(C.c_uint32)
_SIGS = {
    "spmd_error_string": (C.c_char_p, [C.c_int]),
    "spmd_last_error": (C.c_char_p, [C.c_void_p]),
    "spmd_version": (C.c_char_p, []),
    "spmd_create": (C.c_int, [_P(Params), _P(C.c_void_p)]),
    "spmd_create_batch": (C.c_int, [_P(Params), C.c_uint32, _P(C.c_void_p)]),
    "spmd_destroy": (C.c_int, [C.c_void_p]),
    "spmd_num_replicas": (C.c_int, [C.c_void_p, _u32]),
    "spmd_set_md_dt": (C.c_int, [C.c_void_p, C.c_double]),
    "spmd_fluid_initialize": (C.c_int, [C.c_void_p]),
    "spmd_num_fluid": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_uint64)]),
    "spmd_set_fluid": (C.c_int, [C.c_void_p, C.c_uint32, _f, _f, C.c_uint64]),
    "spmd_get_fluid": (C.c_int, [C.c_void_p, C.c_uint32, _f, _f, _u32]),
    "spmd_set_tracking": (C.c_int, [C.c_void_p, C.c_int]),
    "spmd_add_star": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                C.c_double, C.c_double, C.c_double]),
    "spmd_add_star_coil": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_double, C.c_double, C.c_double]),
    "spmd_num_monomers": (C.c_int, [C.c_void_p, C.c_uint32, _u32]),
    "spmd_set_monomers": (C.c_int, [C.c_void_p, C.c_uint32, _d, _d]),
    "spmd_get_monomers": (C.c_int, [C.c_void_p, C.c_uint32, _d, _d, _d]),
    "spmd_get_monomer_masses": (C.c_int, [C.c_void_p, C.c_uint32, _d]),
    "spmd_md_init": (C.c_int, [C.c_void_p]),
    "spmd_md_propagate": (C.c_int, [C.c_void_p, C.c_int]),
    "spmd_mpc_step": (C.c_int, [C.c_void_p, C.c_int64, C.c_double]),
    "spmd_run": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_double]),
    "spmd_sync": (C.c_int, [C.c_void_p]),
    "spmd_set_delrx": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double]),
    "spmd_get_delrx": (C.c_int, [C.c_void_p, C.c_uint32, _d]),
    "spmd_inject_collision_randoms": (C.c_int, [C.c_void_p, C.c_uint32, _d, _d, _d]),
    "spmd_last_collision": (C.c_int, [C.c_void_p, C.c_uint32, _d, _P(C.c_int), _d, _d]),
    "spmd_get_cell_index": (C.c_int, [C.c_void_p, C.c_uint32, _u32, _u32]),
    "spmd_temperature": (C.c_int, [C.c_void_p, C.c_uint32, _d]),
    "spmd_cell_fluid_velocity": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32, C.c_int32, _d]),
    "spmd_epot": (C.c_int, [C.c_void_p, C.c_uint32, _d]),
    "spmd_md_stats": (C.c_int, [C.c_void_p, C.c_uint32, _u32, _u32]),
    "spmd_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "spmd_get_tuning": (C.c_int, [C.c_void_p, _P(C.c_int), _P(C.c_int)]),
    "spmd_set_collision_path": (C.c_int, [C.c_void_p, C.c_int]),
    "spmd_get_collision_path": (C.c_int, [C.c_void_p, _P(C.c_int)]),
    "spmd_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "spmd_profile_read": (C.c_int, [C.c_void_p, C.c_char_p, _d, _P(C.c_int64)]),
    "spmd_profile_reset": (C.c_int, [C.c_void_p]),
    "spmd_mlp_create": (C.c_int, [C.c_uint64, _P(C.c_void_p)]),
    "spmd_mlp_destroy": (C.c_int, [C.c_void_p]),
    "spmd_mlp_last_error": (C.c_char_p, [C.c_void_p]),
    "spmd_mlp_num_params": (C.c_uint32, []),
    "spmd_mlp_get_params": (C.c_int, [C.c_void_p, _f]),
    "spmd_mlp_set_params": (C.c_int, [C.c_void_p, _f]),
    "spmd_mlp_set_norm": (C.c_int, [C.c_void_p, _f, _f, _f]),
    "spmd_mlp_get_norm": (C.c_int, [C.c_void_p, _f, _f, _f]),
    "spmd_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "spmd_mlp_forward_unfused": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "spmd_mlp_gradient": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, _f, _d]),
    "spmd_mlp_gradient_unfused": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, _f, _d]),
    "spmd_mlp_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_double, _d]),
    "spmd_mlp_fit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double,
                               C.c_uint64, _d]),
    "spmd_dev_alloc": (C.c_int, [C.c_uint64, _P(C.c_void_p)]),
    "spmd_dev_free": (C.c_int, [C.c_void_p]),
    "spmd_dev_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "spmd_dev_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "spmd_closure_features": (C.c_int, [C.c_void_p, C.c_void_p]),
    "spmd_closure_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, _f, _f, _u32]),
    "spmd_gyration_tensor": (C.c_int, [C.c_void_p, C.c_uint32, _d]),
    "spmd_rotation_frequency": (C.c_int, [C.c_void_p, C.c_uint32, _d]),
    "spmd_patches": (C.c_int, [C.c_void_p, C.c_uint32, _d, _d]),
    "spmd_statistics": (C.c_int, [C.c_void_p, C.c_uint32, _d]),
    "spmd_velocity_profile": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, C.c_uint32, _P(C.c_int32), _u32, _d]),
    "spmd_velocity_profile_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, C.c_uint32, _P(C.c_int32),
                                                _u32, C.c_void_p]),
    "spmd_fourier_modes": (C.c_int, [C.c_void_p, C.c_uint32, _d]),
    "spmd_speed_histogram": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, C.c_uint32, _P(C.c_uint64)]),
    "spmd_fluid_vacf_sample": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "spmd_fluid_vacf_read": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), _P(C.c_uint64)]),
    "spmd_fluid_vacf_reset": (C.c_int, [C.c_void_p, C.c_uint32]),
    "spmd_get_params": (C.c_int, [C.c_void_p, C.c_uint32, _P(Params)]),
    "spmd_star_layout": (C.c_int, [C.c_void_p, C.c_uint32, _u32, _u32, _u32, _u32]),
    "spmd_write_fluid": (C.c_int, [C.c_void_p, C.c_uint32, C.c_char_p]),
    "spmd_read_fluid": (C.c_int, [C.c_void_p, C.c_uint32, C.c_char_p, _P(C.c_int32), _P(C.c_uint64)]),
    "spmd_write_pdb": (C.c_int, [C.c_void_p, C.c_uint32, C.c_char_p, C.c_int64, C.c_int]),
    "spmd_add_star_pdb": (C.c_int, [C.c_void_p, C.c_uint32, C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_double, C.c_int]),
    "spmd_snapshot_save": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "spmd_snapshot_load": (C.c_int, [C.c_char_p, _P(C.c_void_p), _P(C.c_int64)]),
}

_lib = None


def load() -> C.CDLL:
    """Load libspmd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SpmdError(2, f"{LIB_PATH} not built: run __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def ptr(a: np.ndarray | None, ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))
