"""MI355X-native MPC-MD hot path of StarPolymersMD.

The compute path is libspmd.so (hand-written HIP for gfx950, C ABI in
include/spmd.h).  `MPCSystem` mirrors the reference's MPC / Thermostat_None /
Box calls on top of it.
"""
from ._lib import Params, SpmdError, make_params
from .system import MPCSystem

__all__ = ["MPCSystem", "Params", "SpmdError", "make_params"]
