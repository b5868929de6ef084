"""spmd_velocity_profile_noalloc (no counterpart in the reference: the
velocity profile for callers that cannot allocate device memory) against
spmd_velocity_profile on the same states, through the C ABI.

Bar: bins, keys and sums identical bit for bit (the same fixed-point sums;
only their home and the conversion differ).  Two calls in a row share the
static device scratch, the second on a smaller replica (fewer particles and
bins), so a bin the first call filled must not leak into the second; the
device buffer entries beyond the bins stay as the caller left them.
"""
import ctypes as C

import numpy as np
import pytest

from starpolymersmd_amd import MPCSystem, make_params
from starpolymersmd_amd._lib import load, ptr

pytestmark = pytest.mark.gpu


def _regular(g, r, width, max_bins):
    lib = load()
    b0 = np.zeros(2, np.int32)
    nb = np.zeros(2, np.uint32)
    sums = np.full((2, max_bins, 3), -7.0)
    rc = lib.spmd_velocity_profile(g.h, r, width, max_bins, ptr(b0, C.c_int32), ptr(nb, C.c_uint32),
                                   ptr(sums, C.c_double))
    assert rc == 0, g._lib.spmd_last_error(g.h)
    return b0, nb, sums


def _noalloc(g, r, width, max_bins, dev):
    lib = load()
    b0 = np.zeros(2, np.int32)
    nb = np.zeros(2, np.uint32)
    rc = lib.spmd_velocity_profile_noalloc(g.h, r, width, max_bins, ptr(b0, C.c_int32), ptr(nb, C.c_uint32), dev)
    assert rc == 0, g._lib.spmd_last_error(g.h)
    g.sync()
    sums = np.empty((2, max_bins, 3))
    assert lib.spmd_dev_download(sums.ctypes.data_as(C.c_void_p), dev, sums.nbytes) == 0
    return b0, nb, sums


def test_noalloc_profile_equals_regular_on_two_calls():
    lib = load()
    prm = [make_params(L=(24, 24, 24), n_c=10, shear=0.05, step_update=10, seed=3, md_dt=0.01),
           make_params(L=(16, 16, 16), n_c=10, shear=0.02, step_update=10, seed=4, md_dt=0.01)]
    g = MPCSystem(prm)
    g.add_star(4, 4, 6, 10.0, r=0, coil=True)
    g.add_star(3, 3, 5, 10.0, r=1, coil=True)
    g.md_init()
    g.initialize()
    g.run(0, 31, 0.1)  # three collisions; one substep and delrx advance left pending
    max_bins = 64
    dev = C.c_void_p()
    assert lib.spmd_dev_alloc(2 * max_bins * 3 * 8, C.byref(dev)) == 0
    try:
        fill = np.full((2, max_bins, 3), 123.5)
        assert lib.spmd_dev_upload(dev, fill.ctypes.data_as(C.c_void_p), fill.nbytes) == 0
        # first call: the larger replica, bins of width 0.5 (49 per axis)
        b0n, nbn, sn = _noalloc(g, 0, 0.5, max_bins, dev)
        b0r, nbr, sr = _regular(g, 0, 0.5, max_bins)
        assert np.array_equal(b0n, b0r) and np.array_equal(nbn, nbr) and nbn[0] == 49
        for a in range(2):
            assert np.array_equal(sn[a, :nbn[a]], sr[a, :nbr[a]])
            assert np.all(sn[a, nbn[a]:] == 123.5)
        assert sn[0, :nbn[0], 2].sum() == g.num_fluid(0)
        # second call: the smaller replica, wider bins (17 per axis)
        b0n, nbn, sn = _noalloc(g, 1, 1.0, max_bins, dev)
        b0r, nbr, sr = _regular(g, 1, 1.0, max_bins)
        assert np.array_equal(b0n, b0r) and np.array_equal(nbn, nbr) and nbn[0] == 17
        for a in range(2):
            assert np.array_equal(sn[a, :nbn[a]], sr[a, :nbr[a]])
        assert sn[1, :nbn[1], 2].sum() == g.num_fluid(1)
        # more bins than the scratch holds, or than the caller's buffer: rejected
        b0 = np.zeros(2, np.int32)
        nb = np.zeros(2, np.uint32)
        rc = lib.spmd_velocity_profile_noalloc(g.h, 0, 0.01, 4096, ptr(b0, C.c_int32), ptr(nb, C.c_uint32), dev)
        assert rc != 0 and b"1024 bins" in g._lib.spmd_last_error(g.h)
        rc = lib.spmd_velocity_profile_noalloc(g.h, 0, 0.5, 16, ptr(b0, C.c_int32), ptr(nb, C.c_uint32), dev)
        assert rc != 0 and b"max_bins" in g._lib.spmd_last_error(g.h)
        g.sync()
    finally:
        lib.spmd_dev_free(dev)
