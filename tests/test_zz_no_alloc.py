"""mpc_setstep_noalloc (no allocation, fixed device scratch) against setStep.

Two calls in a row on device copies of the same fields setStep reads from the
host, the second on a smaller field: the error norm agrees to 1e-12 and the
step decision (dzz, bad) is the same.  Sizes beyond the scratch's limit are
rejected before any device work.
"""
import ctypes as C

import numpy as np
import pytest

from mpc_amd import lib

pytestmark = pytest.mark.gpu

MPC_EINVAL = -1


def _dev(a):
    L = lib()
    p = C.c_void_p()
    assert L.mpc_dev_alloc(a.nbytes, C.byref(p)) == 0
    assert L.mpc_dev_upload(p, a.ctypes.data_as(C.c_void_p), a.nbytes) == 0
    return p


def _pair(rng, dims, rel):
    nx, ny, nt = dims
    shape = (nt, ny, nx)
    c2 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    c1 = c2 + rel * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    return np.ascontiguousarray(c1), np.ascontiguousarray(c2)


def _decide_host(dims, c1, c2, dz, def_err, dzmin, y1, y2):
    L = lib()
    d = np.array(dims, dtype=np.int32)
    dzz, bad, err = C.c_double(dz), C.c_int(-7), C.c_double(0)
    L.setStep(d.ctypes.data_as(C.POINTER(C.c_int)), C.byref(dzz), c1.ctypes.data_as(C.c_void_p),
              c2.ctypes.data_as(C.c_void_p), def_err, dzmin, y1.ctypes.data_as(C.POINTER(C.c_double)),
              y2.ctypes.data_as(C.POINTER(C.c_double)), C.byref(bad), C.byref(err))
    assert L.mpc_last_error() == 0
    return dzz.value, bad.value, err.value


def _decide_noalloc(dims, d1, d2, dz, def_err, dzmin, y1, y2):
    L = lib()
    d = np.array(dims, dtype=np.int32)
    dzz, bad, err = C.c_double(dz), C.c_int(-7), C.c_double(0)
    rc = L.mpc_setstep_noalloc(d.ctypes.data_as(C.POINTER(C.c_int)), C.byref(dzz), d1, d2, def_err, dzmin,
                               y1.ctypes.data_as(C.POINTER(C.c_double)), y2.ctypes.data_as(C.POINTER(C.c_double)),
                               C.byref(bad), C.byref(err))
    return rc, dzz.value, bad.value, err.value


def test_noalloc_setstep_matches_setstep_on_two_calls(gpu_lib):
    L = lib()
    rng = np.random.default_rng(31)
    y1 = np.array([0.0, 1e-3, 0.0, 0.0])
    y2 = np.array([0.0, 1e-3, 0.0, 0.0])
    # the first call on a larger field, the second on a smaller one, whose
    # chunks reuse the same scratch slots; errors chosen to take different
    # branches of the decision
    for dims, rel, def_err in (((64, 64, 512), 7e-4, 1e-3), ((16, 16, 256), 3e-3, 1e-3)):
        print(f"case noalloc setStep dims={dims} rel={rel}", flush=True)
        c1, c2 = _pair(rng, dims, rel)
        ref = _decide_host(dims, c1, c2, 1e-3, def_err, 1e-7, y1, y2)
        d1, d2 = _dev(c1), _dev(c2)
        try:
            rc, dzz, bad, err = _decide_noalloc(dims, d1, d2, 1e-3, def_err, 1e-7, y1, y2)
        finally:
            L.mpc_dev_free(d1)
            L.mpc_dev_free(d2)
        print(f"  err {err:.15e} vs {ref[2]:.15e}; dzz {dzz} vs {ref[0]}; bad {bad} vs {ref[1]}", flush=True)
        assert rc == 0
        assert abs(err - ref[2]) <= 1e-12 * ref[2]
        assert (dzz, bad) == (ref[0], ref[1])


def test_noalloc_setstep_rejects_fields_beyond_its_scratch(gpu_lib):
    """2^28 + 1 points and more are refused before the pointers are touched"""
    y = np.zeros(4)
    rc, _, _, _ = _decide_noalloc((1024, 1024, 257), C.c_void_p(16), C.c_void_p(16), 1e-3, 1e-3, 1e-7, y, y)
    assert rc == MPC_EINVAL
    rc, _, _, _ = _decide_noalloc((0, 4, 4), C.c_void_p(16), C.c_void_p(16), 1e-3, 1e-3, 1e-7, y, y)
    assert rc == MPC_EINVAL
