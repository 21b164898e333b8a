"""rsl_solve_device_noalloc (no allocation; its scratch is the library's fixed
device array) against rsl_solve_device (scratch allocated by the handle) on
the same factor and inputs, through the C ABI: equal bit for bit on two calls
in a row, the second with a smaller system and fewer right-hand sides (the
array keeps the first call's contents), and a solve larger than the array is
rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

import rslab_amd as R
from rslab_amd import _lib, matgen

pytestmark = pytest.mark.gpu


def _hip():
    return C.CDLL("libamdhip64.so.7")


def _solve_both(hip, f, B, stream):
    """(x from rsl_solve_device, x from rsl_solve_device_noalloc), both on `stream`."""
    n, k = B.shape
    B = np.ascontiguousarray(B)
    db, dx1, dx2 = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for p in (db, dx1, dx2):
        assert hip.hipMalloc(C.byref(p), C.c_size_t(B.nbytes)) == 0
    try:
        assert hip.hipMemcpy(db, B.ctypes.data_as(C.c_void_p), C.c_size_t(B.nbytes), 1) == 0
        f.solve_device(db.value, k, dx1.value, stream.value)
        f.solve_device_noalloc(db.value, k, dx2.value, stream.value)
        X1, X2 = np.empty_like(B), np.empty_like(B)
        assert hip.hipMemcpyAsync(X1.ctypes.data_as(C.c_void_p), dx1, C.c_size_t(B.nbytes), 2, stream) == 0
        assert hip.hipMemcpyAsync(X2.ctypes.data_as(C.c_void_p), dx2, C.c_size_t(B.nbytes), 2, stream) == 0
        assert hip.hipStreamSynchronize(stream) == 0
    finally:
        for p in (db, dx1, dx2):
            hip.hipFree(p)
    return X1, X2


def test_noalloc_solve_equals_solve_device_on_two_calls():
    hip = _hip()
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    try:
        # call 1: config-3 shape, 20^3, nested dissection, 6 right-hand sides
        a1, B1 = matgen.config3(m=20, nrhs=6)
        st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
        f1 = R.LdltSymbolic.analyze_with(a1, st).factor(a1)
        X1, X1n = _solve_both(hip, f1, B1, stream)
        assert np.array_equal(X1, X1n)
        assert np.array_equal(X1n, f1.solve_many(B1))
        # call 2: a smaller system (2D grid, AMD) with 2 right-hand sides; the
        # fixed array still holds call 1's values
        a2 = matgen.grid_laplacian((30, 30), shift=-2.5)
        B2 = np.stack([matgen.rng_uniform(91 + c, -1, 1, a2.n) for c in range(2)], 1)
        f2 = R.LdltSolver.factor_with(a2)
        X2, X2n = _solve_both(hip, f2, B2, stream)
        assert np.array_equal(X2, X2n)
        assert np.abs(a2.symv(X2n) - B2).max() < 1e-10 * np.abs(B2).max() * max(1.0, np.abs(X2n).max())
    finally:
        hip.hipStreamDestroy(stream)


def test_noalloc_solve_rejects_sizes_beyond_the_array():
    a = matgen.grid_laplacian((30, 30), shift=-2.5)
    f = R.LdltSolver.factor_with(a)
    per_rhs = 8 * (2 * a.n + R.LdltSymbolic.analyze(a).info()["solve_scratch_rows"])
    too_many = int(np.ceil((128 << 20) / per_rhs)) + 1
    with pytest.raises(R.InvalidInput, match="bytes of scratch"):
        # never reaches the device: the size check comes first (pointers unused)
        f.solve_device_noalloc(8, too_many, 8)


def test_header_states_the_scratch_limit():
    import os
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "rslab_amd.h")).read()
    assert "#define RSL_SOLVE_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)" in hdr
    assert "rsl_solve_device_noalloc" in _lib.SIGNATURES
