"""Complex-symmetric field (PARDISO mtype 6) through the device-resident
entry points: rsl_refactor_device and rsl_refine_device on values, b and x
already in device memory against the host-buffer calls on the same inputs.
The device forms the modulus with its own hypot, which may differ from the
host's in the last bit, so the bar is the same pivot sequence, inertia and
perturbation count, L and D to 1e-12 relative, and the refinement's
solution to 1e-12 relative with every column certified."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import rslab_amd as R
from rslab_amd import matgen

from test_gpu_parity import _device_copy, csc

pytestmark = pytest.mark.gpu


def hip():
    return C.CDLL("libamdhip64.so.7")


@pytest.mark.parametrize("pivot", ["fail", "force_accept"])
def test_complex_refactor_device_matches_host_refactor(pivot):
    a, _ = matgen.config2(m=16, nrhs=1)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    if pivot == "force_accept":
        st = st.with_pivot(R.ZeroPivotAction.ForceAccept())
    sym = R.LdltSymbolic.analyze_with(a, st)
    f = sym.factor(a, st)
    ref = f.factors()
    ref_in, ref_np = f.inertia(), f.n_perturbed()
    f.refactor(a.with_values(a.values * 1.5 - 0.25))
    h = hip()
    dv = _device_copy(h, np.ascontiguousarray(a.values_f64()))
    try:
        f.refactor_device(dv.value)
    finally:
        h.hipFree(dv)
    got = f.factors()
    assert np.array_equal(got.perm, ref.perm) and np.array_equal(got.two_by_two, ref.two_by_two)
    assert f.inertia() == ref_in and f.n_perturbed() == ref_np
    lg, lr = csc(got), csc(ref)
    assert abs(lg - lr).max() <= 1e-12 * max(1.0, abs(lr).max())
    assert np.abs(got.d_diag - ref.d_diag).max() <= 1e-12 * np.abs(ref.d_diag).max()


def test_complex_refine_device_matches_host_refine():
    a, B = matgen.config2(m=16, nrhs=3)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    f = R.LdltSymbolic.analyze_with(a, st).factor(a, st)
    mixed = R.MixedLdltSolver(f, a)
    X, infos = mixed.solve_to(a, B, 1e-14)
    assert all(i.certified for i in infos), infos
    B = np.ascontiguousarray(B.astype(np.complex128))
    h = hip()
    da = _device_copy(h, np.ascontiguousarray(a.values_f64()))
    db = _device_copy(h, B)
    dx = _device_copy(h, np.zeros_like(B))
    try:
        dinfos = mixed.solve_to_device(da.value, db.value, 3, dx.value, 1e-14)
        Xd = np.empty_like(B)
        assert h.hipMemcpy(Xd.ctypes.data_as(C.c_void_p), dx, C.c_size_t(B.nbytes), 2) == 0
    finally:
        for d in (da, db, dx):
            h.hipFree(d)
    assert all(i.certified for i in dinfos), dinfos
    assert [i.ir_iters for i in dinfos] == [i.ir_iters for i in infos]
    assert np.abs(Xd - X).max() <= 1e-12 * np.abs(X).max()
    assert O.backward_error(a, Xd, B).max() <= 1e-14
