"""Device paths selected by settings, each against the CPU oracle on the same
symbolic analysis and seeded inputs, through the C ABI:

* drop tolerance at emit (SolverSettings::with_drop_tol,
  multifrontal_ldlt.rs:640-643; applied in ldlt_emit_and_free :2344-2366):
  the device's stored L equals the oracle's with the same tau, and the
  reference's "crippled factor must not certify" case (mixed.rs:504-523);
* the gathered-tile cmod (the path a level takes when its product blocks do
  not fit the product buffer), forced by capping the buffer;
* the rank-deficiency report names the supernode and column of the first
  zero pivot;
* rsl_solve_device's ordering against the caller's stream and later calls on
  the same handle.
"""
import ctypes as C
import dataclasses
import re

import numpy as np
import pytest

import oracle as O
import rslab_amd as R
from rslab_amd import matgen

from test_gpu_parity import assert_same_factor, check_solution, dense_lower_csc, factor_pair, rhs

pytestmark = pytest.mark.gpu


def _factor_pair_drop(a, st, tau):
    sym = R.LdltSymbolic.analyze_with(a, st)
    fg = sym.factor(a, st)
    s = O.onepass_scale(a)
    fo = O.factor(sym.export(), O.scaled(a, s), panel_nb=st.panel_nb, drop_tol=tau)
    return sym, fg, fo, s


@pytest.mark.parametrize("tau", [0.0, 0.05, 0.5])
def test_drop_tol_matches_oracle_dense_front(tau):
    """A dense indefinite front (stable pivot sequence): the dropped factor
    keeps exactly the oracle's entries, and solves equal the oracle's."""
    rng = np.random.default_rng(11)
    M = rng.standard_normal((260, 260))
    M = M + M.T
    a = dense_lower_csc(M)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.Natural).with_drop_tol(tau)
    sym, fg, fo, s = _factor_pair_drop(a, st, tau)
    g = fg.factors()
    assert_same_factor(g, fo)
    assert np.array_equal(g.l_col_ptr, fo.l_col_ptr) and np.array_equal(g.l_row_idx, fo.l_row_idx)
    B = rhs(a, 2, 71)
    X = fg.solve_many(B)
    Xo = fo.solve_many(B * s[:, None]) * s[:, None]
    assert np.abs(X - Xo).max() <= 1e-9 * max(np.abs(Xo).max(), 1.0)
    if tau == 0.0:  # tau = 0 keeps every nonzero: the plain factor
        plain = sym.factor(a, dataclasses.replace(st, drop_tol=None))
        assert np.array_equal(plain.factors().l_values, g.l_values)


def test_drop_tol_matches_oracle_sparse_nd():
    """Config-3 shape (16^3, nested dissection) with tau = 0.02: the same
    entries dropped as the oracle, refinement against A still certifies."""
    a, B = matgen.config3(m=16, nrhs=2)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND).with_drop_tol(0.02)
    sym, fg, fo, s = _factor_pair_drop(a, st, 0.02)
    g = fg.factors()
    assert np.array_equal(g.perm, fo.perm) and g.inertia.as_tuple() == fo.inertia
    assert np.array_equal(g.l_col_ptr, fo.l_col_ptr) and np.array_equal(g.l_row_idx, fo.l_row_idx)
    full = sym.factor(a, dataclasses.replace(st, drop_tol=None)).factors()
    assert g.l_row_idx.size < full.l_row_idx.size  # something was dropped
    X, infos = R.MixedLdltSolver(fg, a).solve_to(a, B, 1e-14)
    assert all(i.certified for i in infos), infos


def test_crippled_factor_reports_uncertified():
    """mixed.rs:504-523: lap2d(90), PerturbToEps(1e-12) and drop tolerance 0.9
    leave a near-useless preconditioner; IR diverges and FGMRES runs out of
    its 120 iterations, so the result must not be certified, and the flag
    agrees with the reported backward error."""
    a = matgen.grid_laplacian((90, 90))
    b = np.array([(i % 5) - 2.0 for i in range(a.n)])
    opts = R.SolverSettings.preconditioner(1e-12).with_drop_tol(0.9).with_precision(R.Precision.FP32_SCHUR)
    m = R.MixedLdltSolver.factor_with(a, opts)
    _, info = m.solve(a, b)
    assert not info.certified, info
    assert info.certified == (info.backward_error <= 64.0 * np.finfo(float).eps)
    assert info.gmres_iters == 120


@pytest.mark.parametrize("cap", [0, "half"])
def test_gathered_cmod_matches_oracle(cap):
    """The gathered-tile cmod (every level, or the levels above half the
    largest level's product blocks) gives the oracle's factor."""
    a, B = matgen.config3(m=24, nrhs=4)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    if cap == "half":
        # the largest level needs this many product elements in the two-phase cmod
        probe = R.LdltSymbolic.analyze_with(a, st)
        cap = probe.cmod_elems_max() // 2
        assert cap > 0
    st = st.with_cmod_cap(cap)
    sym, fg, fo, s = factor_pair(a, st)
    assert_same_factor(fg.factors(), fo)
    check_solution(a, fg, fo, s, B)


def test_factor_from_caller_symbolic_equals_analysis_factor():
    """rsl_plan_create with the permutation, partition and tree of an
    analysis (as the reference's factor_numeric takes its symbolic) drives
    the device factor to the same bits, and to the oracle's factor."""
    a, B = matgen.config3(m=20, nrhs=3)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym = R.LdltSymbolic.analyze_with(a, st)
    e = sym.export()
    sym2 = R.LdltSymbolic.from_symbolic(a, e.perm, e.sn_first, e.sn_ncol, e.sn_parent(), st)
    f1, f2 = sym.factor(a), sym2.factor(a)
    g1, g2 = f1.factors(), f2.factors()
    for k in ("l_col_ptr", "l_row_idx", "l_values", "d_diag", "d_subdiag", "two_by_two", "perm"):
        assert np.array_equal(getattr(g1, k), getattr(g2, k)), k
    assert np.array_equal(f1.solve_many(B), f2.solve_many(B))
    s = O.onepass_scale(a)
    assert_same_factor(g2, O.factor(sym2.export(), O.scaled(a, s)))


@pytest.mark.parametrize("cplx", [False, True])
def test_factor_file_round_trip(tmp_path, cplx):
    """Factor once, solve elsewhere: a factor written to a file and read into
    a handle of a second, independent analysis of the same pattern solves to
    the original's bits without a numeric factorization; the host LdltFactors
    arrays round-trip through .npz; a file for another plan is rejected."""
    if cplx:
        a, B = matgen.config2(m=12, nrhs=2)
    else:
        a, B = matgen.config3(m=14, nrhs=2)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    f = R.LdltSymbolic.analyze_with(a, st).factor(a)
    X = f.solve_many(B)
    path = tmp_path / "factor.rslf"
    f.save(path)
    sym2 = R.LdltSymbolic.analyze_with(a, st)   # as another process would
    g = sym2.load_factor(path)
    assert g.inertia() == f.inertia() and g.n_perturbed() == f.n_perturbed()
    assert np.array_equal(g.solve_many(B), X)
    fa = f.factors()
    fa.save(tmp_path / "factors.npz")
    fb = R.LdltFactors.load(tmp_path / "factors.npz")
    for k in ("l_col_ptr", "l_row_idx", "l_values", "d_diag", "d_subdiag", "two_by_two", "perm"):
        assert np.array_equal(getattr(fa, k), getattr(fb, k)), k
    assert fb.inertia == fa.inertia
    other = R.LdltSymbolic.analyze_with(a, st.with_ordering(R.OrderingMethod.Amd))
    with pytest.raises(R.InvalidInput, match="another symbolic"):
        other.load_factor(path)
    # the loaded handle refactors and refines like the original
    if not cplx:
        a2 = a.with_values(a.values * 0.9 + 0.01)
        g.refactor(a2)
        f.refactor(a2)
        assert np.array_equal(g.solve_many(B), f.solve_many(B))


def test_rank_deficiency_names_supernode_and_column():
    a = R.CscMatrix.from_triplets(3, [0, 1, 2], [0, 1, 2], [2.0, 0.0, 3.0])
    with pytest.raises(R.NumericallyRankDeficient) as ei:
        R.LdltSolver.factor_with(a)
    msg = str(ei.value)
    m = re.search(r"zero pivot at column (\d+) of supernode (\d+) \((\d+) columns, elimination index (\d+)\)", msg)
    assert m, msg
    col, sn, w, e = map(int, m.groups())
    assert col < w and e >= col


def test_solve_device_is_ordered_with_caller_stream_and_later_calls():
    """A device solve enqueued on a caller's stream is ordered after the
    caller's earlier work and before its later work, and a refactor of the
    same handle issued right after it (no synchronization in between) cannot
    change what it reads."""
    hip = C.CDLL("libamdhip64.so.7")
    a = matgen.grid_laplacian((40, 40), shift=-2.5)
    a2 = a.with_values(a.values * 1.25 - 0.1)
    f = R.LdltSolver.factor_with(a)
    B = np.ascontiguousarray(rhs(a, 4, 81))
    X = f.solve_many(B)
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    db, dx = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(db), C.c_size_t(B.nbytes)) == 0
    assert hip.hipMalloc(C.byref(dx), C.c_size_t(B.nbytes)) == 0
    Xd = np.zeros_like(B)
    try:
        assert hip.hipMemsetAsync(dx, 0, C.c_size_t(B.nbytes), stream) == 0
        assert hip.hipMemcpyAsync(db, B.ctypes.data_as(C.c_void_p), C.c_size_t(B.nbytes), 1, stream) == 0
        f.solve_device(db.value, 4, dx.value, stream.value)
        f.refactor(a2)  # same handle, no synchronization in between
        assert hip.hipMemcpyAsync(Xd.ctypes.data_as(C.c_void_p), dx, C.c_size_t(B.nbytes), 2, stream) == 0
        assert hip.hipStreamSynchronize(stream) == 0
    finally:
        hip.hipFree(db)
        hip.hipFree(dx)
        hip.hipStreamDestroy(stream)
    assert np.array_equal(Xd, X)
    # and the handle now holds the refactored matrix
    X2 = f.solve_many(B)
    assert np.abs(a2.symv(X2) - B).max() < 1e-9 * np.abs(B).max() * max(1.0, np.abs(X2).max())
