"""Device factor and solve (librslab_amd.so through the C ABI) against the CPU
oracle on the same symbolic analysis and the same seeded inputs.

Bar, on matrices whose pivot sequence is stable under 1e-15 relative
perturbations of A (random dense fronts, the config-shaped cases): identical
pivot sequence (perm, 2x2 flags), inertia and perturbation count; L and D
equal to 1e-9 relative (FP64 with a different summation order inside the GEMM
tiles, which the reference itself pins only to solution tolerance,
gemm_tuning.rs:116-174); solutions equal to 1e-9 relative; backward error
within 10x of the oracle's.

On constant-coefficient grids the Bunch-Kaufman comparisons meet exact
magnitude ties, and which candidate wins is decided by the last bit of the
Schur updates (the oracle's own pivot sequence changes under a 1e-15
perturbation; tests/test_oracle.py shows it). There the bar is: inertia equal
to the oracle's, P^T A P = L D L^T to rounding for the device factor itself,
and backward error within 10x of the oracle's.
"""
import dataclasses

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import rslab_amd as R
from rslab_amd import matgen

pytestmark = pytest.mark.gpu


def grid2d(m, diag, off):
    a = matgen.grid_laplacian((m, m))
    cols = np.repeat(np.arange(a.n), np.diff(a.col_ptr))
    dt = np.complex128 if isinstance(diag, complex) or isinstance(off, complex) else np.float64
    return a.with_values(np.where(a.row_idx == cols, diag, off).astype(dt))


def dense_lower_csc(M):
    n = M.shape[0]
    r, c = np.tril_indices(n)
    o = np.lexsort((r, c))
    r, c = r[o], c[o]
    cp = np.concatenate([[0], np.cumsum(np.bincount(c, minlength=n))])
    return R.CscMatrix(n, cp, r, M[r, c])


def csc(f):
    return sp.csc_matrix((f.l_values, f.l_row_idx, f.l_col_ptr), shape=(f.n, f.n))


def assert_same_factor(fg, fo, tol=1e-9):
    """Pivot sequence and inertia exact; L and D to `tol` relative."""
    assert np.array_equal(fg.perm, fo.perm), "pivot permutation differs"
    assert np.array_equal(fg.two_by_two, fo.two_by_two), "1x1/2x2 block structure differs"
    assert fg.inertia.as_tuple() == fo.inertia
    assert fg.n_perturbed == fo.n_perturbed
    lg, lo = csc(fg), csc(fo)
    scale_l = max(1.0, abs(lo).max()) if lo.nnz else 1.0
    assert abs(lg - lo).max() <= tol * scale_l if lo.nnz else True
    sd = max(1.0, np.abs(fo.d_diag).max())
    assert np.abs(fg.d_diag - fo.d_diag).max() <= tol * sd
    assert np.abs(fg.d_subdiag - fo.d_subdiag).max() <= tol * sd


def assert_valid_factor(f, a_scaled, tol=1e-12):
    """P^T A_hat P = L D L^T, measured against |L| |D| |L|^T entrywise."""
    n = f.n
    L = csc(f)
    d = sp.diags(f.d_diag) + sp.diags(f.d_subdiag[:-1] * f.two_by_two[:-1], -1) + \
        sp.diags(f.d_subdiag[:-1] * f.two_by_two[:-1], 1)
    cols = np.repeat(np.arange(n), np.diff(a_scaled.col_ptr))
    A = sp.csc_matrix((a_scaled.values, (a_scaled.row_idx, cols)), shape=(n, n))
    A = A + sp.triu(A.T, 1)
    Ap = A[f.perm][:, f.perm]
    R_ = L @ d @ L.T - Ap
    M = abs(L) @ abs(d) @ abs(L).T
    assert abs(R_).max() <= tol * max(M.max(), 1.0)


def factor_pair(a, settings=None, panel_nb=64, pivot=0, abs_floor=0.0):
    st = settings or R.SolverSettings()
    sym = R.LdltSymbolic.analyze_with(a, st)
    fg = sym.factor(a)
    e = sym.export()
    s = O.onepass_scale(a)
    fo = O.factor(e, O.scaled(a, s), panel_nb=panel_nb, pivot=pivot, abs_floor=abs_floor)
    return sym, fg, fo, s


def rhs(a, k, seed):
    cols = [matgen.rng_uniform(seed + c, -1, 1, a.n) for c in range(k)]
    if a.is_complex:
        cols = [c + 1j * matgen.rng_uniform(seed + 100 + i, -1, 1, a.n) for i, c in enumerate(cols)]
    return np.stack(cols, 1)


def check_solution(a, fg, fo, s, B):
    X = fg.solve_many(B)
    Xo = fo.solve_many(B * s[:, None]) * s[:, None]
    assert np.abs(X - Xo).max() <= 1e-9 * max(np.abs(Xo).max(), 1.0)
    bg = O.backward_error(a, X, B)
    bo = O.backward_error(a, Xo, B)
    assert np.all(bg <= 10 * np.maximum(bo, 1e-16)), (bg, bo)
    return X


# ------------------------------------------------------------ known answers
def test_known_answer_complex_2x2():
    # ldlt_generic.rs:896-915: [[0,1],[1,0]] over C
    a = R.CscMatrix.from_triplets(2, [0, 1, 1], [0, 0, 1], np.array([0, 1, 0], dtype=complex))
    f = R.LdltSolver.factor_with(a)
    x = f.solve(np.array([1 + 1j, 2 - 1j]))
    assert np.abs(x - np.array([2 - 1j, 1 + 1j])).max() < 1e-14


def test_diagonal_inertia_and_solve():
    # sparse_solver.rs:1006-1021
    d = np.array([2.0, -3.0, 4.0, -1.0, 5.0])
    a = R.CscMatrix.from_triplets(5, range(5), range(5), d)
    f = R.LdltSolver.factor_with(a)
    assert f.inertia().as_tuple() == (3, 2, 0)
    assert np.allclose(f.solve(d), np.ones(5), atol=1e-15)


def test_one_by_one():
    a = R.CscMatrix.from_triplets(1, [0], [0], [-4.0])
    f = R.LdltSolver.factor_with(a)
    assert f.inertia().as_tuple() == (0, 1, 0)
    assert abs(f.solve(np.array([2.0]))[0] + 0.5) < 1e-15


# ------------------------------------------------------------ error paths
def test_singular_is_rank_deficient():
    # ldlt_generic.rs:975-983
    a = R.CscMatrix.from_triplets(2, [0, 1], [0, 1], [0.0, 0.0])
    with pytest.raises(R.NumericallyRankDeficient):
        R.LdltSolver.factor_with(a)


def test_failed_refactor_blocks_solves_until_refactored():
    a = R.CscMatrix.from_triplets(2, [0, 1], [0, 1], [2.0, 3.0])
    z = a.with_values(np.array([0.0, 0.0]))
    f = R.LdltSolver.factor_with(a)
    with pytest.raises(R.NumericallyRankDeficient):
        f.refactor(z)
    with pytest.raises(R.InvalidInput):
        f.solve(np.ones(2))
    f.refactor(a)
    assert np.allclose(f.solve(np.array([2.0, 3.0])), [1.0, 1.0], atol=1e-15)


def test_perturb_to_eps_rescues_singular_complex():
    # multifrontal_ldlt.rs:4186-4221
    a = R.CscMatrix.from_triplets(3, [0, 2, 1], [0, 0, 1], np.array([2 + 1j, -1 + 0.3j, 0j]))
    with pytest.raises(R.NumericallyRankDeficient):
        R.LdltSolver.factor_with(a)
    st = R.SolverSettings().with_pivot(R.ZeroPivotAction.PerturbToEps(1e-8))
    sym, fg, fo, s = factor_pair(a, st, pivot=1, abs_floor=1e-8)
    assert fg.n_perturbed() >= 1
    assert_same_factor(fg.factors(), fo)
    assert np.all(np.isfinite(fg.solve(np.ones(3, dtype=complex))))


def test_dimension_mismatch():
    a = grid2d(6, 4.0, -1.0)
    f = R.LdltSolver.factor_with(a)
    with pytest.raises(R.DimensionMismatch):
        f.solve(np.ones(a.n + 1))
    with pytest.raises(R.DimensionMismatch):
        f.solve_many(np.ones(a.n * 3), 2)


# ------------------------------------------------------------ reference LL matrices
CASES = {
    "indef10": lambda: grid2d(10, 0.5, -1.0),
    "cplx9": lambda: grid2d(9, 0.4 + 0.3j, -1 + 0.1j),
    "spd12": lambda: grid2d(12, 8.0, -1.0),
    "indef60": lambda: grid2d(60, 0.5, -1.0),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_ll_matrices_match_oracle(case):
    """multifrontal_ldlt.rs:3866-4013: residual <= 1e-9, inertia equal to the
    oracle's, a valid factor (tie-laden constant grids, see module docstring)."""
    a = CASES[case]()
    sym, fg, fo, s = factor_pair(a)
    g = fg.factors()
    assert g.inertia.as_tuple() == fo.inertia
    assert_valid_factor(g, O.scaled(a, s))
    b = np.array([(i % 7) - 3.0 for i in range(a.n)], dtype=a.values.dtype)
    x = fg.solve(b)
    assert np.abs(b - a.symv(x)).max() < 1e-9
    xo = fo.solve_many(b * s) * s
    assert O.backward_error(a, x, b)[0] <= 10 * max(O.backward_error(a, xo, b)[0], 1e-16)


@pytest.mark.parametrize("panel_nb", [16, 32, 64, 128])
@pytest.mark.parametrize("wide", [False, True])
def test_dense_front_panels_and_wide_path(panel_nb, wide):
    """One dense indefinite front (n=300): several nb-panels, deep-row replay
    and deferred Schur; `wide` routes it through the multi-workgroup path."""
    rng = np.random.default_rng(7)
    M = rng.standard_normal((300, 300))
    M = M + M.T
    a = dense_lower_csc(M)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.Natural).with_panel_nb(panel_nb)
    st = dataclasses.replace(st, wide_flops=1.0 if wide else 1e30)
    sym, fg, fo, s = factor_pair(a, st, panel_nb=panel_nb)
    assert_same_factor(fg.factors(), fo)
    ev = np.linalg.eigvalsh(M)
    assert fg.inertia().as_tuple() == (int((ev > 0).sum()), int((ev < 0).sum()), 0)
    check_solution(a, fg, fo, s, rhs(a, 3, 21))


def test_dense_complex_front_multi_panel():
    # multifrontal_ldlt.rs:4040-4066 shape, n = 200 over C
    n = 200
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    Mc = (((np.maximum(i, j) + 3 * np.minimum(i, j)) % 5) - 2.0) + 0.2j
    Mc[np.arange(n), np.arange(n)] = complex(0.5, 1.0)
    a = dense_lower_csc(Mc)
    st = dataclasses.replace(R.SolverSettings(), wide_flops=1.0)
    sym, fg, fo, s = factor_pair(a, st)
    assert_same_factor(fg.factors(), fo)
    check_solution(a, fg, fo, s, rhs(a, 2, 31))


# ------------------------------------------------------------ multi-RHS, refactor, determinism
def test_solve_many_equals_columnwise():
    # ldlt_generic.rs:785-832
    a = grid2d(30, 0.5, -1.0)
    f = R.LdltSolver.factor_with(a)
    B = rhs(a, 5, 41)
    X = f.solve_many(B)
    for c in range(5):
        assert np.abs(X[:, c] - f.solve(B[:, c])).max() <= 1e-13 * np.abs(X).max()


def test_refactor_equals_fresh_factor_and_is_deterministic():
    a1 = matgen.grid_laplacian((20, 20), shift=-2.5)
    a2 = matgen.grid_laplacian((20, 20), shift=-1.1, diag_real=matgen.rng_uniform(5, -0.1, 0.1, 400))
    sym = R.LdltSymbolic.analyze(a1)
    f = sym.factor(a1)
    f.refactor(a2)
    g = sym.factor(a2)
    ff, gf = f.factors(), g.factors()
    assert np.array_equal(ff.l_values, gf.l_values) and np.array_equal(ff.d_diag, gf.d_diag)
    assert f.inertia() == g.inertia()
    b = rhs(a2, 2, 51)
    assert np.array_equal(f.solve_many(b), g.solve_many(b))


def test_solve_device_matches_host():
    """rsl_solve_device on device buffers (allocated through the HIP runtime
    directly) equals the host-buffer solve bit for bit."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    a = grid2d(25, 0.5, -1.0)
    f = R.LdltSolver.factor_with(a)
    B = np.ascontiguousarray(rhs(a, 4, 61))
    X = f.solve_many(B)
    nbytes = B.nbytes
    db, dx = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(db), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dx), C.c_size_t(nbytes)) == 0
    try:
        assert hip.hipMemcpy(db, B.ctypes.data_as(C.c_void_p), C.c_size_t(nbytes), 1) == 0
        f.solve_device(db.value, 4, dx.value)
        assert hip.hipDeviceSynchronize() == 0
        Xd = np.empty_like(B)
        assert hip.hipMemcpy(Xd.ctypes.data_as(C.c_void_p), dx, C.c_size_t(nbytes), 2) == 0
        assert np.array_equal(Xd, X)
    finally:
        hip.hipFree(db)
        hip.hipFree(dx)


def _device_copy(hip, arr):
    import ctypes as C
    d = C.c_void_p()
    assert hip.hipMalloc(C.byref(d), C.c_size_t(max(arr.nbytes, 8))) == 0
    assert hip.hipMemcpy(d, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0
    return d


@pytest.mark.parametrize("pivot", ["fail", "force_accept"])
def test_refactor_device_matches_host_refactor(pivot):
    """rsl_refactor_device on values already in device memory gives the host
    call's factor bit for bit: the device equilibration and ForceAccept floor
    reproduce the host's (maxima are order-free, the guarded 1/sqrt is
    correctly rounded)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    a, _ = matgen.config3(m=20, nrhs=1)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    if pivot == "force_accept":
        st = st.with_pivot(R.ZeroPivotAction.ForceAccept())
    sym = R.LdltSymbolic.analyze_with(a, st)
    f = sym.factor(a, st)
    ref = f.factors()
    ref_in, ref_np = f.inertia(), f.n_perturbed()
    # overwrite the factor with other values, then refactor from device values
    f.refactor(a.with_values(a.values * 1.5 - 0.25))
    vals = np.ascontiguousarray(a.values_f64())
    dv = _device_copy(hip, vals)
    try:
        f.refactor_device(dv.value)
    finally:
        hip.hipFree(dv)
    got = f.factors()
    for field in ("l_col_ptr", "l_row_idx", "l_values", "d_diag", "d_subdiag", "two_by_two", "perm"):
        assert np.array_equal(getattr(got, field), getattr(ref, field)), field
    assert f.inertia() == ref_in and f.n_perturbed() == ref_np


@pytest.mark.parametrize("prec", [R.Precision.FP64, R.Precision.FP32_SCHUR])
def test_refine_device_matches_host_refine(prec):
    """rsl_refine_device with A, b and x in device memory takes the host
    call's steps and returns its x bit for bit (||A||_inf is summed on the
    device in the host loop's order)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    a, B = matgen.config3(m=20, nrhs=3)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND).with_precision(prec)
    f = R.LdltSymbolic.analyze_with(a, st).factor(a, st)
    mixed = R.MixedLdltSolver(f, a)
    X, infos = mixed.solve_to(a, B, 1e-14)
    B = np.ascontiguousarray(B)
    vals = np.ascontiguousarray(a.values_f64())
    da, db, dx = _device_copy(hip, vals), _device_copy(hip, B), _device_copy(hip, np.zeros_like(B))
    try:
        dinfos = mixed.solve_to_device(da.value, db.value, 3, dx.value, 1e-14)
        Xd = np.empty_like(B)
        assert hip.hipMemcpy(Xd.ctypes.data_as(C.c_void_p), dx, C.c_size_t(B.nbytes), 2) == 0
    finally:
        for d in (da, db, dx):
            hip.hipFree(d)
    assert np.array_equal(Xd, X)
    assert dinfos == infos and all(i.certified for i in infos)


# ------------------------------------------------------------ config-shaped cases
def test_config1_matches_oracle():
    """100x100 2D Laplacian, A = L - 2I, AMD, one RHS b = A x_true.

    After equilibration every diagonal is 1 and every coupling -1/2, so a node
    whose four neighbours were eliminated as 1x1 pivots is left with
    1 - 4/4 = 0 up to rounding; with pivoting bounded to the supernode panel
    (no delayed pivots in the left-looking schedule) such pivots are taken as
    near-zero 1x1s whose sign and size are set by rounding: the oracle's own
    inertia moves (8125 -> 8161 -> 8158 positives) under 1e-15 relative
    perturbations of A, and the closed form is (8163, 1837, 0). What is
    pinned: a valid factor of the same symbolic and a backward error of the
    same (poor) order as the oracle's; iterative refinement against A is what
    recovers the solution."""
    a, b, x_true = matgen.config1()
    sym, fg, fo, s = factor_pair(a)
    g = fg.factors()
    assert_valid_factor(g, O.scaled(a, s), tol=1e-10)
    x = fg.solve(b)
    xo = fo.solve_many(b * s) * s
    bg, bo = O.backward_error(a, x, b)[0], O.backward_error(a, xo, b)[0]
    assert bg <= 10 * bo and bo <= 10 * bg
    # Sylvester: only the rounding-level pivots' signs are open, so the closed
    # form (8163, 1837, 0) lies inside [neg, neg + ntiny] for the device factor
    from test_oracle import rounding_level_bracket
    neg, ntiny = rounding_level_bracket(g)
    assert ntiny > 0 and neg <= 1837 <= neg + ntiny, (neg, ntiny, g.inertia)
    # a clean gap between the rounding-level pivots (the device's summation
    # order leaves them a few ulps of 1 rather than the oracle's exact 2^-54)
    # and the O(1e-2) ones
    d = np.abs(g.d_diag[~g.two_by_two])
    assert np.all((d <= 1e-12) | (d >= 1e-3)), np.sort(d[(d > 1e-12) & (d < 1e-3)])[:8]
    # the ladder cannot repair such a factor, on the device as in the oracle:
    # uncertified after the full FGMRES budget, the flag consistent with the
    # backward error; forward error against x_true reported
    xr, info = R.MixedLdltSolver(fg, a).solve_to(a, b, 1e-14)
    _, ir_o, gm_o, be_o, cert_o = O.refine_ladder(a, fo, s, b, 1e-14)
    assert not info.certified and not cert_o
    assert info.gmres_iters == gm_o == 120
    assert info.backward_error == pytest.approx(O.backward_error(a, xr, b)[0], rel=1e-9)
    print(f"config 1: device inertia {g.inertia.as_tuple()} (oracle {fo.inertia}, closed form "
          f"(8163, 1837, 0), {ntiny} rounding-level pivots); ladder berr {info.backward_error:.2e} "
          f"(oracle {be_o:.2e}), forward error {np.abs(xr - x_true).max() / np.abs(x_true).max():.2e}")


@pytest.mark.parametrize("prec", [R.Precision.FP64, R.Precision.FP32_SCHUR, R.Precision.BF16_SCHUR])
def test_ladder_matches_oracle_ladder_config3_shape(prec):
    """The device's refinement ladder against the oracle's restatement
    (refine_ladder, mixed.rs:140-221; FGMRES, iterative.rs:822-1021) on the
    oracle's factor with the same reduced-precision Schur products (operands
    rounded to FP32 / BF16, FP32 sums): same inertia, the same IR and FGMRES
    step counts and certificate per column, backward errors both at the
    target, solutions equal to what the conditioning allows, and a forward
    error against the FP64 oracle solution."""
    a, B = matgen.config3(m=24, nrhs=4)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND).with_precision(prec)
    sym = R.LdltSymbolic.analyze_with(a, st)
    fg = sym.factor(a, st)
    e = sym.export()
    s = O.onepass_scale(a)
    fo = O.factor(e, O.scaled(a, s), prec=int(prec))
    f64 = O.factor(e, O.scaled(a, s))
    if prec == R.Precision.BF16_SCHUR:
        # BF16 products (8-bit significands) perturb the factored matrix by
        # ~1e-3 relative: eigenvalues inside the +-0.01 band of the shift can
        # change sign (the emulation's own inertia is one off the FP64 one),
        # and the device's FP32 sums run in another order than the oracle's
        assert abs(fg.inertia().negative - fo.inertia[1]) <= 2
        assert abs(fg.inertia().negative - f64.inertia[1]) <= 2
    else:
        assert fg.inertia().as_tuple() == fo.inertia == f64.inertia
    X, infos = R.MixedLdltSolver(fg, a).solve_to(a, B, 1e-14)
    Xref = f64.solve_many(B * s[:, None]) * s[:, None]
    for c, info in enumerate(infos):
        xo, ir, gm, be, cert = O.refine_ladder(a, fo, s, B[:, c], 1e-14)
        assert info.certified == cert, (c, info, ir, gm, be)
        gm_slack = 2 if prec != R.Precision.BF16_SCHUR else max(4, gm // 4)
        assert abs(info.ir_iters - ir) <= 1 and abs(info.gmres_iters - gm) <= gm_slack, (c, info, ir, gm)
        assert info.backward_error <= 1e-14 and be <= 1e-14
        assert np.abs(X[:, c] - xo).max() <= 1e-9 * np.abs(xo).max()
        fwd = np.abs(X[:, c] - Xref[:, c]).max() / np.abs(Xref[:, c]).max()
        assert fwd <= 1e-9, fwd


def test_config3_shape_nd_matches_oracle():
    """L - 1.5 I + diag(delta) on 24^3 with nested dissection."""
    a, B = matgen.config3(m=24, nrhs=8)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym, fg, fo, s = factor_pair(a, st)
    assert_same_factor(fg.factors(), fo)
    check_solution(a, fg, fo, s, B)
    m = 24
    t = 2 - 2 * np.cos(np.arange(1, m + 1) * np.pi / (m + 1))
    lam = (t[:, None, None] + t[None, :, None] + t[None, None, :]).ravel() - 1.5
    assert int((lam < -0.01).sum()) <= fg.inertia().negative <= int((lam < 0.01).sum())


def test_config2_shape_complex_nd_matches_oracle():
    """Complex-symmetric Helmholtz with absorbing layer on 16^3, ND, 4 RHS."""
    a, B = matgen.config2(m=16, nrhs=4)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym, fg, fo, s = factor_pair(a, st)
    assert_same_factor(fg.factors(), fo)
    check_solution(a, fg, fo, s, B)


def test_config3_full_size_matches_oracle_fixture():
    """216,000 unknowns (60^3): the oracle's factor of the same matrix
    (tests/golden/config3_60_oracle.json, written by make_config3_oracle.py,
    about 90 s on one core) fixes the inertia, which the device must equal
    exactly (inside the Weyl bracket [7171, 7358]); the device's plain solve
    equals the oracle's at 64 sampled rows of every column to 1e-9 and its
    backward error is within 10x of the oracle's."""
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                     "config3_60_oracle.json")))
    a, B = matgen.config3(m=60, nrhs=8)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    f = R.LdltSymbolic.analyze_with(a, st).factor(a)
    assert f.inertia().as_tuple() == tuple(fx["inertia"])
    assert 7171 <= f.inertia().negative <= 7358
    assert f.n_perturbed() == fx["n_perturbed"]
    X = f.solve_many(B)
    xs = np.array(fx["x_rows"])
    scale = np.array(fx["x_col_inf"])[None, :]
    assert np.abs(X[fx["rows"]] - xs).max() <= 1e-9 * scale.max()
    be = O.backward_error(a, X, B)
    assert np.all(be <= 10 * np.maximum(np.array(fx["backward_error"]), 1e-16)), (be, fx["backward_error"])


# ------------------------------------------------------------ refinement ladder, reduced-precision Schur
def lap2d_ref(k):
    """mixed.rs:389-412: 2D 5-point Laplacian, diagonal 4, couplings -1."""
    return matgen.grid_laplacian((k, k))


@pytest.mark.parametrize("prec", [R.Precision.FP64, R.Precision.FP32_SCHUR, R.Precision.BF16_SCHUR])
def test_mixed_ldlt_certifies_on_laplacian(prec):
    # mixed.rs:415-439: certified, and equal to the FP64 direct solve to 1e-10
    a = lap2d_ref(40)
    b = np.array([(i % 13) - 6.0 for i in range(a.n)])
    m = R.MixedLdltSolver.factor_with(a, R.SolverSettings().with_precision(prec))
    x, info = m.solve(a, b)
    assert info.certified, info
    xd = R.LdltSolver.factor_with(a).solve(b)
    assert np.abs(x - xd).max() <= 1e-10 * max(np.abs(xd).max(), 1.0)


@pytest.mark.parametrize("prec", [R.Precision.FP32_SCHUR, R.Precision.BF16_SCHUR])
def test_reduced_precision_schur_refines_config3_shape(prec):
    """Schur / cmod products in FP32 (or BF16 inputs, FP32 accumulation),
    pivots and D in FP64: the ladder reaches backward error <= 1e-14 on every
    column against the FP64 matrix, and the solution equals the FP64 one to
    what the conditioning allows."""
    a, B = matgen.config3(m=24, nrhs=8)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym = R.LdltSymbolic.analyze_with(a, st)
    lo = sym.factor(a, st.with_precision(prec))
    X, infos = R.MixedLdltSolver(lo, a).solve_to(a, B, 1e-14)
    assert all(i.certified and i.backward_error <= 1e-14 for i in infos), infos
    hi = sym.factor(a)
    Xd, infos_d = R.MixedLdltSolver(hi, a).solve_to(a, B, 1e-14)
    assert all(i.certified for i in infos_d)
    assert np.abs(X - Xd).max() <= 1e-8 * np.abs(Xd).max()
    # the reduced product really is reduced: the raw factors differ
    fl, fh = lo.factors(), hi.factors()
    if np.array_equal(fl.perm, fh.perm):
        d = np.abs(csc(fl) - csc(fh)).max()
        assert d > 0.0


def test_fp64_ladder_certifies_config3_shape_and_matches_oracle_solution():
    a, B = matgen.config3(m=24, nrhs=4)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym, fg, fo, s = factor_pair(a, st)
    X, infos = R.MixedLdltSolver(fg, a).solve_to(a, B, 1e-14)
    assert all(i.certified for i in infos)
    Xo = fo.solve_many(B * s[:, None]) * s[:, None]
    assert np.abs(X - Xo).max() <= 1e-9 * np.abs(Xo).max()
    assert O.backward_error(a, X, B).max() <= 1e-14


# ------------------------------------------------------------ batched independent systems (config 4 shape)
def test_batch_factor_equals_per_system_factors():
    """A batch of uneven shifted 2D Laplacians in one launch sequence: each
    system's inertia equals its own factor's and the closed form
    #{lambda_pq < sigma}; solutions equal the per-system solves."""
    systems = matgen.config4(count=24)
    mats = [a for a, _, _ in systems]
    big = R.CscMatrix.block_diag(mats)
    b = np.concatenate([bb for _, bb, _ in systems])
    sym = R.LdltSymbolic.analyze_batch(mats)
    f = sym.factor(big)
    inert = f.inertia_batch(len(mats))
    x = f.solve(b)
    off = 0
    for (a, bb, sg), ib in zip(systems, inert):
        m = int(round(np.sqrt(a.n)))
        t = 2 - 2 * np.cos(np.arange(1, m + 1) * np.pi / (m + 1))
        lam = (t[:, None] + t[None, :]).ravel() - sg
        fi = R.LdltSolver.factor_with(a)
        assert ib.as_tuple() == fi.inertia().as_tuple()
        if np.abs(lam).min() > 1e-8:
            assert ib.negative == int((lam < 0).sum())
        xi = fi.solve(bb)
        # the batch plan and a system's own plan split the solve between level
        # launches and the dependency-driven launch at different levels, so the
        # two solves sum in different groupings: equal to the conditioning
        # (the shifts sit next to eigenvalues), and equally backward stable
        xs = x[off:off + a.n]
        assert np.abs(xs - xi).max() <= 1e-10 * max(np.abs(xi).max(), 1.0)
        assert O.backward_error(a, xs, bb)[0] <= 10 * max(O.backward_error(a, xi, bb)[0], 1e-15)
        off += a.n
    assert sum(i.total() for i in inert) == big.n


@pytest.mark.parametrize("prec", [R.Precision.FP64, R.Precision.FP32_SCHUR])
def test_batch_refines_to_fp64(prec):
    systems = matgen.config4(count=16)
    mats = [a for a, _, _ in systems]
    big = R.CscMatrix.block_diag(mats)
    b = np.concatenate([bb for _, bb, _ in systems])
    sym = R.LdltSymbolic.analyze_batch(mats)
    f = sym.factor(big, sym.settings.with_precision(prec))
    x, info = R.MixedLdltSolver(f, big).solve_to(big, b, 1e-14)
    assert info.certified
    off = 0
    for a, bb, _ in systems:
        xi = x[off:off + a.n]
        assert O.backward_error(a, xi, bb)[0] <= 1e-13
        off += a.n


# ------------------------------------------------------------ same-pattern refactor sweep (config 5 shape)
def test_refactor_sweep_vector_stencil():
    """3 dof per node, 27-point coupling, shifted by -omega^2 I for several
    omega^2 with the analysis and device storage reused: every refactor
    equals a fresh factor of that matrix (inertia = the oracle's), and the
    FP32-Schur factor refined reaches backward error <= 1e-12 (32 RHS)."""
    om = matgen.config5_omegas(20)[::6]
    base = matgen.vector_stencil27(m=8, omega2=om[0])
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym = R.LdltSymbolic.analyze_with(base, st)
    e = sym.export()
    f64 = sym.factor(base)
    f32 = sym.factor(base, st.with_precision(R.Precision.FP32_SCHUR))
    for k, w2 in enumerate(om):
        a = matgen.vector_stencil27(m=8, omega2=w2)
        f64.refactor(a)
        s = O.onepass_scale(a)
        fo = O.factor(e, O.scaled(a, s))
        assert f64.inertia().as_tuple() == fo.inertia
        B = matgen.config5_rhs(a.n, k, nrhs=32)
        X = f64.solve_many(B)
        Xo = fo.solve_many(B * s[:, None]) * s[:, None]
        assert np.abs(X - Xo).max() <= 1e-9 * np.abs(Xo).max()
        f32.refactor(a)
        X32, infos = R.MixedLdltSolver(f32, a).solve_to(a, B, 1e-12)
        assert all(i.certified for i in infos), infos
        assert O.backward_error(a, X32, B).max() <= 1e-12
