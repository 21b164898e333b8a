"""Device LU twin (supernodal left-looking LU with threshold partial pivoting,
src/numeric/multifrontal_lu.rs) against the CPU oracle (oracle/lu_oracle.c)
and the reference's own LU tests.

Bar: the pivot sequence (perm_row) and the fill-reducing order (perm) equal
the oracle's exactly, the equilibration (d_row, d_col) equals it bit for bit,
L and U equal it to 1e-9 relative (the device sums each update's products in
another order than the oracle's scalar loops), solutions to 1e-9 relative,
with backward error within 10x of the oracle's."""
import dataclasses

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

import oracle as O
import rslab_amd as R
from rslab_amd import matgen
from rslab_amd.lu import GeneralCsc

pytestmark = pytest.mark.gpu

ND = R.OrderingMethod.MetisND


def lu_pair(a, st=None):
    st = st or R.SolverSettings()
    sym = R.LuSymbolic.analyze_with(a, st)
    pivot = st.on_zero_pivot.kind
    fo = O.lu_factor(sym.export(), a, pivot=pivot, abs_floor=st.on_zero_pivot.abs_floor, pivot_u=st.pivot_u,
                     drop_tol=st.drop_tol)
    fd = sym.factor(a, st)
    return sym, fd, fo


def sparse_of(n, ptr, idx, val, csr):
    m = (sp.csr_matrix if csr else sp.csc_matrix)((val, idx, ptr), shape=(n, n))
    return m


def assert_same_lu(g, fo, tol=1e-9):
    assert np.array_equal(g.perm, fo.perm)
    assert np.array_equal(g.perm_row, fo.perm_row)
    assert np.array_equal(g.d_row, fo.d_row) and np.array_equal(g.d_col, fo.d_col)
    n = g.n
    Ld = sparse_of(n, g.l_col_ptr, g.l_row_idx, g.l_values, False)
    Lo = sparse_of(n, fo.l_col_ptr, fo.l_row_idx, fo.l_values, False)
    Ud = sparse_of(n, g.u_row_ptr, g.u_col_idx, g.u_values, True)
    Uo = sparse_of(n, fo.u_row_ptr, fo.u_col_idx, fo.u_values, True)
    assert abs(Ld - Lo).max() <= tol * max(abs(Lo).max(), 1.0)
    assert abs(Ud - Uo).max() <= tol * max(abs(Uo).max(), 1.0)
    # each L column starts with its unit diagonal, each U row with its pivot
    e = np.arange(n)
    assert np.array_equal(g.l_row_idx[g.l_col_ptr[:-1]], e) and np.all(g.l_values[g.l_col_ptr[:-1]] == 1)
    assert np.array_equal(g.u_col_idx[g.u_row_ptr[:-1]], e)


def check_solve(a, fd, fo, B, tol=1e-9):
    X = fd.solve_many(B)
    Xo = fo.solve_many(B)
    assert np.abs(X - Xo).max() <= tol * max(np.abs(Xo).max(), 1e-300)
    bd, bo = O.lu_backward_error(a, X, B), O.lu_backward_error(a, Xo, B)
    assert (bd <= 10 * np.maximum(bo, 1e-16)).all(), (bd, bo)
    return X


def rhs(n, k, seed, cplx=False):
    rng = np.random.default_rng(seed)
    B = rng.uniform(-1, 1, (n, k))
    return B + 1j * rng.uniform(-1, 1, (n, k)) if cplx else B


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("ordering", [R.OrderingMethod.Amd, ND])
@pytest.mark.parametrize("u", [0.0, 0.1, 1.0])
def test_lu_matches_oracle_on_convection_diffusion(cplx, ordering, u):
    a, B = matgen.lu_workload(m=10, nrhs=3, cplx=cplx)
    st = R.SolverSettings().with_ordering(ordering).with_pivot_u(u)
    sym, fd, fo = lu_pair(a, st)
    assert sym.info()["nsuper"] > 10
    assert_same_lu(fd.factors(), fo)
    check_solve(a, fd, fo, B)


def test_lu_larger_grid_exercises_wide_panels():
    """16^3 (n = 4096): the top separators span several panels, and
    pivoting with u = 1 moves rows in most panels."""
    a, B = matgen.lu_workload(m=16, nrhs=4)
    st = R.SolverSettings().with_ordering(ND).with_pivot_u(1.0)
    sym, fd, fo = lu_pair(a, st)
    g = fd.factors()
    assert (g.perm_row != g.perm).any()
    assert_same_lu(g, fo)
    check_solve(a, fd, fo, B)


@pytest.mark.parametrize("n", [200, 600])
def test_dense_front_pivots_equal_lapack(n):
    """One dense supernode (Natural ordering): with u = 1 the row order is
    LAPACK getrf's partial pivoting (the device's 16-column panels, the
    oracle's 32 or 128, LAPACK's own blocking: the same pivots)."""
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    a = GeneralCsc.from_dense(M)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.Natural).with_pivot_u(1.0)
    sym = R.LuSymbolic.analyze_with(a, st)
    assert sym.info()["nsuper"] == 1
    g = sym.factor(a, st).factors()
    lu, piv = sla.lu_factor(M * g.d_row[:, None] * g.d_col[None, :])
    order = np.arange(n)
    for i, p in enumerate(piv):
        order[[i, p]] = order[[p, i]]
    assert np.array_equal(g.perm_row, order)
    L, U = g.dense_lu()
    assert np.abs(L - (np.tril(lu, -1) + np.eye(n))).max() <= 1e-10
    assert np.abs(U - np.triu(lu)).max() <= 1e-10 * np.abs(U).max()


def test_lu_zero_pivot_policies():
    a = GeneralCsc.from_dense(np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 2.0]]))
    st = R.SolverSettings().with_ordering(R.OrderingMethod.Natural).with_pivot_u(0.0)
    sym = R.LuSymbolic.analyze_with(a, st)
    with pytest.raises(R.NumericallyRankDeficient, match="zero pivot"):
        sym.factor(a, st)
    f = sym.factor(a, st.with_pivot(R.ZeroPivotAction.PerturbToEps(1e-3)))
    assert f.n_perturbed() >= 1
    g = sym.factor(a, st.with_pivot_u(0.1))
    fo = O.lu_factor(sym.export(), a, pivot_u=0.1)
    gf = g.factors()
    assert g.n_perturbed() == 0 and np.array_equal(gf.perm_row, fo.perm_row)
    assert (gf.perm_row != gf.perm).sum() == 2  # the zero diagonal's rows interchanged


def test_lu_drop_tol_matches_oracle_and_refines():
    """Incomplete LU (drop_tol at emit): the same dropped pattern as the oracle's."""
    a, B = matgen.lu_workload(m=10, nrhs=1, cplx=True)
    st = R.SolverSettings().with_ordering(ND).with_drop_tol(5e-2)
    sym, fd, fo = lu_pair(a, st)
    g = fd.factors()
    full = sym.factor(a, R.SolverSettings().with_ordering(ND))
    assert g.factor_nnz() < full.factor_nnz()
    assert np.array_equal(g.l_col_ptr, fo.l_col_ptr) and np.array_equal(g.u_row_ptr, fo.u_row_ptr)
    assert_same_lu(g, fo)


def device_solve(f, B):
    """rsl_lu_solve_device on buffers allocated through the HIP runtime."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    B = np.ascontiguousarray(B)
    db, dx = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(db), C.c_size_t(B.nbytes)) == 0
    assert hip.hipMalloc(C.byref(dx), C.c_size_t(B.nbytes)) == 0
    try:
        assert hip.hipMemcpy(db, B.ctypes.data_as(C.c_void_p), C.c_size_t(B.nbytes), 1) == 0
        f.solve_device(db.value, B.shape[1], dx.value)
        assert hip.hipDeviceSynchronize() == 0
        X = np.empty_like(B)
        assert hip.hipMemcpy(X.ctypes.data_as(C.c_void_p), dx, C.c_size_t(B.nbytes), 2) == 0
        return X
    finally:
        hip.hipFree(db)
        hip.hipFree(dx)


def test_lu_device_solve_and_refactor():
    a, B = matgen.lu_workload(m=12, nrhs=5)
    st = R.SolverSettings().with_ordering(ND)
    sym = R.LuSymbolic.analyze_with(a, st)
    f = sym.factor(a, st)
    X = f.solve_many(B)
    assert np.array_equal(X, device_solve(f, B))
    # refactor with new values on the same pattern equals a fresh factor
    a2 = a.with_values(a.values * (1.0 + 0.1 * np.cos(np.arange(a.nnz))))
    f.refactor(a2)
    g = sym.factor(a2, st)
    assert np.array_equal(f.solve_many(B), g.solve_many(B))
    fo = O.lu_factor(sym.export(), a2)
    check_solve(a2, f, fo, B)


# ---- the reference's LU tests (multifrontal_lu.rs:3336-3925), on the device
def grid2d(m, diag, right, left, down, up):
    rr, cc, vv = [], [], []
    for i in range(m):
        for j in range(m):
            p = i * m + j
            rr.append(p); cc.append(p); vv.append(diag)
            if j + 1 < m:
                rr += [p, p + 1]; cc += [p + 1, p]; vv += [right, left]
            if i + 1 < m:
                rr += [p, p + m]; cc += [p + m, p]; vv += [down, up]
    return GeneralCsc.from_triplets(m * m, rr, cc, np.array(vv))


def resid(a, x, b):
    return float(np.abs(a.matvec(x) - b).max())


def test_ref_tridiag_and_solve_many():
    n = 20
    rr, cc, vv = [], [], []
    for i in range(n):
        rr.append(i); cc.append(i); vv.append(4.0)
        if i + 1 < n:
            rr += [i + 1, i]; cc += [i, i + 1]; vv += [-1.0, -2.0]
    a = GeneralCsc.from_triplets(n, rr, cc, vv)
    b = np.arange(n) - 9.5
    f = R.LuSolver.factor(a)
    assert resid(a, f.solve(b), b) < 1e-10
    B = (np.arange(n * 5) % 7 - 3.0).reshape(n, 5)
    X = f.solve_many(B)
    for c in range(5):
        assert np.abs(X[:, c] - f.solve(B[:, c])).max() < 1e-10


@pytest.mark.parametrize("diag", [0.3 + 0.05j, 0.05 + 0.01j])
def test_ref_pivoting_small_diagonal(diag):
    a = grid2d(6, diag, 2.0 + 0.3j, 1.5 - 0.2j, 1.8 + 0.1j, 2.2 + 0.4j)
    b = np.array([(i % 5) - 2.0 + 1.0j for i in range(a.n)])
    x = R.LuSolver.factor(a).solve(b)
    assert resid(a, x, b) < 1e-9


@pytest.mark.parametrize("u", [0.0, 0.1, 0.5, 1.0])
def test_ref_pivot_u_knob(u):
    a = grid2d(7, 12.0 + 1.0j, -1.0 + 0.2j, -1.3 - 0.1j, -1.1 + 0.3j, -0.9 + 0.15j)
    b = np.array([(i % 5) - 2.0 + 1.0j for i in range(a.n)])
    x = R.LuSolver.factor(a, R.SolverSettings().with_pivot_u(u)).solve(b)
    assert resid(a, x, b) < 1e-9
    assert R.SolverSettings().with_pivot_u(5.0).pivot_u == 1.0
    assert R.SolverSettings().with_pivot_u(-2.0).pivot_u == 0.0


def test_ref_static_pivot_reuse_across_value_sweep():
    t = grid2d(8, 1.0, 1.0, 1.0, 1.0, 1.0)
    sym = R.LuSymbolic.analyze(t.with_values(t.values.astype(complex)))
    b = np.array([i - 4.0 + 0.7j for i in range(t.n)])
    cols = np.repeat(np.arange(t.n), np.diff(t.col_ptr))
    st = R.SolverSettings().with_pivot_u(0.0)
    for shift in (0.0, 1.5, -0.8, 3.0):
        a = t.with_values(np.where(t.row_idx == cols, 9.0 + shift + 1.0j, -1.0 + 0.2j))
        f = sym.factor(a, st)
        assert resid(a, f.solve_refined(a, b, 2), b) < 1e-9


def test_ref_complex_grid_and_incomplete_lu():
    a = grid2d(14, 8.0 + 1.0j, -1.0 + 0.2j, -2.0 + 0.1j, -1.5 + 0.3j, -0.5 + 0.4j)
    b = np.array([(i % 5) - 2.0 + 1.0j for i in range(a.n)])
    full = R.LuSolver.factor(a)
    assert resid(a, full.solve(b), b) < 1e-9
    inc = R.LuSolver.factor(a, dataclasses.replace(R.SolverSettings(), drop_tol=5e-2))
    assert inc.factor_nnz() < full.factor_nnz()
    assert resid(a, inc.solve_refined(a, b, 10), b) < 1e-6


def test_ref_phased_analyze_once_factor_many():
    t = grid2d(7, 1.0, 1.0, 1.0, 1.0, 1.0)
    sym = R.LuSymbolic.analyze(t)
    b = np.array([i - 4.0 + 1.0j for i in range(t.n)])
    cols = np.repeat(np.arange(t.n), np.diff(t.col_ptr))
    for shift in (0.0, 2.5, -1.0):
        a = t.with_values(np.where(t.row_idx == cols, 8.0 + shift + 1.0j, -1.0 + 0.2j))
        xp = sym.factor(a).solve(b)
        xo = R.LuSolver.factor(a).solve(b)
        assert np.abs(xp - xo).max() < 1e-10
        assert resid(a, xp, b) < 1e-8


def test_lu_config3_grid_full_size_refines():
    """60^3 convection-diffusion (n = 216,000, the LU twin's workload): the
    device factor with a few refinement steps reaches backward error 1e-14
    on 8 right-hand sides, and the device solve equals the host-buffer one."""
    a, B = matgen.lu_workload(m=60, nrhs=8)
    st = R.SolverSettings().with_ordering(ND)
    f = R.LuSymbolic.analyze_with(a, st).factor(a, st)
    X = f.solve_refined(a, B, 3)
    assert O.lu_backward_error(a, X, B).max() <= 1e-14
    assert np.array_equal(device_solve(f, B), f.solve_many(B))
