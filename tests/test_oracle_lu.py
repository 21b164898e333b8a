"""Pins of the LU twin's CPU oracle (oracle/lu_oracle_body.h), independent of
its own formulas:

* reconstruction: L U equals the equilibrated, permuted matrix
  (D_r A D_c)[perm_row][:, perm] to rounding, for real and complex,
  orderings AMD and ND, thresholds 0.1 and 1 (a dropped update, a wrong
  index or a missed interchange breaks it);
* the pivot rule: with threshold 1 on a dense front (one supernode) the row
  order is LAPACK getrf's partial pivoting (scipy.linalg.lu_factor) and L, U
  equal its factors; with 0.1 the diagonal is kept exactly when
  |a_kk| >= 0.1 max |a_ik| (2x2 closed cases);
* the reference's own known-answer tests for the LU (multifrontal_lu.rs
  :3336-3925): residual bars, solve_many against single solves, the u knob,
  static pivots reused across a value sweep, incomplete LU with refinement;
* the ZeroPivotAction policies on an exactly singular pivot.
"""
import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O
import rslab_amd as R
from rslab_amd import matgen
from rslab_amd.lu import GeneralCsc


def analysis(a: GeneralCsc, ordering=None):
    """Export of the analysis of the symmetrized pattern (what LuSymbolic
    analyses, multifrontal_lu.rs:850-869)."""
    cp, ri = O.symmetrized_lower(a.n, a.col_ptr, a.row_idx)
    pat = R.CscMatrix(a.n, cp, ri, np.ones(ri.size))
    st = R.SolverSettings()
    if ordering is not None:
        st = st.with_ordering(ordering)
    return R.LdltSymbolic.analyze_with(pat, st).export()


def scaled_permuted(a: GeneralCsc, f):
    M = a.to_dense() * f.d_row[:, None] * f.d_col[None, :]
    return M[np.ix_(f.perm_row, f.perm)]


def resid(a, x, b):
    return float(np.abs(a.matvec(x) - b).max())


def grid2d(m, diag, right, left, down, up):
    """The reference tests' 2D grids: diagonal `diag`, (p, p+1) `right`,
    (p+1, p) `left`, (p, p+m) `down`, (p+m, p) `up`, p = a*m + b."""
    rr, cc, vv = [], [], []
    for a in range(m):
        for b in range(m):
            p = a * m + b
            rr.append(p); cc.append(p); vv.append(diag)
            if b + 1 < m:
                q = p + 1
                rr += [p, q]; cc += [q, p]; vv += [right, left]
            if a + 1 < m:
                q = p + m
                rr += [p, q]; cc += [q, p]; vv += [down, up]
    return GeneralCsc.from_triplets(m * m, rr, cc, np.array(vv))


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("ordering", [R.OrderingMethod.Amd, R.OrderingMethod.MetisND])
@pytest.mark.parametrize("u", [0.1, 1.0])
def test_lu_reconstructs_equilibrated_permuted_matrix(cplx, ordering, u):
    a, _ = matgen.lu_workload(m=6, nrhs=1, cplx=cplx)
    f = O.lu_factor(analysis(a, ordering), a, pivot_u=u)
    L, U = f.dense_lu()
    assert np.allclose(np.diag(L), 1.0) and np.abs(np.triu(L, 1)).max() == 0 and np.abs(np.tril(U, -1)).max() == 0
    M = scaled_permuted(a, f)
    assert np.abs(L @ U - M).max() <= 1e-13 * np.abs(M).max()
    # both permutations are permutations; columns follow the fill-reducing order
    assert sorted(f.perm) == list(range(a.n)) and sorted(f.perm_row) == list(range(a.n))
    if u == 1.0:
        assert (f.perm_row != f.perm).any()  # partial pivoting moved rows


def test_lu_oracle_detects_a_planted_mistake():
    """The reconstruction check fails when one L entry is perturbed."""
    a, _ = matgen.lu_workload(m=5, nrhs=1)
    f = O.lu_factor(analysis(a), a, pivot_u=1.0)
    L, U = f.dense_lu()
    M = scaled_permuted(a, f)
    i = int(np.argmax(np.abs(np.tril(L, -1)).sum(axis=1)))
    L[i, np.argmax(np.abs(np.tril(L, -1)[i]))] *= 1.0 + 1e-9
    assert np.abs(L @ U - M).max() > 1e-13 * np.abs(M).max()


@pytest.mark.parametrize("seed", [1, 2])
def test_full_pivoting_threshold_equals_lapack_on_a_dense_front(seed):
    n = 60
    rng = np.random.default_rng(seed)
    Md = rng.standard_normal((n, n))
    a = GeneralCsc.from_dense(Md)
    sym = analysis(a, R.OrderingMethod.Natural)
    assert len(sym.sn_first) == 1  # one dense front
    f = O.lu_factor(sym, a, pivot_u=1.0)
    As = Md * f.d_row[:, None] * f.d_col[None, :]
    lu, piv = sla.lu_factor(As)
    order = np.arange(n)
    for i, p in enumerate(piv):
        order[[i, p]] = order[[p, i]]
    assert np.array_equal(f.perm, np.arange(n))
    assert np.array_equal(f.perm_row, order)
    L, U = f.dense_lu()
    assert np.abs(L - (np.tril(lu, -1) + np.eye(n))).max() <= 1e-12
    assert np.abs(U - np.triu(lu)).max() <= 1e-12 * np.abs(U).max()


@pytest.mark.parametrize("a11,swap", [(0.05, True), (0.0999, True), (0.1001, False), (0.2, False)])
def test_threshold_rule_on_2x2(a11, swap):
    """[[a11, 1], [1, 1]] (already equilibrated: every row and column max is
    1): u = 0.1 interchanges the rows iff |a11| < 0.1; u = 0 never does."""
    a = GeneralCsc.from_dense(np.array([[a11, 1.0], [1.0, 1.0]]))
    sym = analysis(a, R.OrderingMethod.Natural)
    f = O.lu_factor(sym, a, pivot_u=0.1)
    assert np.array_equal(f.d_row, [1.0, 1.0]) and np.array_equal(f.d_col, [1.0, 1.0])
    assert list(f.perm_row) == ([1, 0] if swap else [0, 1])
    g = O.lu_factor(sym, a, pivot_u=0.0)
    assert list(g.perm_row) == [0, 1]


def test_zero_pivot_actions():
    """[[0, 1], [1, 0]] with static pivots: Fail reports rank deficiency,
    PerturbToEps lifts the zero to the floor (perturb_pivot), a threshold
    interchange avoids it altogether."""
    a = GeneralCsc.from_dense(np.array([[0.0, 1.0], [1.0, 0.0]]))
    sym = analysis(a, R.OrderingMethod.Natural)
    with pytest.raises(O.RankDeficient):
        O.lu_factor(sym, a, pivot_u=0.0)
    f = O.lu_factor(sym, a, pivot=1, abs_floor=1e-3, pivot_u=0.0)
    assert f.n_perturbed >= 1
    L, U = f.dense_lu()
    assert U[0, 0] == 1e-3
    g = O.lu_factor(sym, a, pivot_u=0.1)
    assert g.n_perturbed == 0 and list(g.perm_row) == [1, 0]
    h = O.lu_factor(sym, a, pivot=2, pivot_u=0.0)  # ForceAccept: max(|A|, 1) eps
    assert h.n_perturbed >= 1 and np.diag(h.dense_lu()[1])[0] == np.finfo(float).eps


# ---- the reference's known-answer tests (multifrontal_lu.rs:3336-3925)
def test_ref_f64_unsymmetric_tridiag():
    n = 20
    rr, cc, vv = [], [], []
    for i in range(n):
        rr.append(i); cc.append(i); vv.append(4.0)
        if i + 1 < n:
            rr += [i + 1, i]; cc += [i, i + 1]; vv += [-1.0, -2.0]
    a = GeneralCsc.from_triplets(n, rr, cc, vv)
    b = np.arange(n) - 9.5
    x = O.lu_factor(analysis(a), a).solve_many(b)
    assert resid(a, x, b) < 1e-10


def test_ref_solve_many_matches_single():
    n = 12
    rr, cc, vv = [], [], []
    for i in range(n):
        rr.append(i); cc.append(i); vv.append(5.0)
        if i + 1 < n:
            rr += [i + 1, i]; cc += [i, i + 1]; vv += [-1.0, -2.0]
    a = GeneralCsc.from_triplets(n, rr, cc, vv)
    f = O.lu_factor(analysis(a), a)
    B = (np.arange(n * 5) % 7 - 3.0).reshape(n, 5)
    X = f.solve_many(B)
    for c in range(5):
        assert np.abs(X[:, c] - f.solve_many(B[:, c])).max() < 1e-10


@pytest.mark.parametrize("diag", [0.3 + 0.05j, 0.05 + 0.01j])
def test_ref_pivoting_small_diagonal(diag):
    a = grid2d(6, diag, 2.0 + 0.3j, 1.5 - 0.2j, 1.8 + 0.1j, 2.2 + 0.4j)
    b = np.array([(i % 5) - 2.0 + 1.0j for i in range(a.n)])
    f = O.lu_factor(analysis(a), a)
    if abs(diag) < 0.1:
        assert (f.perm_row != f.perm).any()  # the tiny diagonal is interchanged
    assert resid(a, f.solve_many(b), b) < 1e-9


@pytest.mark.parametrize("u", [0.0, 0.1, 0.5, 1.0])
def test_ref_pivot_u_knob(u):
    a = grid2d(7, 12.0 + 1.0j, -1.0 + 0.2j, -1.3 - 0.1j, -1.1 + 0.3j, -0.9 + 0.15j)
    b = np.array([(i % 5) - 2.0 + 1.0j for i in range(a.n)])
    f = O.lu_factor(analysis(a), a, pivot_u=u)
    assert resid(a, f.solve_many(b), b) < 1e-9


def refined(f, a, b, max_iter):
    """solve_lu_refined (multifrontal_lu.rs:3285-3322): keep the best iterate."""
    x = f.solve_many(b)
    best, best_res = x.copy(), np.inf
    for it in range(max_iter + 1):
        r = b - a.matvec(x)
        res = np.abs(r).max()
        if res < best_res:
            best_res, best = res, x.copy()
        if res == 0.0 or it == max_iter:
            break
        x = x + f.solve_many(r)
    return best


def test_ref_static_pivot_reuse_across_value_sweep():
    m = 8
    t = grid2d(m, 1.0, 1.0, 1.0, 1.0, 1.0)
    sym = analysis(t)
    b = np.array([i - 4.0 + 0.7j for i in range(t.n)])
    cols = np.repeat(np.arange(t.n), np.diff(t.col_ptr))
    for shift in (0.0, 1.5, -0.8, 3.0):
        vals = np.where(t.row_idx == cols, 9.0 + shift + 1.0j, -1.0 + 0.2j)
        a = t.with_values(vals)
        f = O.lu_factor(sym, a, pivot_u=0.0)
        assert resid(a, refined(f, a, b, 2), b) < 1e-9


def test_ref_complex_unsymmetric_2d_grid():
    a = grid2d(8, 8.0 + 1.0j, -1.0 + 0.2j, -2.0 + 0.1j, -1.5 + 0.3j, -0.5 + 0.4j)
    b = np.array([(i % 5) - 2.0 + 1.0j for i in range(a.n)])
    assert resid(a, O.lu_factor(analysis(a), a).solve_many(b), b) < 1e-9


def test_ref_incomplete_lu_reduces_fill_and_still_solves():
    a = grid2d(14, 8.0 + 1.0j, -1.0 + 0.2j, -2.0 + 0.1j, -1.5 + 0.3j, -0.5 + 0.4j)
    b = np.array([(i % 5) - 2.0 + 1.0j for i in range(a.n)])
    sym = analysis(a)
    full = O.lu_factor(sym, a)
    inc = O.lu_factor(sym, a, drop_tol=5e-2)
    assert inc.factor_nnz() < full.factor_nnz()
    assert resid(a, refined(inc, a, b, 10), b) < 1e-6


def test_ref_phased_analyze_once_factor_many():
    m = 7
    t = grid2d(m, 1.0, 1.0, 1.0, 1.0, 1.0)
    sym = analysis(t)
    b = np.array([i - 4.0 + 1.0j for i in range(t.n)])
    cols = np.repeat(np.arange(t.n), np.diff(t.col_ptr))
    for shift in (0.0, 2.5, -1.0):
        a = t.with_values(np.where(t.row_idx == cols, 8.0 + shift + 1.0j, -1.0 + 0.2j))
        xp = O.lu_factor(sym, a).solve_many(b)
        xo = O.lu_factor(analysis(a), a).solve_many(b)
        assert np.abs(xp - xo).max() < 1e-10
        assert resid(a, xp, b) < 1e-8


def test_symmetrized_lower_pattern():
    """A + A^T lower pattern against a dense construction."""
    rng = np.random.default_rng(5)
    M = (rng.random((30, 30)) < 0.1) * 1.0 + np.eye(30)
    a = GeneralCsc.from_dense(M)
    cp, ri = O.symmetrized_lower(a.n, a.col_ptr, a.row_idx)
    S = np.tril((M + M.T) != 0)
    c, r = np.nonzero(S.T)
    assert np.array_equal(ri, r) and np.array_equal(np.diff(cp), np.bincount(c, minlength=30))
