"""Pins for the CPU oracle (oracle/ldlt_oracle.c), independent of the oracle's
own formulas: the reference's known answers, dense LAPACK Bunch-Kaufman
(scipy.linalg.ldl, the ?sytrf control flow the reference follows for one
front), eigenvalue inertia, closed forms, and planted mistakes."""
import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O
import rslab_amd as R
from rslab_amd import matgen


def analyze(a, **kw):
    st = R.SolverSettings()
    for k, v in kw.items():
        st = getattr(st, "with_" + k)(v)
    return R.LdltSymbolic.analyze_with(a, st).export()


def dense(a):
    cols = np.repeat(np.arange(a.n), np.diff(a.col_ptr))
    M = np.zeros((a.n, a.n), dtype=a.values.dtype)
    M[a.row_idx, cols] = a.values
    M[cols, a.row_idx] = a.values
    return M


def dense_lower_csc(M):
    n = M.shape[0]
    r, c = np.tril_indices(n)
    o = np.lexsort((r, c))
    r, c = r[o], c[o]
    cp = np.concatenate([[0], np.cumsum(np.bincount(c, minlength=n))])
    return R.CscMatrix(n, cp, r, M[r, c])


def grid2d(m, diag, off):
    a = matgen.grid_laplacian((m, m))
    cols = np.repeat(np.arange(a.n), np.diff(a.col_ptr))
    dt = np.complex128 if isinstance(diag, complex) or isinstance(off, complex) else np.float64
    return a.with_values(np.where(a.row_idx == cols, diag, off).astype(dt))


def reconstruct(f):
    n = f.n
    L = np.zeros((n, n), dtype=f.l_values.dtype)
    for j in range(n):
        for k in range(f.l_col_ptr[j], f.l_col_ptr[j + 1]):
            L[f.l_row_idx[k], j] = f.l_values[k]
    D = np.diag(f.d_diag).astype(f.l_values.dtype)
    for k in np.nonzero(f.two_by_two)[0]:
        D[k + 1, k] = D[k, k + 1] = f.d_subdiag[k]
    return L @ D @ L.T


def test_known_answer_2x2_real_and_complex():
    # ldlt_generic.rs:836-850 and :896-915
    a = R.CscMatrix.from_triplets(2, [0, 1, 1], [0, 0, 1], [0.0, 1.0, 0.0])
    x, f = O.solve(a, np.array([3.0, 5.0]), analyze(a))
    assert np.allclose(x, [5.0, 3.0], atol=1e-12)
    assert f.inertia == (1, 1, 0) and f.two_by_two.any()
    ac = a.with_values(np.array([0, 1, 0], dtype=np.complex128))
    x, f = O.solve(ac, np.array([1 + 1j, 2 - 1j]), analyze(ac))
    assert np.abs(x - np.array([2 - 1j, 1 + 1j])).max() < 1e-12


def test_diagonal_inertia_and_singular_rejected():
    # sparse_solver.rs:1006-1021 and ldlt_generic.rs:975-983
    d = [2.0, -3.0, 4.0, -1.0, 5.0]
    a = R.CscMatrix.from_triplets(5, range(5), range(5), d)
    _, f = O.solve(a, np.ones(5), analyze(a))
    assert f.inertia == (3, 2, 0)
    z = R.CscMatrix.from_triplets(2, [0, 1], [0, 1], [0.0, 0.0])
    with pytest.raises(O.RankDeficient):
        O.solve(z, np.ones(2), analyze(z))


def test_perturb_rescues_singular_complex():
    # multifrontal_ldlt.rs:4186-4221
    a = R.CscMatrix.from_triplets(3, [0, 2, 1], [0, 0, 1], np.array([2 + 1j, -1 + 0.3j, 0j]))
    e = analyze(a)
    with pytest.raises(O.RankDeficient):
        O.solve(a, np.ones(3), e)
    x, f = O.solve(a, np.ones(3, dtype=complex), e, pivot=1, abs_floor=1e-8)
    assert f.n_perturbed >= 1 and np.all(np.isfinite(x))


@pytest.mark.parametrize("n,seed", [(40, 0), (64, 1), (150, 2)])
def test_single_front_matches_lapack_bunch_kaufman(n, seed):
    """One dense front: within a single nb-panel the reference's getf2 is
    LAPACK ?sytf2 lower; over several panels it stays a Bunch-Kaufman LDL^T.
    Pinned by inertia (Sylvester vs eigenvalues), the 2x2 count of LAPACK for
    n <= nb, and the reconstruction P^T A P = L D L^T."""
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    M = M + M.T
    a = dense_lower_csc(M)
    e = analyze(a, ordering=R.OrderingMethod.Natural)
    assert len(e.sn_first) == 1
    f = O.factor(e, a, panel_nb=64)
    ev = np.linalg.eigvalsh(M)
    assert f.inertia == (int((ev > 0).sum()), int((ev < 0).sum()), 0)
    Ae = M[np.ix_(f.perm, f.perm)]
    assert np.abs(reconstruct(f) - Ae).max() < 1e-10 * np.abs(M).max() * n
    if n <= 64:
        lu, d, perm = sla.ldl(M, lower=True)
        n2_lapack = int(np.count_nonzero(np.abs(np.diag(d, -1)) > 0))
        assert int(f.two_by_two.sum()) == n2_lapack


@pytest.mark.parametrize("case", ["indef10", "cplx9", "spd12", "indef60"])
def test_reference_ll_test_matrices(case):
    """multifrontal_ldlt.rs:3866-4013: the left-looking tests' matrices,
    residual bar 1e-9, inertia against eigenvalues (real cases)."""
    a = {"indef10": lambda: grid2d(10, 0.5, -1.0), "cplx9": lambda: grid2d(9, 0.4 + 0.3j, -1 + 0.1j),
         "spd12": lambda: grid2d(12, 8.0, -1.0), "indef60": lambda: grid2d(60, 0.5, -1.0)}[case]()
    n = a.n
    b = np.array([(i % 7) - 3.0 for i in range(n)], dtype=a.values.dtype)
    x, f = O.solve(a, b, analyze(a))
    assert np.abs(b - a.symv(x)).max() < 1e-9
    if not np.iscomplexobj(a.values):
        ev = np.linalg.eigvalsh(dense(a))
        assert f.inertia == (int((ev > 0).sum()), int((ev < 0).sum()), 0)
    if case.startswith("indef") or case == "cplx9":
        assert f.two_by_two.any()


def test_multi_panel_dense_fronts():
    # multifrontal_ldlt.rs:4015-4066
    n = 100
    M = np.array([[n + 1.0 if i == j else ((max(i, j) + 2 * min(i, j)) % 5) - 2.0 for j in range(n)] for i in range(n)])
    a = dense_lower_csc(M)
    b = np.array([(i % 7) - 3.0 for i in range(n)])
    x, _ = O.solve(a, b, analyze(a))
    assert np.abs(b - M @ x).max() < 1e-9
    n = 90
    Mc = np.array([[complex(n, 1.0) if i == j else complex(((max(i, j) + 3 * min(i, j)) % 5) - 2.0, 0.2)
                    for j in range(n)] for i in range(n)])
    ac = dense_lower_csc(Mc)
    bc = np.full(n, 1 + 0.5j)
    x, _ = O.solve(ac, bc, analyze(ac))
    assert np.abs(bc - Mc @ x).max() < 1e-9


def test_config3_like_inertia_within_weyl_bounds():
    """3D L - 1.5 I + diag(delta): the inertia is bracketed by the eigenvalues
    of L - 1.5 I shifted by +-0.01 (Weyl), computed in closed form."""
    m = 12
    delta = matgen.rng_uniform(3, -0.01, 0.01, m ** 3)
    a = matgen.grid_laplacian((m, m, m), shift=-1.5, diag_real=delta)
    x, f = O.solve(a, matgen.rng_uniform(4, -1, 1, a.n), analyze(a, ordering=R.OrderingMethod.MetisND))
    t = 2 - 2 * np.cos(np.arange(1, m + 1) * np.pi / (m + 1))
    lam = (t[:, None, None] + t[None, :, None] + t[None, None, :]).ravel() - 1.5
    lo, hi = int((lam < -0.01).sum()), int((lam < 0.01).sum())
    assert lo <= f.inertia[1] <= hi
    ev = np.linalg.eigvalsh(dense(a))
    assert f.inertia[1] == int((ev < 0).sum())


def test_config1_closed_form_inertia():
    """Config 1's closed form (+8163, -1837, 0) from the 2D Dirichlet spectrum."""
    p, nneg, z, mn = O.laplacian2d_inertia(100, 2.0)
    assert (p, nneg, z) == (8163, 1837, 0) and mn > 1e-3


def test_solve_many_equals_columnwise():
    # ldlt_generic.rs:785-832: multi-RHS equals per-column solves
    a = grid2d(20, 0.5, -1.0)
    e = analyze(a)
    B = np.stack([matgen.rng_uniform(9 + c, -1, 1, a.n) for c in range(5)], 1)
    X, f = O.solve(a, B, e)
    for c in range(5):
        xc, _ = O.solve(a, B[:, c], e)
        assert np.abs(X[:, c] - xc).max() < 1e-12


def test_onepass_scale_known_answer():
    # s_i = 1/sqrt(max_j |A_ij|), guarded to 1 for empty rows
    a = R.CscMatrix.from_triplets(3, [0, 1, 1], [0, 0, 1], [4.0, -9.0, 1.0])
    s = O.onepass_scale(a)
    assert np.allclose(s, [1 / 3.0, 1 / 3.0, 1.0])


def test_planted_mistakes_are_caught():
    """The checks used for the device path catch a dropped term, a wrong sign
    or a transposed operand when planted in an oracle factor."""
    a = grid2d(12, 0.5, -1.0)
    b = np.array([(i % 7) - 3.0 for i in range(a.n)])
    e = analyze(a)
    s = O.onepass_scale(a)
    ah = O.scaled(a, s)
    good = O.factor(e, ah)
    xg = good.solve_arrays(b * s) * s
    assert np.abs(xg - good.solve_many(b * s) * s).max() < 1e-12
    assert np.abs(b - a.symv(xg)).max() < 1e-9
    # dropped term in L
    bad = O.factor(e, ah)
    k = int(np.argmax(np.abs(bad.l_values) * (bad.l_values != 1.0)))
    bad.l_values[k] = 0.0
    x = bad.solve_arrays(b * s) * s
    assert np.abs(b - a.symv(x)).max() > 1e-6
    # wrong sign of a 1x1 pivot: residual and inertia both move
    bad2 = O.factor(e, ah)
    j = int(np.nonzero(~bad2.two_by_two & ~np.concatenate([[False], bad2.two_by_two[:-1]]))[0][3])
    bad2.d_diag[j] = -bad2.d_diag[j]
    x = bad2.solve_arrays(b * s) * s
    assert np.abs(b - a.symv(x)).max() > 1e-6
    # transposed 2x2 coupling (d21 applied as -d21)
    bad3 = O.factor(e, ah)
    k2 = int(np.nonzero(bad3.two_by_two)[0][0])
    bad3.d_subdiag[k2] = -bad3.d_subdiag[k2]
    x = bad3.solve_arrays(b * s) * s
    assert np.abs(b - a.symv(x)).max() > 1e-6


def perturbed(a, seed):
    rng = np.random.default_rng(seed)
    return a.with_values(a.values * (1 + 1e-15 * rng.standard_normal(a.nnz)))


def test_pivot_sequence_stability_decides_the_parity_bar():
    """On a constant-coefficient grid the Bunch-Kaufman magnitude comparisons
    meet exact ties, so a 1e-15 relative perturbation of A changes the pivot
    sequence (inertia unchanged): two correct implementations with different
    summation orders need not pick the same pivots. On the config-3 shape
    (random diagonal) the sequence is stable, so there the device path must
    reproduce it exactly."""
    a = grid2d(30, 0.5, -1.0)
    e = analyze(a)
    f0 = O.factor(e, O.scaled(a, O.onepass_scale(a)))
    a1 = perturbed(a, 0)
    f1 = O.factor(e, O.scaled(a1, O.onepass_scale(a1)))
    assert not np.array_equal(f0.perm, f1.perm) and f0.inertia == f1.inertia
    a3, _ = matgen.config3(m=16, nrhs=1)
    e3 = analyze(a3, ordering=R.OrderingMethod.MetisND)
    g0 = O.factor(e3, O.scaled(a3, O.onepass_scale(a3)))
    for seed in range(3):
        a4 = perturbed(a3, seed)
        g1 = O.factor(e3, O.scaled(a4, O.onepass_scale(a4)))
        assert np.array_equal(g0.perm, g1.perm) and np.array_equal(g0.two_by_two, g1.two_by_two)


def _random_sym(n, seed, cplx, density=0.05):
    """Random symmetric (complex-symmetric when cplx) lower CSC with a full diagonal."""
    rng = np.random.default_rng(seed)
    M = np.where(rng.random((n, n)) < density, rng.standard_normal((n, n)), 0.0)
    if cplx:
        M = M + 1j * np.where(M != 0, rng.standard_normal((n, n)), 0.0)
    M = np.tril(M, -1)
    M = M + M.T  # plain transpose: complex-symmetric, never Hermitian
    M[np.arange(n), np.arange(n)] = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)
    return dense_lower_csc(M), M


@pytest.mark.parametrize("cplx", [False, True])
def test_symv_and_norm_inf_equal_dense_products(cplx):
    """CscMatrix::symv (csc.rs:294-309) applies the stored lower triangle as
    the full symmetric matrix, and sym_norm_inf (mixed.rs:358-371) is its
    largest absolute row sum: both equal numpy on the dense matrix. Every
    backward-error bar of the device tests goes through these two."""
    a, M = _random_sym(157, 3 + cplx, cplx)
    rng = np.random.default_rng(9)
    x = rng.standard_normal(a.n) + (1j * rng.standard_normal(a.n) if cplx else 0)
    y = O.symv(a, x)
    ref = M @ x
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()
    nrm = O.sym_norm_inf(a)
    assert nrm == pytest.approx(np.abs(M).sum(1).max(), rel=1e-14)
    # the Python mirror (used by MixedLdltSolver) agrees with both
    assert np.abs(a.symv(x) - ref).max() <= 1e-13 * np.abs(ref).max()
    assert a.norm_inf() == pytest.approx(nrm, rel=1e-14)
    # planted mistakes the comparison must catch: the transposed half dropped
    # (lower-triangle product), a conjugated transpose (Hermitian reading),
    # and the row sum of the stored triangle only
    cols = np.repeat(np.arange(a.n), np.diff(a.col_ptr))
    Ml = np.zeros_like(M)
    Ml[a.row_idx, cols] = a.values
    assert np.abs(Ml @ x - ref).max() > 1e-3 * np.abs(ref).max()
    assert np.abs(np.abs(Ml).sum(1).max() - nrm) > 1e-6 * nrm or np.abs(np.abs(Ml).sum(0).max() - nrm) > 1e-6 * nrm
    if cplx:
        Mh = np.tril(M) + np.tril(M, -1).conj().T
        assert np.abs(Mh @ x - ref).max() > 1e-3 * np.abs(ref).max()


def test_backward_error_known_answer():
    """Normwise backward error |b - A x|inf / (|A|inf |x|inf + |b|inf)
    (mixed.rs:151-155) on a hand-computed 2x2 case: A = [[2, 1], [1, -3]],
    x = (1, 1), b = (3, -1): r = (0, 1), |A|inf = 4, so 1 / (4 + 3)."""
    a = R.CscMatrix.from_triplets(2, [0, 1, 1], [0, 0, 1], [2.0, 1.0, -3.0])
    be = O.backward_error(a, np.array([1.0, 1.0]), np.array([3.0, -1.0]))
    assert be[0] == pytest.approx(1.0 / 7.0, rel=1e-15)
    assert O.sym_norm_inf(a) == 4.0


# ------------------------------------------------------------ ladder, FGMRES, reduced-precision emulation
def _krylov_min_residual(A, b, k):
    """min over x in K_k(A, b) of ||b - A x||_2, by least squares on an
    orthonormal basis of the Krylov space (x0 = 0, no preconditioner)."""
    Q = np.zeros((b.size, k))
    v = b / np.linalg.norm(b)
    for j in range(k):
        for i in range(j):  # twice-orthogonalised power basis
            v = v - (Q[:, i] @ v) * Q[:, i]
        for i in range(j):
            v = v - (Q[:, i] @ v) * Q[:, i]
        Q[:, j] = v / np.linalg.norm(v)
        v = A @ Q[:, j]
    y = np.linalg.lstsq(A @ Q, b, rcond=None)[0]
    return np.linalg.norm(b - A @ Q @ y)


@pytest.mark.parametrize("k", [1, 4, 9])
def test_gmres_iterate_minimizes_residual_over_krylov_space(k):
    """GMRES's defining property: after k steps (no restart, x0 = 0, identity
    preconditioner) its iterate minimizes ||b - A x|| over K_k(A, b)."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((40, 40)) + 6 * np.eye(40)
    b = rng.standard_normal(40)
    x, its, _ = O.gmres(lambda v: A @ v, b, lambda v: v, 1e-300, max_iter=k, restart=30)
    assert its == k
    assert np.linalg.norm(b - A @ x) == pytest.approx(_krylov_min_residual(A, b, k), rel=1e-9)
    # an exact preconditioner solves in one iteration; a restart keeps the iterate
    Ai = np.linalg.inv(A)
    x1, its1, conv = O.gmres(lambda v: A @ v, b, lambda v: Ai @ v, 1e-12)
    assert conv and its1 == 1 and np.linalg.norm(b - A @ x1) <= 1e-12 * np.linalg.norm(b)
    x2, its2, _ = O.gmres(lambda v: A @ v, b, lambda v: v, 1e-300, max_iter=12, restart=4)
    assert its2 == 12 and np.linalg.norm(b - A @ x2) < np.linalg.norm(b)


def test_gmres_complex_symmetric_system():
    """The complex field: GMRES orthogonalises with the conjugated inner product
    (dotc) even for a complex-symmetric operator."""
    rng = np.random.default_rng(8)
    M = rng.standard_normal((30, 30)) + 1j * rng.standard_normal((30, 30))
    A = M + M.T + 10 * np.eye(30)
    b = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    x, its, conv = O.gmres(lambda v: A @ v, b, lambda v: v, 1e-12, max_iter=60, restart=30)
    assert conv and np.linalg.norm(b - A @ x) <= 1e-12 * np.linalg.norm(b) * 1.01
    x5, _, _ = O.gmres(lambda v: A @ v, b, lambda v: v, 1e-300, max_iter=5, restart=30)
    assert np.linalg.norm(b - A @ x5) == pytest.approx(_krylov_min_residual_c(A, b, 5), rel=1e-9)


def _krylov_min_residual_c(A, b, k):
    K = np.zeros((b.size, k), dtype=complex)
    v = b.copy()
    for j in range(k):
        K[:, j] = v
        v = A @ v
    Q, _ = np.linalg.qr(K)
    y = np.linalg.lstsq(A @ Q, b, rcond=None)[0]
    return np.linalg.norm(b - A @ Q @ y)


def _oracle_factor(a, prec=0, pivot=0, abs_floor=0.0, drop_tol=None, **kw):
    e = analyze(a, **kw)
    s = O.onepass_scale(a)
    return O.factor(e, O.scaled(a, s), pivot=pivot, abs_floor=abs_floor, drop_tol=drop_tol, prec=prec), s


def test_ladder_certifies_laplacian_and_matches_direct_solve():
    """mixed.rs:415-439 (lap2d(40)): certified, and equal to the FP64 direct
    solve to 1e-10; the FP64 factor certifies without any GMRES step."""
    a = matgen.grid_laplacian((40, 40))
    b = np.array([(i % 13) - 6.0 for i in range(a.n)])
    for prec in (O.PREC_FP64, O.PREC_FP32, O.PREC_BF16):
        fo, s = _oracle_factor(a, prec=prec)
        x, ir, gm, be, cert = O.refine_ladder(a, fo, s, b, O.CERT_FACTOR * np.finfo(float).eps)
        assert cert and be <= 64 * np.finfo(float).eps, (prec, ir, gm, be)
        f64, s64 = _oracle_factor(a)
        xd = f64.solve_many(b * s64) * s64
        assert np.abs(x - xd).max() <= 1e-10 * np.abs(xd).max()
        if prec == O.PREC_FP64:
            assert gm == 0 and ir <= 1


def test_ladder_reports_uncertified_for_crippled_factor():
    """mixed.rs:504-523: lap2d(90) with PerturbToEps(1e-12) and drop
    tolerance 0.9: IR stalls, FGMRES spends its 120 iterations, not certified,
    and the flag agrees with the backward error."""
    a = matgen.grid_laplacian((90, 90))
    b = np.array([(i % 5) - 2.0 for i in range(a.n)])
    fo, s = _oracle_factor(a, prec=O.PREC_FP32, pivot=1, abs_floor=1e-12, drop_tol=0.9)
    x, ir, gm, be, cert = O.refine_ladder(a, fo, s, b, 64 * np.finfo(float).eps)
    assert not cert and gm == 120
    assert cert == (be <= 64 * np.finfo(float).eps)
    assert be == pytest.approx(O.backward_error(a, x, b)[0], rel=1e-12)


def test_lowp_rounding_known_answers():
    """FP32 and BF16 operand rounding (round to nearest, ties to even)."""
    assert O.round_lowp(1.0 + 2.0 ** -30, 1) == 1.0
    assert O.round_lowp(0.1, 1) == float(np.float32(0.1))
    assert O.round_lowp(1.0 + 2.0 ** -8, 2) == 1.0               # tie -> even (1.0)
    assert O.round_lowp(1.0 + 2.0 ** -7 + 2.0 ** -8, 2) == 1.0 + 2.0 ** -6  # tie -> even
    assert O.round_lowp(1.0 + 3 * 2.0 ** -9, 2) == 1.0 + 2.0 ** -7
    assert O.round_lowp(-3.0, 2) == -3.0
    assert O.round_lowp(1.0 + 2.0 ** -30, 0) == float(np.float32(1.0 + 2.0 ** -30))


def test_lowp_emulation_reduces_only_the_schur_products():
    """One dense front narrower than a panel has no Schur products: the
    reduced-precision factor equals the FP64 one bit for bit. On a sparse
    multi-level tree the products are reduced: with FP32 the pivot sequence
    is the FP64 one and D moves by a few FP32 roundoffs, with BF16 (8-bit
    significands) some Bunch-Kaufman choices change; the inertia holds."""
    rng = np.random.default_rng(3)
    M = rng.standard_normal((48, 48))
    a1 = dense_lower_csc(M + M.T)
    f_hi, _ = _oracle_factor(a1, ordering=R.OrderingMethod.Natural)
    for prec in (O.PREC_FP32, O.PREC_BF16):
        f_lo, _ = _oracle_factor(a1, prec=prec, ordering=R.OrderingMethod.Natural)
        assert np.array_equal(f_lo.l_values, f_hi.l_values) and np.array_equal(f_lo.d_diag, f_hi.d_diag)
    a, _ = matgen.config3(m=12, nrhs=1)
    f_hi, _ = _oracle_factor(a, ordering=R.OrderingMethod.MetisND)
    f32, _ = _oracle_factor(a, prec=O.PREC_FP32, ordering=R.OrderingMethod.MetisND)
    assert f32.inertia == f_hi.inertia and np.array_equal(f32.perm, f_hi.perm)
    rel = np.abs(f32.d_diag - f_hi.d_diag).max() / np.abs(f_hi.d_diag).max()
    assert 1e-9 < rel < 1e-4, rel
    f16, _ = _oracle_factor(a, prec=O.PREC_BF16, ordering=R.OrderingMethod.MetisND)
    assert f16.inertia == f_hi.inertia
    assert not np.array_equal(f16.d_diag, f_hi.d_diag)


def rounding_level_bracket(f, tiny=1e-10):
    """(negatives among the pivots above `tiny`, number of 1x1 pivots at or
    below it): Sylvester's law pins the closed-form negative count inside
    [neg, neg + ntiny] when only the rounding-level pivots' signs are open."""
    two = np.asarray(f.two_by_two, dtype=bool)
    first = np.nonzero(two)[0]
    one = np.ones(f.n, dtype=bool)
    one[first] = False
    one[first + 1] = False
    d = np.real(f.d_diag)
    dd = d[first] * d[first + 1] - np.real(f.d_subdiag[first]) ** 2
    tr = d[first] + d[first + 1]
    neg2 = int((dd < 0).sum() + 2 * ((dd > 0) & (tr < 0)).sum() + ((dd == 0) & (tr < 0)).sum())
    big = one & (np.abs(d) > tiny)
    return int((d[big] < 0).sum()) + neg2, int((one & (np.abs(d) <= tiny)).sum())


def test_config1_inertia_is_fixed_up_to_rounding_level_pivots():
    """Config 1 (A = L - 2I on the 100x100 grid, AMD): after equilibration a
    node whose four neighbours were eliminated as 1x1 pivots is left with
    1 - 4/4 = 0 up to rounding, and the panel-bounded Bunch-Kaufman (no
    delayed pivots in the left-looking schedule) takes it as a 1x1 pivot of
    size 2^-54. Every other pivot is O(1e-2) or larger, and the closed-form
    inertia (8163, 1837, 0) is reachable by the rounding-level pivots' signs
    alone. Refinement cannot repair such a factor: the ladder's FGMRES spends
    its 120 iterations and reports an uncertified backward error of ~0.2."""
    a, b, xt = matgen.config1()
    fo, s = _oracle_factor(a)
    neg, ntiny = rounding_level_bracket(fo)
    pos_cf, neg_cf, _, _ = O.laplacian2d_inertia(100, 2.0)
    assert (pos_cf, neg_cf) == (8163, 1837)
    assert ntiny > 0 and neg <= neg_cf <= neg + ntiny
    d = np.abs(fo.d_diag[~fo.two_by_two])
    assert np.all((d <= 1e-15) | (d >= 1e-3))  # a clean gap: tiny pivots are 2^-54
    x, ir, gm, be, cert = O.refine_ladder(a, fo, s, b, 1e-14)
    assert not cert and gm == 120 and be > 1e-3
