"""Pins for the oracle's COCG and COCR restatements (oracle.cocg, oracle.cocr,
after src/numeric/iterative.rs:402-583), independent of their own
recurrences: for a real SPD matrix COCG is CG, whose k-th iterate minimizes
the A-norm of the error over the Krylov space, and COCR is CR, whose k-th
iterate minimizes the residual 2-norm there; an exact preconditioner solves
in one step; a zero bilinear denominator is a breakdown; and a planted
mistake (conjugated products, which make COCG the wrong method for a
complex-symmetric matrix) is caught."""
import numpy as np
import pytest

import oracle as O


def _spd(n=60, seed=2):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    return M @ M.T / n + 0.5 * np.eye(n), rng.standard_normal(n)


def _krylov_basis(A, b, k):
    K = np.zeros((b.size, k), dtype=np.result_type(A, b))
    v = b.copy()
    for j in range(k):
        K[:, j] = v
        v = A @ v
    Q, _ = np.linalg.qr(K)
    return Q


@pytest.mark.parametrize("k", [1, 3, 7])
def test_cocg_is_cg_on_spd(k):
    A, b = _spd()
    x, it, conv, res, stop = O.cocg(lambda v: A @ v, b, lambda v: v, 1e-300, k)
    assert it == k and stop == "max_iter" and not conv
    Q = _krylov_basis(A, b, k)
    xg = Q @ np.linalg.solve(Q.T @ A @ Q, Q.T @ b)  # Galerkin = A-norm minimizer over K_k
    assert np.linalg.norm(x - xg) <= 1e-9 * np.linalg.norm(xg)
    assert res == pytest.approx(np.linalg.norm(b - A @ x) / np.linalg.norm(b), rel=1e-8)


@pytest.mark.parametrize("k", [1, 3, 7])
def test_cocr_is_cr_on_spd(k):
    A, b = _spd()
    x, it, conv, res, stop = O.cocr(lambda v: A @ v, b, lambda v: v, 1e-300, k)
    assert it == k
    Q = _krylov_basis(A, b, k)
    y = np.linalg.lstsq(A @ Q, b, rcond=None)[0]  # residual minimizer over K_k
    assert np.linalg.norm(b - A @ x) == pytest.approx(np.linalg.norm(b - A @ Q @ y), rel=1e-8)


def test_exact_preconditioner_and_breakdown_complex_symmetric():
    rng = np.random.default_rng(4)
    M = rng.standard_normal((30, 30)) + 1j * rng.standard_normal((30, 30))
    A = M + M.T + 8 * np.eye(30)
    b = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    Ai = np.linalg.inv(A)
    for fn in (O.cocg, O.cocr):
        x, it, conv, res, stop = fn(lambda v: A @ v, b, lambda v: Ai @ v, 1e-12, 50)
        assert conv and it == 1 and stop == "converged"
        assert np.linalg.norm(b - A @ x) <= 1e-12 * np.linalg.norm(b)
    # [[0, 1], [1, 0]] with b = e0: p^T A p = 0 at the first step
    S = np.array([[0.0, 1.0], [1.0, 0.0]], dtype=complex)
    x, it, conv, res, stop = O.cocg(lambda v: S @ v, np.array([1.0, 0.0], dtype=complex), lambda v: v, 1e-12, 10)
    assert stop == "breakdown" and it == 0 and not conv


def test_cocg_converges_on_complex_symmetric_and_conjugation_is_caught():
    """Unpreconditioned COCG on a complex-symmetric, diagonally dominant
    matrix converges; replacing the bilinear products by Hermitian ones (a
    plausible slip) turns it into a method that does not, on the same budget."""
    rng = np.random.default_rng(6)
    n = 80
    M = 0.3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    A = M + M.T + (3 + 1j) * np.eye(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x, it, conv, res, _ = O.cocg(lambda v: A @ v, b, lambda v: v, 1e-10, 60)
    assert conv and np.linalg.norm(b - A @ x) <= 1.01e-10 * np.linalg.norm(b)

    def cocg_conj(op, b, tol, max_iter):  # the planted mistake: vdot instead of sum
        x = np.zeros_like(b)
        r = b.copy()
        p = r.copy()
        rho = np.vdot(r, r)
        for _ in range(max_iter):
            q = op(p)
            alpha = rho / np.vdot(p, q)
            x, r = x + alpha * p, r - alpha * q
            if np.linalg.norm(r) <= tol * np.linalg.norm(b):
                return x, True
            rho_new = np.vdot(r, r)
            p = r + rho_new / rho * p
            rho = rho_new
        return x, False

    _, ok = cocg_conj(lambda v: A @ v, b, 1e-10, it)
    assert not ok
