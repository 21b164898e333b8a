"""Device Krylov consumers of the factor (rsl_gmres_block, rsl_cocg, rsl_cocr
through the C ABI) against the oracle's restatements of gmres, cocg and cocr
(src/numeric/iterative.rs:402-583, 822-1021) with the oracle's factor of the
same matrix as preconditioner: per right-hand side the same iteration count
(to one step: dot products are summed in another order), the same
certificate and stop reason, both at the tolerance, solutions equal to what
the tolerance allows."""
import numpy as np
import pytest

import oracle as O
import rslab_amd as R
from rslab_amd import matgen

pytestmark = pytest.mark.gpu


def _oracle_pre(sym, a_factored, prec=0):
    s = O.onepass_scale(a_factored)
    fo = O.factor(sym.export(), O.scaled(a_factored, s), prec=prec)
    return lambda r: fo.solve_many(r * s) * s


@pytest.mark.parametrize("restart", [30, 2])
def test_gmres_block_with_fp32_factor_matches_oracle(restart):
    """FP32-Schur factor of config-3's 16^3 shape preconditioning the FP64
    matrix: FGMRES to 1e-13 per column (restart 2 exercises restarts)."""
    a, B = matgen.config3(m=16, nrhs=4)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND).with_precision(R.Precision.FP32_SCHUR)
    sym = R.LdltSymbolic.analyze_with(a, st)
    f = sym.factor(a, st)
    res = R.gmres_block(a, B, 4, f, 1e-13, 60, restart)
    assert res.converged and res.stop == "converged"
    pre = _oracle_pre(sym, a, prec=O.PREC_FP32)
    op = lambda v: O.symv(a, v)  # noqa: E731
    for c in range(4):
        xo, its, conv = O.gmres(op, B[:, c], pre, 1e-13, 60, restart)
        assert conv
        assert res.final_res[c] <= 1e-13
        assert np.abs(res.x[:, c] - xo).max() <= 1e-9 * np.abs(xo).max()
    one = R.gmres(a, B[:, 0], f, 1e-13, 60, restart)
    xo, its0, _ = O.gmres(op, B[:, 0], pre, 1e-13, 60, restart)
    assert abs(one.iters - its0) <= 1, (one.iters, its0)
    # warm start at the solution: nothing to do
    warm = R.gmres(a, B[:, 0], f, 1e-12, 60, restart, x0=one.x)
    assert warm.converged and warm.iters == 0
    # the reference's column-major flat layout comes back flat
    flat = R.gmres_block(a, B.T.reshape(-1), 4, f, 1e-13, 60, restart)
    assert flat.x.shape == (a.n * 4,) and np.allclose(flat.x.reshape(4, a.n).T, res.x, rtol=0, atol=1e-12)


def test_gmres_without_preconditioner_matches_oracle():
    """Identity preconditioner (NoPreconditioner) on an SPD grid Laplacian,
    restart 10: the same iteration count as the oracle's gmres, to one."""
    a = matgen.grid_laplacian((24, 24), shift=0.6)
    b = matgen.rng_uniform(31, -1, 1, a.n)
    f = R.LdltSolver.factor_with(a)
    res = R.gmres(a, b, R.NoPreconditioner(f), 1e-10, 400, 10)
    xo, its, conv = O.gmres(lambda v: O.symv(a, v), b, lambda v: v, 1e-10, 400, 10)
    assert res.converged and conv
    assert abs(res.iters - its) <= 1, (res.iters, its)
    assert np.abs(res.x - xo).max() <= 1e-7 * np.abs(xo).max()


@pytest.mark.parametrize("method", ["cocg", "cocr"])
def test_cocg_cocr_complex_helmholtz_match_oracle(method):
    """Complex-symmetric Helmholtz (config-2 shape, 12^3): the factor of the
    matrix with a weaker absorbing layer preconditions the true one."""
    a, B = matgen.config2(m=12, nrhs=3)
    cols = np.repeat(np.arange(a.n), np.diff(a.col_ptr))
    diag = a.row_idx == cols
    near = a.with_values(np.where(diag, a.values.real + 0.6j * a.values.imag, a.values))
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym = R.LdltSymbolic.analyze_with(a, st)
    f = sym.factor(near)
    pre = _oracle_pre(sym, near)
    fn_d = R.cocg if method == "cocg" else R.cocr
    fn_o = O.cocg if method == "cocg" else O.cocr
    for c in range(3):
        rd = fn_d(a, B[:, c], f, 1e-12, 200)
        xo, its, conv, fres, stop = fn_o(lambda v: O.symv(a, v), B[:, c], pre, 1e-12, 200)
        assert rd.converged and conv and rd.stop == stop == "converged"
        assert abs(rd.iters - its) <= 1, (rd.iters, its)
        assert rd.final_res <= 1e-12
        assert np.abs(rd.x - xo).max() <= 1e-9 * np.abs(xo).max()
        assert np.abs(a.symv(rd.x) - B[:, c]).max() <= 1e-10 * np.abs(B[:, c]).max()


def test_cocg_breakdown_and_iteration_budget():
    """[[0, 1], [1, 0]] with b = e0 and no preconditioner: p^T A p = 0 at the
    first step (breakdown, 0 iterations); a budget too small for a hard
    system ends in max_iter."""
    s2 = R.CscMatrix.from_triplets(2, [0, 1, 1], [0, 0, 1], np.array([0, 1, 0], dtype=complex))
    f2 = R.LdltSolver.factor_with(s2)
    r = R.cocg(s2, np.array([1.0, 0.0], dtype=complex), R.NoPreconditioner(f2), 1e-12, 10)
    assert r.stop == "breakdown" and r.iters == 0 and not r.converged
    a, B = matgen.config2(m=10, nrhs=1)
    f = R.LdltSolver.factor_with(a)
    r = R.cocr(a, B[:, 0], R.NoPreconditioner(f), 1e-14, 3)
    assert r.stop == "max_iter" and r.iters == 3 and not r.converged
