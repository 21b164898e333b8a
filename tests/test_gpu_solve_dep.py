"""Dependency-driven solve (one launch per sweep, units ordered by device
counters; the default for solves of up to 8 right-hand sides, wider ones run
the level launches) against the level-by-level launches of the same factor
and against the oracle's solve (solve_ldlt_many, ldlt_generic.rs:564-714).

The two device paths read the same factor and sum each entry's terms in
different groupings, so their solutions agree to rounding (1e-11 relative
on these well-conditioned cases); each meets the oracle bar of the parity
tests (1e-9 relative, backward error within 10x of the oracle's)."""
import dataclasses

import numpy as np
import pytest

import oracle as O
import rslab_amd as R
from rslab_amd import matgen

from test_gpu_parity import check_solution, dense_lower_csc, factor_pair, rhs

pytestmark = pytest.mark.gpu

FIELDS = ("l_col_ptr", "l_row_idx", "l_values", "d_diag", "d_subdiag", "two_by_two", "perm")


def dense_front(n=300, seed=7):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    return dense_lower_csc(M + M.T)


CASES = {
    # one supernode of 5 own blocks (2x2 pivots straddle block edges)
    "dense300": lambda: (dense_front(), R.SolverSettings().with_ordering(R.OrderingMethod.Natural)),
    "config3_24": lambda: (matgen.config3(m=24, nrhs=1)[0],
                           R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)),
    "config2_16": lambda: (matgen.config2(m=16, nrhs=1)[0],
                           R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)),
    "stencil27": lambda: (matgen.vector_stencil27(m=10, omega2=0.3),
                          R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("nrhs", [1, 5, 8, 13])
@pytest.mark.parametrize("split", [None, 0, 2])
def test_dependency_solve_equals_level_solve_and_oracle(case, nrhs, split):
    """split: the first tree level of the dependency-driven launch (None:
    automatic; 0: every level; 2: the two lowest levels as launches)."""
    a, st = CASES[case]()
    st = dataclasses.replace(st, solve_split=split)
    sym, fg, fo, s = factor_pair(a, st)
    sym0 = sym.replan(dataclasses.replace(st, solve_dep=False))
    f0 = sym0.factor(a)
    g, h = fg.factors(), f0.factors()
    for k in FIELDS:
        assert np.array_equal(getattr(g, k), getattr(h, k)), k
    B = rhs(a, nrhs, 71)
    X = check_solution(a, fg, fo, s, B)
    X0 = f0.solve_many(B)
    assert np.abs(X - X0).max() <= 1e-11 * np.abs(X0).max()
    # run to run identical
    assert np.array_equal(fg.solve_many(B), X)


def test_dependency_solve_batch_of_systems():
    systems = matgen.config4(count=48)
    mats = [x for x, _, _ in systems]
    big = R.CscMatrix.block_diag(mats)
    b = np.concatenate([bb for _, bb, _ in systems])
    sym = R.LdltSymbolic.analyze_batch(mats)
    f = sym.factor(big)
    x = f.solve(b)
    x0 = sym.replan(dataclasses.replace(R.SolverSettings(), solve_dep=False)).factor(big).solve(b)
    off = 0
    for a, bb, _ in systems:
        xs, x0s = x[off:off + a.n], x0[off:off + a.n]
        # shifts next to an eigenvalue make some systems ill-conditioned:
        # the bar is the level-launch solve's own backward error
        assert O.backward_error(a, xs, bb).max() <= 10 * max(O.backward_error(a, x0s, bb).max(), 1e-15)
        off += a.n


def test_dependency_solve_full_size_refines_config3():
    """60^3: the FP64 ladder certifies at 1e-14 with the dependency-driven
    solve, and its plain solve equals the level-launch one to 1e-10."""
    a, B = matgen.config3(m=60, nrhs=8)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym = R.LdltSymbolic.analyze_with(a, st)
    f = sym.factor(a)
    X, infos = R.MixedLdltSolver(f, a).solve_to(a, B, 1e-14)
    assert all(i.certified for i in infos), infos
    f0 = sym.replan(dataclasses.replace(st, solve_dep=False)).factor(a)
    assert np.abs(f.solve_many(B) - f0.solve_many(B)).max() <= 1e-10 * np.abs(X).max()
