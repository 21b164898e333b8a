"""Right-looking schedule (FactorMethod::RightLooking,
multifrontal_ldlt.rs:1950-1977; same factor as the left-looking one,
:257-295) against the oracle and against the left-looking device factor.

The reference pins RightLooking to its Multifrontal factor bit for bit
(multifrontal_ldlt.rs:3730-3760) because both run the same per-front
arithmetic. Here the two schedules sum a target's updates in different
groupings (all updaters at once, or one updater level at a time), so the bar
is the parity bar of the left-looking path: identical pivot sequence,
inertia and perturbation count, L and D to 1e-9 relative, solutions to 1e-9
relative and backward error within 10x of the oracle's."""
import dataclasses

import numpy as np
import pytest

import oracle as O
import rslab_amd as R
from rslab_amd import matgen

from test_gpu_parity import assert_same_factor, check_solution, factor_pair, rhs

pytestmark = pytest.mark.gpu

RL = R.FactorMethod.RightLooking
FIELDS = ("l_col_ptr", "l_row_idx", "l_values", "d_diag", "d_subdiag", "two_by_two", "perm")


def nd(method=RL):
    return R.SolverSettings().with_ordering(R.OrderingMethod.MetisND).with_method(method)


@pytest.mark.parametrize("cap", [None, 0])
def test_right_looking_config3_shape_matches_oracle(cap):
    """L - 1.5 I + diag(delta) on 24^3, ND; cap 0 forces the gathered-tile
    cmod on every level instead of the two-phase one."""
    a, B = matgen.config3(m=24, nrhs=8)
    st = nd().with_cmod_cap(cap)
    sym, fg, fo, s = factor_pair(a, st)
    assert_same_factor(fg.factors(), fo)
    check_solution(a, fg, fo, s, B)


def test_right_looking_complex_config2_shape_matches_oracle():
    a, B = matgen.config2(m=16, nrhs=4)
    sym, fg, fo, s = factor_pair(a, nd())
    assert_same_factor(fg.factors(), fo)
    check_solution(a, fg, fo, s, B)


@pytest.mark.parametrize("wide", [False, True])
def test_right_looking_equals_left_looking_on_vector_stencil(wide):
    """27-point 3-dof stencil (many updaters per target, several levels).
    Its constant coefficients give exact Bunch-Kaufman magnitude ties, so the
    pivot choice follows the last bit of the updates, which the two
    groupings of the cmod sums round differently: the bar is the tie-laden
    one of the left-looking parity tests -- equal inertia, both factors valid
    (backward error <= 1e-13) and equal solutions to the conditioning."""
    a = matgen.vector_stencil27(m=10, omega2=0.3)
    st = nd()
    if wide:
        st = dataclasses.replace(st, wide_flops=1.0)
    sym_rl = R.LdltSymbolic.analyze_with(a, st)
    sym_ll = sym_rl.replan(st.with_method(R.FactorMethod.LeftLooking))
    frl, fll = sym_rl.factor(a), sym_ll.factor(a)
    assert frl.inertia() == fll.inertia()
    B = rhs(a, 3, 61)
    X, Y = frl.solve_many(B), fll.solve_many(B)
    assert O.backward_error(a, X, B).max() <= 1e-13
    assert O.backward_error(a, Y, B).max() <= 1e-13
    assert np.abs(X - Y).max() <= 1e-8 * np.abs(Y).max()


def test_right_looking_fp32_schur_refines_to_1e_14():
    a, B = matgen.config3(m=24, nrhs=4)
    st = nd()
    sym = R.LdltSymbolic.analyze_with(a, st)
    lo = sym.factor(a, st.with_precision(R.Precision.FP32_SCHUR))
    X, infos = R.MixedLdltSolver(lo, a).solve_to(a, B, 1e-14)
    assert all(i.certified and i.backward_error <= 1e-14 for i in infos), infos


def test_right_looking_graph_replay_equals_eager():
    """Refactors replayed as a captured graph give the eager right-looking
    factor bit for bit across changing values."""
    oms = matgen.config5_omegas(20)
    mats = [matgen.vector_stencil27(m=8, omega2=oms[k]) for k in (0, 7, 19, 7)]
    st = nd()
    sym = R.LdltSymbolic.analyze_with(mats[0], st)
    g = sym.factor(mats[0], st)
    e = sym.factor(mats[0], st.with_graphs(False))
    for a in mats:
        g.refactor(a)
        e.refactor(a)
        fg, fe = g.factors(), e.factors()
        for k in FIELDS:
            assert np.array_equal(getattr(fg, k), getattr(fe, k)), k
        assert g.inertia() == e.inertia()


def test_right_looking_config3_full_size_matches_oracle_fixture():
    """60^3 (n = 216,000): the right-looking factor's inertia equals the
    oracle's exactly and its plain solve equals the oracle's at the fixture's
    sampled rows to 1e-9."""
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                     "config3_60_oracle.json")))
    a, B = matgen.config3(m=60, nrhs=8)
    f = R.LdltSymbolic.analyze_with(a, nd()).factor(a)
    assert f.inertia().as_tuple() == tuple(fx["inertia"])
    assert f.n_perturbed() == fx["n_perturbed"]
    X = f.solve_many(B)
    scale = np.array(fx["x_col_inf"])[None, :]
    assert np.abs(X[fx["rows"]] - np.array(fx["x_rows"])).max() <= 1e-9 * scale.max()
