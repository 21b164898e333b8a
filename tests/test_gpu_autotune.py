"""The autotuner on the device (rslab_amd.autotune): it times every candidate
cutover on the matrix and returns one of them; the knobs choose how the device
computes, not what, so the tuned factor has the default factor's pivot
sequence and inertia, the oracle's, and the same solutions to rounding."""
import numpy as np
import pytest

import oracle as O
import rslab_amd as R
from rslab_amd import matgen
from rslab_amd.autotune import SOLVE_DEP, SOLVE_WIDE, WIDE_FLOPS, autotune

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_candidate_and_keeps_the_factor():
    a, B = matgen.config3(m=24, nrhs=4)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym = R.LdltSymbolic.analyze_with(a, st)
    res = autotune(a, sym, st, nrhs=4, reps=2)
    assert res.settings.wide_flops in WIDE_FLOPS and res.settings.solve_wide in SOLVE_WIDE
    assert set(res.factor_ms) == set(WIDE_FLOPS) and set(res.solve_ms) == {SOLVE_DEP, *SOLVE_WIDE}
    assert all(t > 0 for t in list(res.factor_ms.values()) + list(res.solve_ms.values()))
    tuned = sym.replan(res.settings).factor(a, res.settings)
    plain = sym.factor(a)
    ft, fp = tuned.factors(), plain.factors()
    assert np.array_equal(ft.perm, fp.perm) and np.array_equal(ft.two_by_two, fp.two_by_two)
    assert tuned.inertia() == plain.inertia()
    s = O.onepass_scale(a)
    fo = O.factor(sym.export(), O.scaled(a, s))
    assert tuned.inertia().as_tuple() == fo.inertia
    X1, X2 = tuned.solve_many(B), plain.solve_many(B)
    assert np.abs(X1 - X2).max() <= 1e-10 * np.abs(X2).max()
