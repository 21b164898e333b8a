"""Same-pattern refactors replayed as a captured HIP graph (settings.graphs,
the default) against eager launches of the same sequence (graphs off), on
one analysis: every replay gives the eager factor bit for bit, across
changing values, interleaved eager work and host synchronizations, the
device-resident entry with changing value buffers, a failing refactor in
between, and a pivot policy whose floor changes with the values (the graph
is captured again then)."""
import ctypes as C

import numpy as np
import pytest

import rslab_amd as R
from rslab_amd import matgen

pytestmark = pytest.mark.gpu

FIELDS = ("l_col_ptr", "l_row_idx", "l_values", "d_diag", "d_subdiag", "two_by_two", "perm")


def same(f, g):
    a, b = f.factors(), g.factors()
    for k in FIELDS:
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert f.inertia() == g.inertia() and f.n_perturbed() == g.n_perturbed()


def sweep_values(m=10, count=4):
    oms = matgen.config5_omegas(20)
    return [matgen.vector_stencil27(m=m, omega2=oms[k]) for k in np.linspace(0, 19, count).astype(int)]


@pytest.mark.parametrize("prec", [R.Precision.FP64, R.Precision.FP32_SCHUR])
def test_graph_replay_equals_eager_across_a_sweep(prec):
    mats = sweep_values()
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND).with_precision(prec)
    sym = R.LdltSymbolic.analyze_with(mats[0], st)
    g = sym.factor(mats[0], st)                       # graphs on (default)
    e = sym.factor(mats[0], st.with_graphs(False))   # eager
    B = matgen.config5_rhs(mats[0].n, 0, nrhs=3)
    for rep in range(2):
        for a in mats + [mats[1]]:
            g.refactor(a)   # captured on the first refactor, replayed after
            e.refactor(a)
            same(g, e)
            assert np.array_equal(g.solve_many(B), e.solve_many(B))


def test_graph_replay_device_values_and_failure_in_between():
    hip = C.CDLL("libamdhip64.so.7")
    a, _ = matgen.config3(m=16, nrhs=1)
    a2 = a.with_values(a.values * 1.1 - 0.05)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym = R.LdltSymbolic.analyze_with(a, st)
    g = sym.factor(a, st)
    e = sym.factor(a, st.with_graphs(False))
    bufs = []
    try:
        for vals in (a, a2, a, a2):
            v = np.ascontiguousarray(vals.values_f64())
            d = C.c_void_p()
            assert hip.hipMalloc(C.byref(d), C.c_size_t(v.nbytes)) == 0
            bufs.append(d)
            assert hip.hipMemcpy(d, v.ctypes.data_as(C.c_void_p), C.c_size_t(v.nbytes), 1) == 0
            g.refactor_device(d.value)   # a new device buffer every time
            e.refactor(vals)
            same(g, e)
        # a refactor that fails (a zero row makes a zero pivot), then a good one
        z = a.with_values(np.where(a.row_idx == 0, 0.0, a.values))
        cols = np.repeat(np.arange(a.n), np.diff(a.col_ptr))
        z = z.with_values(np.where(cols == 0, 0.0, z.values))
        with pytest.raises(R.NumericallyRankDeficient):
            g.refactor(z)
        g.refactor(a2)
        e.refactor(a2)
        same(g, e)
    finally:
        for d in bufs:
            hip.hipFree(d)


def test_graph_recaptured_when_the_pivot_floor_changes():
    """ForceAccept's floor is max(|A_hat|max, 1) eps of the values at hand
    (without equilibration it follows the scale of A): a refactor whose floor
    differs captures the sequence again."""
    a, _ = matgen.config3(m=14, nrhs=1)
    st = (R.SolverSettings().with_ordering(R.OrderingMethod.MetisND).with_pivot(R.ZeroPivotAction.ForceAccept())
          .with_scaling("Identity"))
    sym = R.LdltSymbolic.analyze_with(a, st)
    g = sym.factor(a, st)
    e = sym.factor(a, st.with_graphs(False))
    for scale in (1.0, 3.0, 1.0):
        v = a.with_values(a.values * scale)
        g.refactor(v)
        e.refactor(v)
        same(g, e)
