"""The autotuner's plumbing on the CPU: a re-plan keeps the analysis
(permutation, partition, pattern) and rebuilds only the device plan; the
tuning knobs reach the plan (rsl_symbolic_replan, rslab_amd.autotune)."""
import dataclasses

import numpy as np

import rslab_amd as R
from rslab_amd import matgen


def test_replan_keeps_the_analysis():
    a, _ = matgen.config3(m=14, nrhs=1)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym = R.LdltSymbolic.analyze_with(a, st)
    e = sym.export()
    for wf, sw in ((1.0, 0), (1e30, 1000), (2.5e5, 32)):
        s2 = sym.replan(dataclasses.replace(st, wide_flops=wf, solve_wide=sw))
        e2 = s2.export()
        for k in ("perm", "perm_inv", "sn_first", "sn_ncol", "sn_nrow", "ch_ptr", "ch_idx", "pp_colp", "pp_rowi"):
            assert np.array_equal(getattr(e, k), getattr(e2, k)), k
        i1, i2 = sym.info(), s2.info()
        assert i1["factor_flops"] == i2["factor_flops"] and i1["levels"] == i2["levels"]


def test_settings_carry_the_solve_cutover():
    s = R.SolverSettings().to_c()
    assert s.solve_wide == 32 and s.graphs == 1 and s.drop_tol == -1.0 and s.cmod_cap == -1
    assert dataclasses.replace(R.SolverSettings(), solve_wide=0).to_c().solve_wide == 0
    assert s.solve_dep == 1 and s.method == 0
    assert dataclasses.replace(R.SolverSettings(), solve_dep=False).to_c().solve_dep == 0
