"""Autotuner for the device plan's tuning knobs, by measurement on the matrix
at hand (the reference tunes its CPU kernels' GEMM thresholds and panel width
per matrix, src/numeric/gemm_tuning.rs:16-50, from a cost model and a
calibration, src/tuning.rs, src/auto_tune.rs; on the GPU the knobs are the
device plan's cutovers, and they are timed directly):

* ``wide_flops``: a supernode with h*w^2 above it factors its panels in the
  multi-workgroup batched steps, below it in one workgroup (the small/large
  front cutover);
* the solve schedule: the dependency-driven sweeps (one launch per sweep,
  key ``SOLVE_DEP``), or the level-by-level launches with ``solve_wide``
  (own blocks wider than it are solved in multi-workgroup 64-column steps,
  narrower ones in one workgroup).

Each candidate reuses the analysis (``LdltSymbolic.replan``: same ordering
and partition, another device plan). The factor knob is timed by same-pattern
refactors (graph replay, as a sweep runs them), the solve knob by multi-RHS
solves; the fastest of each is kept. The pivot sequence does not depend on
either knob (both select how, not what, the device computes); only the
summation order inside the kernels changes.
"""
from __future__ import annotations

import dataclasses
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from .solver import CscMatrix, LdltSymbolic, SolverSettings

WIDE_FLOPS = (6.4e4, 2.5e5, 1.0e6, 4.0e6)
SOLVE_WIDE = (0, 32, 128)
SOLVE_DEP = -1  # solve candidate key of the dependency-driven sweeps


@dataclasses.dataclass
class TuneResult:
    settings: SolverSettings
    factor_ms: Dict[float, float]   # wide_flops -> best refactor time
    solve_ms: Dict[int, float]      # SOLVE_DEP or solve_wide (level launches) -> best solve time
    nrhs: int


def _best(f, reps):
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        f()
        t.append((time.perf_counter() - t0) * 1e3)
    return min(t)


def autotune(a: CscMatrix, sym: LdltSymbolic, opts: Optional[SolverSettings] = None, nrhs: int = 1,
             wide_flops: Sequence[float] = WIDE_FLOPS, solve_wide: Sequence[int] = SOLVE_WIDE,
             reps: int = 3) -> TuneResult:
    """Times the candidates on `a` (its pattern analysed by `sym`) and returns
    `opts` with the fastest wide_flops and solve schedule."""
    opts = opts or sym.settings
    b = np.ones((a.n, nrhs), dtype=np.complex128 if a.is_complex else np.float64)
    fac: Dict[float, float] = {}
    handles: List = []
    for wf in wide_flops:
        st = dataclasses.replace(opts, wide_flops=float(wf))
        s2 = sym.replan(st)
        f = s2.factor(a, st)
        f.refactor(a)  # captures the graph
        fac[float(wf)] = _best(lambda: f.refactor(a), reps)
        handles.append((s2, f))
    best_wf = min(fac, key=fac.get)
    sol: Dict[int, float] = {}
    for sw in (SOLVE_DEP,) + tuple(solve_wide):
        if sw == SOLVE_DEP:
            st = dataclasses.replace(opts, wide_flops=best_wf, solve_dep=True)
        else:
            st = dataclasses.replace(opts, wide_flops=best_wf, solve_dep=False, solve_wide=int(sw))
        s2 = sym.replan(st)
        f = s2.factor(a, st)
        f.solve_many(b)
        sol[int(sw)] = _best(lambda: f.solve_many(b), reps)
        handles.append((s2, f))
    best_sw = min(sol, key=sol.get)
    del handles
    if best_sw == SOLVE_DEP:
        tuned = dataclasses.replace(opts, wide_flops=best_wf, solve_dep=True)
    else:
        tuned = dataclasses.replace(opts, wide_flops=best_wf, solve_dep=False, solve_wide=best_sw)
    return TuneResult(tuned, fac, sol, nrhs)
