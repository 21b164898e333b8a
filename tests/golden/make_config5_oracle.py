"""Writes config5_48_oracle.json: the CPU oracle's FP64 factor of the config-5
matrix (3 dof per node on 48^3 nodes, n = 331,776, 27-point coupling with
off-diagonal blocks -(1/26)(I + 0.25 d d^T/|d|^2), diagonal blocks minus the
row's off-diagonal sum plus 0.01 I, shifted by -omega^2 I; rslab-metis nested
dissection seed 1) at the first and the last of the 20 omega^2 values, with
the inertia, 2x2-block count, perturbation count, stored-L size and, for the
first 4 of the 32 right-hand sides of that omega^2, the plain solve's backward
errors and its solution at 64 fixed rows. The GPU test refactors with the same
symbolic and compares. Run from the repository root (OpenMP, all cores; about
ten minutes on 8 cores, 12 GB of memory):
python tests/golden/make_config5_oracle.py"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402
import rslab_amd as R  # noqa: E402
from rslab_amd import matgen  # noqa: E402

M = 48
K = (0, 19)      # omega^2 indices
NRHS_FIXTURE = 4


def main():
    oms = matgen.config5_omegas(20)
    a0 = matgen.vector_stencil27(m=M, omega2=oms[0])
    sym = R.LdltSymbolic.analyze_with(a0, R.SolverSettings().with_ordering(R.OrderingMethod.MetisND))
    e = sym.export()
    rows = [int(r) for r in np.linspace(0, a0.n - 1, 64)]
    out = dict(workload=f"config 5, 3 dof on {M}^3 nodes (n={a0.n}), nested dissection seed 1",
               n=a0.n, threads=O.get_threads(), rows=rows, omegas={})
    for k in K:
        a = matgen.vector_stencil27(m=M, omega2=oms[k])
        B = matgen.config5_rhs(a.n, k, nrhs=32)[:, :NRHS_FIXTURE]
        t0 = time.perf_counter()
        s = O.onepass_scale(a)
        fo = O.factor(e, O.scaled(a, s))
        t_f = time.perf_counter() - t0
        X = fo.solve_many(B * s[:, None]) * s[:, None]
        dt = time.perf_counter() - t0
        be = O.backward_error(a, X, B)
        out["omegas"][str(k)] = dict(
            omega2=float(oms[k]), inertia=list(fo.inertia), two_by_two_blocks=int(fo.two_by_two.sum()),
            n_perturbed=fo.n_perturbed, nnz_l=int(fo.l_row_idx.size), factor_seconds=round(t_f, 1),
            seconds=round(dt, 1), backward_error=[float(v) for v in be],
            x_rows=[[float(X[r, c]) for c in range(NRHS_FIXTURE)] for r in rows],
            x_col_inf=[float(np.abs(X[:, c]).max()) for c in range(NRHS_FIXTURE)])
        print(f"omega2[{k}] = {oms[k]:.4f}: inertia {fo.inertia}, factor {t_f:.1f} s, berr {be.max():.1e}",
              flush=True)
        del fo
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config5_48_oracle.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
