"""Writes config3_60_oracle.json: the CPU oracle's FP64 factor of the
216,000-unknown config-3 matrix (3D 7-point L - 1.5 I + diag(delta) on 60^3,
delta ~ U[-0.01, 0.01] from rslab's xorshift64* seed 3, rslab-metis nested
dissection seed 1) and its 8-RHS solve (seed 4): inertia, 2x2-block count,
perturbation count, stored-L size, the plain solve's backward errors, and the
solution at 64 fixed rows of every column. The GPU test compares the device
factor's inertia with it exactly and the device solution with the sampled
rows. Run from the repository root: python tests/golden/make_config3_oracle.py
(about a minute on one core)."""
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

M = 60
ROWS = [int(r) for r in np.linspace(0, M ** 3 - 1, 64)]


def main():
    a, B = matgen.config3(m=M, nrhs=8)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym = R.LdltSymbolic.analyze_with(a, st)
    t0 = time.perf_counter()
    s = O.onepass_scale(a)
    fo = O.factor(sym.export(), O.scaled(a, s))
    X = fo.solve_many(B * s[:, None]) * s[:, None]
    dt = time.perf_counter() - t0
    be = O.backward_error(a, X, B)
    out = dict(
        workload=f"config 3, {M}^3, nested dissection seed 1, delta seed 3, 8 RHS seed 4",
        n=a.n, inertia=list(fo.inertia), two_by_two_blocks=int(fo.two_by_two.sum()),
        n_perturbed=fo.n_perturbed, nnz_l=int(fo.l_row_idx.size), oracle_seconds=round(dt, 1),
        backward_error=[float(v) for v in be], rows=ROWS,
        x_rows=[[float(X[r, c]) for c in range(X.shape[1])] for r in ROWS],
        x_col_inf=[float(np.abs(X[:, c]).max()) for c in range(X.shape[1])],
    )
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config3_60_oracle.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {path}: inertia {out['inertia']}, {out['two_by_two_blocks']} 2x2 blocks, {dt:.1f} s")


if __name__ == "__main__":
    main()
