"""Writes config2_40_oracle.json: the CPU oracle's factor of the 64,000-unknown
complex-symmetric config-2 matrix (3D 7-point Helmholtz L - 1.2 I + i diag(eta)
on 40^3, eta rising quadratically to 0.5 across a 4-cell absorbing layer,
rslab-metis nested dissection seed 1) and its 16-RHS solve (real and
imaginary parts from rslab's xorshift64* seeds): the pivot structure (2x2
count, inertia by real part, perturbation count, stored-L size), the plain
solve's backward errors and the solution at 64 fixed rows of every column.
The GPU test compares the device factor and solve with it. Run from the
repository root: python tests/golden/make_config2_oracle.py (about a minute)."""
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

M = 40
ROWS = [int(r) for r in np.linspace(0, M ** 3 - 1, 64)]


def main():
    a, B = matgen.config2(m=M, nrhs=16)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    sym = R.LdltSymbolic.analyze_with(a, st)
    t0 = time.perf_counter()
    s = O.onepass_scale(a)
    fo = O.factor(sym.export(), O.scaled(a, s))
    X = fo.solve_many(B * s[:, None]) * s[:, None]
    dt = time.perf_counter() - t0
    be = O.backward_error(a, X, B)
    out = dict(
        workload=f"config 2, {M}^3 complex symmetric, nested dissection seed 1, 16 RHS",
        n=a.n, inertia_by_real_part=list(fo.inertia), two_by_two_blocks=int(fo.two_by_two.sum()),
        n_perturbed=fo.n_perturbed, nnz_l=int(fo.l_row_idx.size), oracle_seconds=round(dt, 1),
        backward_error=[float(v) for v in be], rows=ROWS,
        x_rows_re=[[float(X[r, c].real) for c in range(X.shape[1])] for r in ROWS],
        x_rows_im=[[float(X[r, c].imag) for c in range(X.shape[1])] for r in ROWS],
        x_col_inf=[float(np.abs(X[:, c]).max()) for c in range(X.shape[1])],
    )
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config2_40_oracle.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {path}: {out['two_by_two_blocks']} 2x2 blocks, nnz(L) {out['nnz_l']}, {dt:.1f} s")


if __name__ == "__main__":
    main()
