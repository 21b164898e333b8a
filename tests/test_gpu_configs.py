"""The configurations at their full workload sizes, on the device through the
C ABI, against the CPU oracle's fixtures (tests/golden/make_config*_oracle.py
wrote them from the same inputs and symbolic) and closed forms:

* config 2: 40^3 complex-symmetric Helmholtz (n = 64,000), nested dissection,
  16 right-hand sides;
* config 4: 1,024 independent shifted 2D Laplacians of sides 16..96
  (sum n = 3.77 M), AMD, one batched analysis / factor / solve;
* config 5: 3 dof on 48^3 nodes (n = 331,776), 27-point vector stencil,
  refactor sweep with the analysis reused, first and last of the 20 omega^2,
  32 right-hand sides, FP32-Schur factor refined to backward error <= 1e-12.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import rslab_amd as R
from rslab_amd import matgen

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    with open(os.path.join(GOLD, name)) as fh:
        return json.load(fh)


def test_config2_full_size_matches_oracle_fixture():
    fx = fixture("config2_40_oracle.json")
    a, B = matgen.config2(m=40, nrhs=16)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    f = R.LdltSymbolic.analyze_with(a, st).factor(a)
    g = f.factors()
    assert f.inertia().as_tuple() == tuple(fx["inertia_by_real_part"])
    assert int(g.two_by_two.sum()) == fx["two_by_two_blocks"]
    assert f.n_perturbed() == fx["n_perturbed"]
    X = f.solve_many(B)
    xs = np.array(fx["x_rows_re"]) + 1j * np.array(fx["x_rows_im"])
    assert np.abs(X[fx["rows"]] - xs).max() <= 1e-9 * max(fx["x_col_inf"])
    be = O.backward_error(a, X, B)
    assert np.all(be <= 10 * np.maximum(np.array(fx["backward_error"]), 1e-16)), (be, fx["backward_error"])
    # refined on the device against the FP64 matrix
    Xr, infos = R.MixedLdltSolver(f, a).solve_to(a, B, 1e-14)
    assert all(i.certified for i in infos), infos
    assert O.backward_error(a, Xr, B).max() <= 1e-14


@pytest.mark.parametrize("prec", [R.Precision.FP64, R.Precision.FP32_SCHUR])
def test_config4_all_1024_systems(prec):
    """Per-system inertia equal to the closed form #{lambda_pq < sigma}
    wherever the spectrum keeps away from sigma, every system's backward error
    at FP64 level after the batched ladder, and the batch's inertias summing
    to n."""
    systems = matgen.config4(count=1024)
    mats = [a for a, _, _ in systems]
    assert sum(a.n for a in mats) > 3.7e6
    big = R.CscMatrix.block_diag(mats)
    b = np.concatenate([bb for _, bb, _ in systems])
    sym = R.LdltSymbolic.analyze_batch(mats)
    f = sym.factor(big, sym.settings.with_precision(prec))
    inert = f.inertia_batch(len(mats))
    assert sum(i.total() for i in inert) == big.n
    checked = 0
    for (a, _, sg), ib in zip(systems, inert):
        m = int(round(np.sqrt(a.n)))
        t = 2 - 2 * np.cos(np.arange(1, m + 1) * np.pi / (m + 1))
        lam = (t[:, None] + t[None, :]).ravel() - sg
        if np.abs(lam).min() > 1e-6:
            assert ib.as_tuple() == (int((lam > 0).sum()), int((lam < 0).sum()), 0), (a.n, sg)
            checked += 1
    assert checked >= 1000
    x, info = R.MixedLdltSolver(f, big).solve_to(big, b, 1e-14)
    assert info.certified, info
    off = 0
    worst = 0.0
    for a, bb, _ in systems:
        worst = max(worst, O.backward_error(a, x[off:off + a.n], bb)[0])
        off += a.n
    assert worst <= 1e-13, worst


def test_config5_full_size_sweep_ends():
    fx = fixture("config5_48_oracle.json")
    oms = matgen.config5_omegas(20)
    st = R.SolverSettings().with_ordering(R.OrderingMethod.MetisND)
    a0 = matgen.vector_stencil27(m=48, omega2=oms[0])
    sym = R.LdltSymbolic.analyze_with(a0, st)
    f64 = sym.factor(a0)
    f32 = sym.factor(a0, st.with_precision(R.Precision.FP32_SCHUR))
    for key, e in sorted(fx["omegas"].items(), key=lambda kv: int(kv[0])):
        k = int(key)
        a = matgen.vector_stencil27(m=48, omega2=oms[k])
        f64.refactor(a)
        assert f64.inertia().as_tuple() == tuple(e["inertia"]), (k, f64.inertia(), e["inertia"])
        assert f64.n_perturbed() == e["n_perturbed"]
        B = matgen.config5_rhs(a.n, k, nrhs=32)
        X = f64.solve_many(B[:, :4])
        xs = np.array(e["x_rows"])
        assert np.abs(X[fx["rows"]] - xs).max() <= 1e-9 * max(e["x_col_inf"])
        be = O.backward_error(a, X, B[:, :4])
        assert np.all(be <= 10 * np.maximum(np.array(e["backward_error"]), 1e-16))
        f32.refactor(a)
        X32, infos = R.MixedLdltSolver(f32, a).solve_to(a, B, 1e-12)
        assert all(i.certified for i in infos), infos
        assert O.backward_error(a, X32, B).max() <= 1e-12
        fwd = np.abs(X32[fx["rows"], :4] - xs).max() / max(e["x_col_inf"])
        print(f"config 5 omega2[{k}]: inertia {f64.inertia().as_tuple()}, FP32+IR "
              f"{max(i.ir_iters for i in infos)} IR steps, forward error vs oracle {fwd:.1e}")
