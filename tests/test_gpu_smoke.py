import numpy as np
import pytest

import oracle as O
import rslab_amd as R
from rslab_amd import matgen


@pytest.mark.gpu
def test_known_answer_2x2():
    # ldlt_generic.rs:836-850: [[0,1],[1,0]] x = [3,5] -> x = [5,3]; inertia (1,1,0)
    a = R.CscMatrix.from_triplets(2, [0, 1], [0, 0], [0.0, 1.0])
    f = R.LdltSolver.factor_with(a)
    x = f.solve(np.array([3.0, 5.0]))
    assert np.allclose(x, [5.0, 3.0], atol=1e-14)
    assert f.inertia().as_tuple() == (1, 1, 0)
    assert f.factors().two_by_two.any()


@pytest.mark.gpu
def test_indefinite_grid_matches_oracle():
    a = matgen.grid_laplacian((24, 24), shift=-3.5)
    b = matgen.rng_uniform(11, -1, 1, a.n)
    sym = R.LdltSymbolic.analyze_with(a, R.SolverSettings())
    f = sym.factor(a)
    x = f.solve(b)
    xo, fo = O.solve(a, b, sym.export())
    assert f.inertia().as_tuple() == fo.inertia
    assert np.abs(x - xo).max() <= 1e-9 * np.abs(xo).max()
