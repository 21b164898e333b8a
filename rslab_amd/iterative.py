"""Krylov consumers of the device factor, mirroring rslab's iterative module
(src/numeric/iterative.rs): a factored ``LdltSolver`` is a ``Preconditioner``
(:178-191; its block apply is the multi-RHS device solve), and

* ``gmres``        flexible restarted GMRES (:822-1021), optional warm start
* ``gmres_block``  the same for s right-hand sides (:1881); every column keeps
                   its own Arnoldi basis, Hessenberg and rotations and follows
                   the single-RHS recurrence (MGS with DGKS), the columns
                   sharing the batched operator and preconditioner applies
* ``cocg``         COCG for complex-symmetric A = A^T (:402-489)
* ``cocr``         COCR (:495-583)

run on the GPU through the C ABI (rsl_gmres_block, rsl_cocg, rsl_cocr), with
the operator A given by its values in the factor's analysed pattern (they may
differ from the factored matrix: a factor of a nearby matrix preconditions A).
``NoPreconditioner(solver)`` is the identity (:167-176) on the device context
of ``solver``.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import List, Optional, Union

import numpy as np

from ._lib import lib, ptr
from .solver import CscMatrix, DimensionMismatch, InvalidInput, LdltSolver, _check

_STOP = {0: "converged", 1: "max_iter", 2: "breakdown"}


@dataclasses.dataclass
class KrylovResult:
    """KrylovResult (iterative.rs:379-392)."""

    x: np.ndarray
    iters: int
    converged: bool
    final_res: float
    stop: str


@dataclasses.dataclass
class BlockKrylovResult:
    """BlockKrylovResult (iterative.rs:1764-1776): x in the layout b came in."""

    x: np.ndarray
    iters: int
    converged: bool
    final_res: List[float]
    stop: str


class NoPreconditioner:
    """The identity preconditioner M = I, run on `solver`'s device context."""

    def __init__(self, solver: LdltSolver):
        self.solver = solver


Precond = Union[LdltSolver, NoPreconditioner]


def _handle(precond: Precond):
    if isinstance(precond, NoPreconditioner):
        return precond.solver, 0
    if isinstance(precond, LdltSolver):
        return precond, 1
    raise InvalidInput("precond must be an LdltSolver or NoPreconditioner(LdltSolver)")


def _run(method, a: CscMatrix, B: np.ndarray, precond: Precond, tol, max_iter, restart=30, x0=None):
    f, use = _handle(precond)
    n = f.n()
    if a.n != n or a.nnz != f._sym._nnz:
        raise InvalidInput("the operator must have the factor's analysed pattern")
    if B.ndim != 2 or B.shape[0] != n:
        raise DimensionMismatch(f"expected {n} rows, got {B.shape}")
    k = B.shape[1]
    if k == 0:
        raise DimensionMismatch("no right-hand sides")
    if a.is_complex != f._cplx:
        raise InvalidInput("operator and factor must be over the same field")
    bb, bv = f._as_c(B)
    x = np.zeros_like(bb)
    xv = x.view(np.float64) if f._cplx else x
    info = np.zeros(4 * k, dtype=np.float64)
    vals = a.values_f64()
    if method == 0:
        x0v = None
        if x0 is not None:
            x0c, x0v = f._as_c(np.asarray(x0).reshape(n, k))
        _check(lib.rsl_gmres_block(f._h, ptr(vals, C.c_double), ptr(bv, C.c_double), k, float(tol),
                                   int(max_iter), int(restart), use, ptr(x0v, C.c_double),
                                   ptr(xv, C.c_double), ptr(info, C.c_double)))
    else:
        fn = lib.rsl_cocg if method == 1 else lib.rsl_cocr
        _check(fn(f._h, ptr(vals, C.c_double), ptr(bv, C.c_double), k, float(tol), int(max_iter), use,
                  ptr(xv, C.c_double), ptr(info, C.c_double)))
    cols = [(int(info[4 * c]), float(info[4 * c + 1]), bool(info[4 * c + 2]), _STOP[int(info[4 * c + 3])])
            for c in range(k)]
    return x, cols


def _single(method, a, b, precond, tol, max_iter, restart=30, x0=None) -> KrylovResult:
    b = np.asarray(b)
    if b.ndim != 1:
        raise DimensionMismatch("b must be a vector")
    x, cols = _run(method, a, b.reshape(-1, 1), precond, tol, max_iter, restart,
                   None if x0 is None else np.asarray(x0).reshape(-1, 1))
    it, fr, conv, stop = cols[0]
    return KrylovResult(x.reshape(-1), it, conv, fr, stop)


def gmres(a: CscMatrix, b, precond: Precond, tol: float, max_iter: int, restart: int,
          x0=None) -> KrylovResult:
    return _single(0, a, b, precond, tol, max_iter, restart, x0)


def cocg(a: CscMatrix, b, precond: Precond, tol: float, max_iter: int) -> KrylovResult:
    return _single(1, a, b, precond, tol, max_iter)


def cocr(a: CscMatrix, b, precond: Precond, tol: float, max_iter: int) -> KrylovResult:
    return _single(2, a, b, precond, tol, max_iter)


def gmres_block(a: CscMatrix, b, s: int, precond: Precond, tol: float, max_iter: int, restart: int,
                x0=None) -> BlockKrylovResult:
    """b: the reference's column-major n*s vector, or an (n, s) array; x comes
    back in the same layout."""
    b = np.asarray(b)
    n = a.n
    flat = b.ndim == 1
    if flat:
        if b.shape[0] != n * s:
            raise DimensionMismatch(f"expected {n * s}, got {b.shape[0]}")
        B = b.reshape(s, n).T
        X0 = None if x0 is None else np.asarray(x0).reshape(s, n).T
    else:
        if b.shape != (n, s):
            raise DimensionMismatch(f"expected ({n}, {s}), got {b.shape}")
        B = b
        X0 = None if x0 is None else np.asarray(x0)
    x, cols = _run(0, a, np.ascontiguousarray(B), precond, tol, max_iter, restart, X0)
    conv = all(c[2] for c in cols)
    out = x.T.reshape(-1) if flat else x
    return BlockKrylovResult(out, max(c[0] for c in cols), conv, [c[1] for c in cols],
                             "converged" if conv else "max_iter")
