"""The unsymmetric LU twin (src/numeric/multifrontal_lu.rs): GeneralCsc
(src/sparse/general.rs), LuSymbolic (:827-1044), LuSolver (:1046-1253) and
LuFactors (:321-362), over the device LU in librslab_amd.so.

The analysis runs on the symmetrized pattern A + A^T (the LDL^T analysis,
ordering and supernodes); the numeric factor is the supernodal left-looking
LU with threshold partial pivoting restricted to each supernode's own rows,
on the device; the solve runs on the device.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from ._lib import lib, ptr
from .solver import (CscMatrix, DimensionMismatch, InvalidInput, SolverSettings, SymbolicExport, _check)


@dataclasses.dataclass
class GeneralCsc:
    """Full (both triangles) CSC of a general square matrix: column pointers,
    strictly increasing rows within a column (GeneralCsc, src/sparse/general.rs)."""

    n: int
    col_ptr: np.ndarray
    row_idx: np.ndarray
    values: np.ndarray

    def __post_init__(self):
        self.col_ptr = np.ascontiguousarray(self.col_ptr, dtype=np.int64)
        self.row_idx = np.ascontiguousarray(self.row_idx, dtype=np.int64)
        self.values = np.ascontiguousarray(self.values)

    @property
    def is_complex(self) -> bool:
        return np.iscomplexobj(self.values)

    @property
    def nnz(self) -> int:
        return int(self.row_idx.shape[0])

    @staticmethod
    def from_triplets(n, rows, cols, vals) -> "GeneralCsc":
        """Entries anywhere in the matrix; duplicates summed."""
        rows = np.asarray(rows, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        vals = np.asarray(vals)
        if not (rows.shape == cols.shape == vals.shape):
            raise InvalidInput("triplet arrays must have equal length")
        if rows.size and (cols.max() >= n or rows.max() >= n or cols.min() < 0 or rows.min() < 0):
            raise InvalidInput("index out of bounds")
        key = cols * n + rows
        order = np.argsort(key, kind="stable")
        key = key[order]
        uniq, start = np.unique(key, return_index=True)
        summed = np.add.reduceat(vals[order], start) if uniq.size else vals[:0]
        col_ptr = np.zeros(n + 1, dtype=np.int64)
        np.add.at(col_ptr, uniq // n + 1, 1)
        return GeneralCsc(n, np.cumsum(col_ptr), uniq % n, summed)

    @staticmethod
    def from_dense(M) -> "GeneralCsc":
        M = np.asarray(M)
        c, r = np.nonzero(M.T)
        return GeneralCsc.from_triplets(M.shape[0], r, c, M[r, c])

    @staticmethod
    def from_symmetric(a: CscMatrix) -> "GeneralCsc":
        """Both triangles of a symmetric matrix stored as its lower part."""
        cols = np.repeat(np.arange(a.n), np.diff(a.col_ptr))
        off = a.row_idx != cols
        return GeneralCsc.from_triplets(a.n, np.concatenate([a.row_idx, cols[off]]),
                                        np.concatenate([cols, a.row_idx[off]]),
                                        np.concatenate([a.values, a.values[off]]))

    def with_values(self, values) -> "GeneralCsc":
        return GeneralCsc(self.n, self.col_ptr, self.row_idx, values)

    def values_f64(self) -> np.ndarray:
        """Values as the C ABI takes them (complex as interleaved re/im)."""
        if self.is_complex:
            return np.ascontiguousarray(self.values.astype(np.complex128)).view(np.float64)
        return np.ascontiguousarray(self.values, dtype=np.float64)

    def matvec(self, x):
        """y = A x (GeneralCsc::matvec)."""
        x = np.asarray(x)
        cols = np.repeat(np.arange(self.n), np.diff(self.col_ptr))
        vv = self.values.reshape((-1,) + (1,) * (x.ndim - 1))
        y = np.zeros((self.n,) + x.shape[1:], dtype=np.result_type(self.values, x))
        np.add.at(y, self.row_idx, vv * x[cols])
        return y

    def transpose(self) -> "GeneralCsc":
        cols = np.repeat(np.arange(self.n), np.diff(self.col_ptr))
        return GeneralCsc.from_triplets(self.n, cols, self.row_idx, self.values)

    def to_dense(self):
        M = np.zeros((self.n, self.n), dtype=self.values.dtype)
        cols = np.repeat(np.arange(self.n), np.diff(self.col_ptr))
        M[self.row_idx, cols] = self.values
        return M

    def norm_inf(self) -> float:
        r = np.zeros(self.n)
        np.add.at(r, self.row_idx, np.abs(self.values))
        return float(r.max()) if self.n else 0.0


@dataclasses.dataclass
class LuFactors:
    """LuFactors (multifrontal_lu.rs:321-352): L in CSC (unit lower, the
    diagonal first in each column), U in CSR (upper, the pivot first in each
    row), both in factorization order; perm / perm_row map a position to the
    original column / row; the factor is of diag(d_row) A diag(d_col)."""

    n: int
    l_col_ptr: np.ndarray
    l_row_idx: np.ndarray
    l_values: np.ndarray
    u_row_ptr: np.ndarray
    u_col_idx: np.ndarray
    u_values: np.ndarray
    perm: np.ndarray
    perm_row: np.ndarray
    d_row: np.ndarray
    d_col: np.ndarray
    n_perturbed: int

    def factor_nnz(self) -> int:
        """nnz(L) + nnz(U) (LuFactors::factor_nnz, :356-358)."""
        return int(self.l_values.size + self.u_values.size)

    def dense_lu(self):
        n = self.n
        L = np.zeros((n, n), dtype=self.l_values.dtype)
        U = np.zeros((n, n), dtype=self.u_values.dtype)
        for e in range(n):
            s, t = self.l_col_ptr[e], self.l_col_ptr[e + 1]
            L[self.l_row_idx[s:t], e] = self.l_values[s:t]
            s, t = self.u_row_ptr[e], self.u_row_ptr[e + 1]
            U[e, self.u_col_idx[s:t]] = self.u_values[s:t]
        return L, U


class LuSymbolic:
    """LuSymbolic (multifrontal_lu.rs:827-1044): the analysis of the
    symmetrized pattern A + A^T and the device plan; analyse once, factor
    many value sets of the same pattern."""

    def __init__(self, h, n, nnz, settings):
        self._h = h
        self._n = n
        self._nnz = nnz
        self.settings = settings

    @staticmethod
    def analyze(a: GeneralCsc) -> "LuSymbolic":
        return LuSymbolic.analyze_with(a, SolverSettings())

    @staticmethod
    def analyze_with(a: GeneralCsc, opts: SolverSettings) -> "LuSymbolic":
        """LuSymbolic::analyze_with (:850-869)."""
        h = C.c_void_p()
        s = opts.to_c()
        _check(lib.rsl_lu_analyze(a.n, ptr(a.col_ptr, C.c_int64), ptr(a.row_idx, C.c_int64), C.byref(s),
                                  C.byref(h)))
        return LuSymbolic(h, a.n, a.nnz, opts)

    def __del__(self):
        if getattr(self, "_h", None):
            lib.rsl_lu_free_symbolic(self._h)
            self._h = None

    def n(self) -> int:
        return self._n

    def info(self) -> dict:
        out = np.zeros(8, dtype=np.int64)
        fl = np.zeros(1)
        _check(lib.rsl_lu_symbolic_info(self._h, ptr(out, C.c_int64), ptr(fl, C.c_double)))
        keys = ("n", "nnz", "nsuper", "levels", "l_entries", "u_entries", "cmod_tasks", "panel_units")
        d = {k: int(v) for k, v in zip(keys, out)}
        d["factor_flops"] = float(fl[0])
        return d

    def export(self) -> SymbolicExport:
        """The analysis of A + A^T (the layout of LdltSymbolic.export)."""
        inf = self.info()
        n, ns = inf["n"], inf["nsuper"]
        arr = lambda k, dt=np.int32: np.zeros(max(k, 1), dtype=dt)  # noqa: E731
        e = SymbolicExport(n, arr(n), arr(n), arr(ns), arr(ns), arr(ns), arr(ns + 1), arr(ns),
                           arr(n + 1, np.int64), arr(2 * self._nnz + n))
        _check(lib.rsl_lu_symbolic_export(
            self._h, ptr(e.perm, C.c_int32), ptr(e.perm_inv, C.c_int32), ptr(e.sn_first, C.c_int32),
            ptr(e.sn_ncol, C.c_int32), ptr(e.sn_nrow, C.c_int32), ptr(e.ch_ptr, C.c_int32),
            ptr(e.ch_idx, C.c_int32), ptr(e.pp_colp, C.c_int64), ptr(e.pp_rowi, C.c_int32)))
        e.perm, e.perm_inv = e.perm[:n], e.perm_inv[:n]
        e.sn_first, e.sn_ncol, e.sn_nrow = e.sn_first[:ns], e.sn_ncol[:ns], e.sn_nrow[:ns]
        e.ch_ptr = e.ch_ptr[:ns + 1]
        e.ch_idx = e.ch_idx[:max(int(e.ch_ptr[-1]), 1)] if ns else e.ch_idx
        e.pp_rowi = e.pp_rowi[:max(int(e.pp_colp[-1]), 1)]
        return e

    def factor(self, a: GeneralCsc, opts: SolverSettings = None) -> "LuSolver":
        """LuSymbolic::factor (:880-907): equilibrate, then the device LU."""
        opts = opts or self.settings
        if a.n != self._n or a.nnz != self._nnz:
            raise InvalidInput("factor_general_lu_numeric: matrix does not match the analyzed pattern")
        h = C.c_void_p()
        s = opts.to_c()
        vals = a.values_f64()
        _check(lib.rsl_lu_factor(self._h, 1 if a.is_complex else 0, ptr(vals, C.c_double), C.byref(s),
                                 C.byref(h)))
        return LuSolver(h, self, a.n, a.is_complex)


class LuSolver:
    """LuSolver<T> (multifrontal_lu.rs:1046-1253) over the device factor."""

    def __init__(self, h, sym: LuSymbolic, n: int, cplx: bool):
        self._h = h
        self._sym = sym  # the symbolic must outlive the factor
        self._n = n
        self._cplx = cplx

    def __del__(self):
        if getattr(self, "_h", None):
            lib.rsl_lu_free_factor(self._h)
            self._h = None

    @staticmethod
    def factor(a: GeneralCsc, opts: SolverSettings = None) -> "LuSolver":
        """One-shot analyse + equilibrate + factor (:1061-1066)."""
        opts = opts or SolverSettings()
        return LuSymbolic.analyze_with(a, opts).factor(a, opts)

    def n(self) -> int:
        return self._n

    def refactor(self, a: GeneralCsc):
        vals = a.values_f64()
        _check(lib.rsl_lu_refactor(self._h, ptr(vals, C.c_double)))

    def _dtype(self):
        return np.complex128 if self._cplx else np.float64

    def solve_many(self, b):
        """solve_lu_many (:3158-3207): b row-major n x nrhs (or a vector)."""
        b = np.asarray(b)
        single = b.ndim == 1
        B = np.ascontiguousarray(b.reshape(self._n, -1), dtype=self._dtype())
        if B.shape[0] != self._n:
            raise DimensionMismatch(f"expected {self._n} rows, got {b.shape[0]}")
        X = np.zeros_like(B)
        bv = B.view(np.float64) if self._cplx else B
        xv = X.view(np.float64) if self._cplx else X
        _check(lib.rsl_lu_solve(self._h, ptr(bv, C.c_double), B.shape[1], ptr(xv, C.c_double)))
        return X.reshape(self._n) if single else X

    def solve(self, b):
        """LuSolver::solve (:1205-1210)."""
        return self.solve_many(b)

    def solve_device(self, d_b: int, nrhs: int, d_x: int, stream: int = 0):
        """solve_lu_many on device buffers: d_b, d_x are device addresses of
        row-major n x nrhs arrays; ordered after the work on `stream` (0 = the
        null stream), which then waits for the solve."""
        if nrhs <= 0:
            raise DimensionMismatch("nrhs must be positive")
        _check(lib.rsl_lu_solve_device(self._h, C.c_void_p(d_b), int(nrhs), C.c_void_p(d_x),
                                       C.c_void_p(stream) if stream else None))

    def solve_refined(self, a: GeneralCsc, b, max_iter: int, info: bool = False):
        """solve_lu_refined (:3285-3322) on the device, per column: x =
        (LU)^-1 b, then up to max_iter corrections x += (LU)^-1 (b - A x),
        every correction evaluated, keeping each column's iterate of least
        max |b - A x|. info=True also returns per column (corrections
        applied, that least max |r|)."""
        if a.n != self._n:
            raise DimensionMismatch(f"expected a matrix of dimension {self._n}, got {a.n}")
        b = np.asarray(b)
        single = b.ndim == 1
        B = np.ascontiguousarray(b.reshape(self._n, -1), dtype=self._dtype())
        X = np.zeros_like(B)
        inf = np.zeros(2 * B.shape[1])
        bv = B.view(np.float64) if self._cplx else B
        xv = X.view(np.float64) if self._cplx else X
        vals = a.values_f64()
        _check(lib.rsl_lu_refine(self._h, ptr(vals, C.c_double), ptr(bv, C.c_double), B.shape[1], int(max_iter),
                                 ptr(xv, C.c_double), ptr(inf, C.c_double)))
        X = X.reshape(self._n) if single else X
        if info:
            return X, [(int(inf[2 * c]), float(inf[2 * c + 1])) for c in range(B.shape[1])]
        return X

    def info(self) -> dict:
        out = np.zeros(6, dtype=np.int64)
        t = np.zeros(1)
        _check(lib.rsl_lu_info(self._h, ptr(out, C.c_int64), ptr(t, C.c_double)))
        return {"n": int(out[0]), "n_perturbed": int(out[1]), "nnz_l": int(out[2]), "nnz_u": int(out[3]),
                "factor_ms": float(t[0])}

    def n_perturbed(self) -> int:
        return self.info()["n_perturbed"]

    def factor_nnz(self) -> int:
        i = self.info()
        return i["nnz_l"] + i["nnz_u"]

    def factors(self) -> LuFactors:
        i = self.info()
        n, nl, nu = i["n"], i["nnz_l"], i["nnz_u"]
        nd = 2 if self._cplx else 1
        lcp, ucp = np.zeros(n + 1, np.int64), np.zeros(n + 1, np.int64)
        lri, uci = np.zeros(max(nl, 1), np.int32), np.zeros(max(nu, 1), np.int32)
        lv, uv = np.zeros(max(nl * nd, 1)), np.zeros(max(nu * nd, 1))
        pm, pr = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
        dr, dc = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        _check(lib.rsl_lu_factor_download(self._h, ptr(lcp, C.c_int64), ptr(lri, C.c_int32), ptr(lv, C.c_double),
                                          ptr(ucp, C.c_int64), ptr(uci, C.c_int32), ptr(uv, C.c_double),
                                          ptr(pm, C.c_int32), ptr(pr, C.c_int32), ptr(dr, C.c_double),
                                          ptr(dc, C.c_double)))
        lv, uv = lv[:nl * nd], uv[:nu * nd]
        if self._cplx:
            lv, uv = lv.view(np.complex128), uv.view(np.complex128)
        return LuFactors(n, lcp, lri[:nl], lv, ucp, uci[:nu], uv, pm[:n], pr[:n], dr[:n], dc[:n], i["n_perturbed"])


def factor_general_lu(a: GeneralCsc, opts: SolverSettings = None) -> LuFactors:
    """factor_general_lu (:1255-1261): analyse + factor, the raw factors."""
    return LuSolver.factor(a, opts).factors()
