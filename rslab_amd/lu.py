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
