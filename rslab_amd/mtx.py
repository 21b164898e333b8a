"""Matrix Market reader for symmetric matrices (src/io/mtx.rs).

`read_mtx` / `parse_mtx` accept `%%MatrixMarket matrix coordinate real
symmetric` (`read_mtx_complex` / `parse_mtx_complex`: `complex symmetric`,
data lines `i j re im`), with the reference parser's rules (mtx.rs:182-360):
the banner compared token by token, case-insensitively; comment and blank
lines skipped before the size line `rows cols nnz` (square only); after it,
blank lines skipped and every other line must be `i j value`; the entry
count must equal the declared nnz; 1-based
indices checked against n; non-finite values rejected; upper-triangle
entries folded to the lower triangle. `MtxMatrix.to_csc` is
`CscMatrix::from_triplets` (duplicates summed, rows sorted per column).
Errors raise IoError, as RslabError::IoError.
"""
from __future__ import annotations

import dataclasses
import math
import os

import numpy as np

from .solver import CscMatrix, IoError


@dataclasses.dataclass
class MtxMatrix:
    """Coordinate entries, 0-based, lower triangle (i >= j) (mtx.rs:8-17)."""

    n: int
    rows: np.ndarray
    cols: np.ndarray
    values: np.ndarray

    def to_csc(self) -> CscMatrix:
        return CscMatrix.from_triplets(self.n, self.rows, self.cols, self.values)


def _parse(contents: str, source: str, field: str, nvt: int) -> MtxMatrix:
    lines = contents.splitlines()
    if not lines:
        raise IoError(f"{source}: empty file")
    banner = ["%%matrixmarket", "matrix", "coordinate", field, "symmetric"]
    toks = lines[0].split()
    if len(toks) != len(banner) or any(t.lower() != b for t, b in zip(toks, banner)):
        raise IoError(f"{source}: unsupported header '{lines[0].strip()}' "
                      f"(expected: %%MatrixMarket matrix coordinate {field} symmetric)")
    k = 1
    while k < len(lines) and (not lines[k].strip() or lines[k].strip().startswith("%")):
        k += 1
    if k >= len(lines):
        raise IoError(f"{source}: missing size line")
    parts = lines[k].split()
    if len(parts) != 3:
        raise IoError(f"{source}: line {k + 1}: expected 'rows cols nnz', got '{lines[k].strip()}'")
    try:
        nr, nc, nnz = (int(p) for p in parts)
    except ValueError:
        raise IoError(f"{source}: line {k + 1}: invalid size line '{lines[k].strip()}'") from None
    if nr != nc or nr < 0:
        raise IoError(f"{source}: symmetric matrix must be square, got {nr}x{nc}")
    n = nr
    rows, cols, vals = [], [], []
    for ln in range(k + 1, len(lines)):
        t = lines[ln].split()
        if not t:
            continue
        if len(t) != 2 + nvt:
            raise IoError(f"{source}: line {ln + 1}: expected 'i j {'re im' if nvt == 2 else 'value'}', "
                          f"got '{lines[ln].strip()}'")
        try:
            i, j = int(t[0]), int(t[1])
        except ValueError:
            raise IoError(f"{source}: line {ln + 1}: invalid index") from None
        try:
            v = complex(float(t[2]), float(t[3])) if nvt == 2 else float(t[2])
        except ValueError:
            raise IoError(f"{source}: line {ln + 1}: invalid value '{' '.join(t[2:])}'") from None
        if not (math.isfinite(v.real) and math.isfinite(v.imag) if nvt == 2 else math.isfinite(v)):
            raise IoError(f"{source}: line {ln + 1}: non-finite value '{' '.join(t[2:])}'")
        if i < 1 or j < 1 or i > n or j > n:
            raise IoError(f"{source}: line {ln + 1}: index ({i}, {j}) out of bounds for {n}x{n} matrix")
        i0, j0 = i - 1, j - 1
        if i0 < j0:
            i0, j0 = j0, i0
        rows.append(i0)
        cols.append(j0)
        vals.append(v)
    if len(rows) != nnz:
        # the size line's nnz is the number of entries that follow (mtx.rs:374-381)
        raise IoError(f"{source}: declared nnz {nnz} does not match the {len(rows)} entries in the file")
    dt = np.complex128 if nvt == 2 else np.float64
    return MtxMatrix(n, np.array(rows, dtype=np.int64), np.array(cols, dtype=np.int64), np.array(vals, dtype=dt))


def parse_mtx(contents: str, source: str = "<string>") -> MtxMatrix:
    """parse_mtx (mtx.rs:119-124): real symmetric."""
    return _parse(contents, source, "real", 1)


def parse_mtx_complex(contents: str, source: str = "<string>") -> MtxMatrix:
    """parse_mtx_complex (mtx.rs:128-138): complex symmetric, `i j re im`."""
    return _parse(contents, source, "complex", 2)


def _read(path) -> str:
    try:
        with open(path) as fh:
            return fh.read()
    except OSError as e:
        raise IoError(f"{os.fspath(path)}: {e}") from None


def read_mtx(path) -> MtxMatrix:
    """read_mtx (mtx.rs:54-58)."""
    return parse_mtx(_read(path), os.fspath(path))


def read_mtx_complex(path) -> MtxMatrix:
    """read_mtx_complex (mtx.rs:64-68)."""
    return parse_mtx_complex(_read(path), os.fspath(path))
