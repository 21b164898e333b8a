"""Host-side mirror of rslab's symmetric solver surface over the C ABI.

Names, argument meaning and error behaviour follow the Rust API:

* ``CscMatrix``            src/sparse/csc.rs:15-20 (lower triangle, rows sorted)
* ``SolverSettings``       src/numeric/multifrontal_ldlt.rs:302-394 (+ builders :616-779)
* ``LdltSymbolic``         src/numeric/sparse_solver.rs:479-712 (analyze / factor)
* ``LdltSolver``           src/numeric/sparse_solver.rs:40-376 (solve / solve_many / inertia)
* ``LdltFactors``          src/dense/ldlt_generic.rs:31-66 (downloaded CSC factor)
* ``MixedLdltSolver``      src/numeric/mixed.rs:225-291 (IR -> GMRES-IR ladder)
* ``RslabError`` family    src/error.rs:3-43

Every numeric call runs on the GPU through librslab_amd.so; nothing here
computes a factor or a solve on the CPU.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import enum
from typing import Optional

import numpy as np

from . import _lib
from ._lib import lib, ptr


# ---------------------------------------------------------------- errors
class RslabError(Exception):
    """Base error (RslabError)."""


class NumericallyRankDeficient(RslabError):
    pass


class InvalidInput(RslabError):
    pass


class DimensionMismatch(RslabError):
    pass


class DeviceError(InvalidInput):
    """rslab has no device variant; it reports InvalidInput("device: ...")."""


class OutOfMemory(RslabError):
    pass


class IoError(RslabError):
    """RslabError::IoError (error.rs:16): unreadable or malformed input files."""


_ERR = {
    _lib.RSL_NUMERICALLY_RANK_DEFICIENT: NumericallyRankDeficient,
    _lib.RSL_INVALID_INPUT: InvalidInput,
    _lib.RSL_DIMENSION_MISMATCH: DimensionMismatch,
    _lib.RSL_DEVICE_ERROR: DeviceError,
    _lib.RSL_OUT_OF_MEMORY: OutOfMemory,
}


def _check(rc: int):
    if rc != _lib.RSL_OK:
        msg = lib.rsl_last_error().decode(errors="replace")
        raise _ERR.get(rc, RslabError)(msg)


# ---------------------------------------------------------------- matrix
@dataclasses.dataclass
class CscMatrix:
    """Lower-triangle CSC of a symmetric (or complex-symmetric) matrix."""

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
    def from_triplets(n, rows, cols, vals) -> "CscMatrix":
        """CscMatrix::from_triplets (csc.rs:41-106): lower entries, duplicates summed."""
        rows = np.asarray(rows, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        vals = np.asarray(vals)
        if not (rows.shape == cols.shape == vals.shape):
            raise InvalidInput("triplet arrays must have equal length")
        if rows.size and (cols.max() >= n or rows.max() >= n or cols.min() < 0 or rows.min() < 0):
            raise InvalidInput("index out of bounds")
        if np.any(rows < cols):
            raise InvalidInput("triplet is upper-triangle; CscMatrix stores only the lower triangle")
        key = cols * n + rows
        order = np.argsort(key, kind="stable")
        key = key[order]
        uniq, start = np.unique(key, return_index=True)
        summed = np.add.reduceat(vals[order], start) if uniq.size else vals[:0]
        c = uniq // n
        r = uniq % n
        col_ptr = np.zeros(n + 1, dtype=np.int64)
        np.add.at(col_ptr, c + 1, 1)
        return CscMatrix(n, np.cumsum(col_ptr), r, summed)

    @staticmethod
    def block_diag(mats) -> "CscMatrix":
        """diag(A_0, A_1, ...) in lower CSC (the layout rsl_analyze_batch uses)."""
        n = sum(m.n for m in mats)
        cps, ris, vs = [np.zeros(1, dtype=np.int64)], [], []
        cbase, rbase = 0, 0
        for m in mats:
            cps.append(m.col_ptr[1:] + cbase)
            ris.append(m.row_idx + rbase)
            vs.append(m.values)
            cbase += m.nnz
            rbase += m.n
        return CscMatrix(n, np.concatenate(cps), np.concatenate(ris), np.concatenate(vs))

    def with_values(self, values) -> "CscMatrix":
        return CscMatrix(self.n, self.col_ptr, self.row_idx, values)

    def values_f64(self) -> np.ndarray:
        """Values as the C ABI takes them (complex as interleaved re/im)."""
        if self.is_complex:
            return np.ascontiguousarray(self.values.astype(np.complex128)).view(np.float64)
        return np.ascontiguousarray(self.values, dtype=np.float64)

    def symv(self, x):
        """CscMatrix::symv (csc.rs:294-309): y = A x from the stored lower part."""
        cols = np.repeat(np.arange(self.n), np.diff(self.col_ptr))
        rows = self.row_idx
        v = self.values
        x = np.asarray(x)
        y = np.zeros((self.n,) + x.shape[1:], dtype=np.result_type(v, x))
        vv = v.reshape((-1,) + (1,) * (x.ndim - 1))
        np.add.at(y, rows, vv * x[cols])
        off = rows != cols
        np.add.at(y, cols[off], vv[off] * x[rows[off]])
        return y

    def norm_inf(self) -> float:
        """sym_norm_inf (mixed.rs:358-371)."""
        cols = np.repeat(np.arange(self.n), np.diff(self.col_ptr))
        m = np.abs(self.values)
        rs = np.zeros(self.n)
        np.add.at(rs, self.row_idx, m)
        off = self.row_idx != cols
        np.add.at(rs, cols[off], m[off])
        return float(rs.max()) if self.n else 0.0


# ---------------------------------------------------------------- settings
class OrderingMethod(enum.IntEnum):
    Amd = 0
    MetisND = 1
    Natural = 2


class FactorMethod(enum.IntEnum):
    """FactorMethod (multifrontal_ldlt.rs:257-295): both schedules give the same
    factor. LeftLooking pulls each level's updates before its cdiv;
    RightLooking pushes a level's updates into its ancestors right after its
    cdiv. Multifrontal is named for API parity and rejected by the device path."""

    LeftLooking = 0
    RightLooking = 1
    Multifrontal = 2


class Precision(enum.IntEnum):
    """Precision of the Schur / contribution GEMMs; pivots, D and residuals stay FP64."""

    FP64 = 0
    FP32_SCHUR = 1
    BF16_SCHUR = 2


@dataclasses.dataclass
class ZeroPivotAction:
    """ZeroPivotAction (multifrontal_ldlt.rs:165-180)."""

    kind: int = 0  # 0 Fail, 1 PerturbToEps, 2 ForceAccept
    abs_floor: float = 0.0

    @staticmethod
    def Fail():
        return ZeroPivotAction(0, 0.0)

    @staticmethod
    def PerturbToEps(abs_floor: float):
        return ZeroPivotAction(1, float(abs_floor))

    @staticmethod
    def ForceAccept():
        return ZeroPivotAction(2, 0.0)


@dataclasses.dataclass
class SolverSettings:
    """SolverSettings with the defaults of multifrontal_ldlt.rs:583-614."""

    ordering: OrderingMethod = OrderingMethod.Amd
    nemin: int = 16
    relax: Optional[tuple] = (256, 64)
    panel_nb: int = 64
    on_zero_pivot: ZeroPivotAction = dataclasses.field(default_factory=ZeroPivotAction.Fail)
    scaling: str = "OnePassInfNorm"
    precision: Precision = Precision.FP64
    nd_seed: int = 1
    wide_flops: float = 2.5e5
    drop_tol: Optional[float] = None
    # cmod product-buffer cap in elements (None: automatic); levels needing
    # more take the gathered-tile cmod, 0 forces it everywhere
    cmod_cap: Optional[int] = None
    # same-pattern refactors replay the factor's launch sequence as a HIP graph
    graphs: bool = True
    # own blocks wider than this are solved in multi-workgroup steps (tuning knob)
    solve_wide: int = 32
    method: FactorMethod = FactorMethod.LeftLooking
    # solves of up to 8 right-hand sides run each sweep as one launch ordered
    # by device counters (False, and wider solves: one launch per level and
    # block step; tuning knob)
    solve_dep: bool = True
    # first tree level of the dependency-driven launch (None: automatic)
    solve_split: Optional[int] = None
    # LU twin's threshold partial-pivoting tolerance u (multifrontal_ldlt.rs:382)
    pivot_u: float = 0.1

    # builders (multifrontal_ldlt.rs:616-779)
    def with_method(self, m):
        """SolverSettings::with_method (multifrontal_ldlt.rs:257-295)."""
        return dataclasses.replace(self, method=FactorMethod(m))

    def with_graphs(self, on: bool):
        return dataclasses.replace(self, graphs=bool(on))

    def with_pivot_u(self, u: float):
        """SolverSettings::with_pivot_u (multifrontal_ldlt.rs:742-745): clamped to [0, 1]."""
        return dataclasses.replace(self, pivot_u=min(1.0, max(0.0, float(u))))

    def with_drop_tol(self, tau: float):
        """Incomplete-factor threshold dropping (with_drop_tol, multifrontal_ldlt.rs:640-643):
        stored L entries below tau times the largest in their column are dropped."""
        return dataclasses.replace(self, drop_tol=float(tau))

    def with_cmod_cap(self, elems: Optional[int]):
        return dataclasses.replace(self, cmod_cap=None if elems is None else int(elems))

    def with_ordering(self, o):
        return dataclasses.replace(self, ordering=OrderingMethod(o))

    def with_nemin(self, n):
        return dataclasses.replace(self, nemin=int(n))

    def with_relax(self, r):
        return dataclasses.replace(self, relax=r)

    def with_panel_nb(self, nb):
        return dataclasses.replace(self, panel_nb=int(nb))

    def with_pivot(self, p: ZeroPivotAction):
        return dataclasses.replace(self, on_zero_pivot=p)

    def with_scaling(self, s: str):
        return dataclasses.replace(self, scaling=s)

    def with_precision(self, p):
        return dataclasses.replace(self, precision=Precision(p))

    @staticmethod
    def preconditioner(abs_floor: float) -> "SolverSettings":
        return SolverSettings(on_zero_pivot=ZeroPivotAction.PerturbToEps(abs_floor))

    def to_c(self) -> _lib.Settings:
        s = _lib.Settings()
        lib.rsl_settings_default(C.byref(s))
        s.ordering = int(self.ordering)
        s.nemin = int(self.nemin)
        if self.relax is None:
            s.relax = 0
        else:
            s.relax = 1
            s.relax_width, s.relax_rows = int(self.relax[0]), int(self.relax[1])
        s.panel_nb = int(self.panel_nb)
        s.pivot = int(self.on_zero_pivot.kind)
        s.abs_floor = float(self.on_zero_pivot.abs_floor)
        if self.scaling not in ("OnePassInfNorm", "Identity"):
            raise InvalidInput(f"scaling strategy {self.scaling!r} is not supported by the device path")
        s.scaling = 1 if self.scaling == "OnePassInfNorm" else 0
        s.precision = int(self.precision)
        s.nd_seed = int(self.nd_seed)
        s.wide_flops = float(self.wide_flops)
        if self.drop_tol is not None and not (self.drop_tol >= 0.0):
            raise InvalidInput(f"drop_tol must be a non-negative number, got {self.drop_tol!r}")
        s.drop_tol = -1.0 if self.drop_tol is None else float(self.drop_tol)
        s.cmod_cap = -1 if self.cmod_cap is None else int(self.cmod_cap)
        s.graphs = 1 if self.graphs else 0
        s.solve_wide = int(self.solve_wide)
        if FactorMethod(self.method) == FactorMethod.Multifrontal:
            raise InvalidInput("FactorMethod.Multifrontal is not provided; use LeftLooking or RightLooking")
        s.method = int(self.method)
        s.solve_dep = 1 if self.solve_dep else 0
        s.solve_split = -1 if self.solve_split is None else int(self.solve_split)
        s.pivot_u = min(1.0, max(0.0, float(self.pivot_u)))
        return s


# ---------------------------------------------------------------- results
@dataclasses.dataclass
class Inertia:
    positive: int
    negative: int
    zero: int

    def total(self) -> int:
        return self.positive + self.negative + self.zero

    def as_tuple(self):
        return (self.positive, self.negative, self.zero)


@dataclasses.dataclass
class LdltFactors:
    """LdltFactors (ldlt_generic.rs:31-66), elimination order."""

    n: int
    l_col_ptr: np.ndarray
    l_row_idx: np.ndarray
    l_values: np.ndarray
    d_diag: np.ndarray
    d_subdiag: np.ndarray
    two_by_two: np.ndarray
    perm: np.ndarray
    n_perturbed: int
    inertia: Inertia

    def save(self, path):
        """The arrays of LdltFactors (ldlt_generic.rs:31-66) in one .npz
        (plain arrays, no pickled objects)."""
        np.savez(path, n=np.int64(self.n), l_col_ptr=self.l_col_ptr, l_row_idx=self.l_row_idx,
                 l_values=self.l_values, d_diag=self.d_diag, d_subdiag=self.d_subdiag,
                 two_by_two=self.two_by_two, perm=self.perm, n_perturbed=np.int64(self.n_perturbed),
                 inertia=np.array(self.inertia.as_tuple(), dtype=np.int64))

    @staticmethod
    def load(path) -> "LdltFactors":
        with np.load(path, allow_pickle=False) as z:
            return LdltFactors(int(z["n"]), z["l_col_ptr"], z["l_row_idx"], z["l_values"], z["d_diag"],
                               z["d_subdiag"], z["two_by_two"], z["perm"], int(z["n_perturbed"]),
                               Inertia(*map(int, z["inertia"])))


@dataclasses.dataclass
class MixedInfo:
    """MixedInfo (mixed.rs:81-92)."""

    ir_iters: int
    gmres_iters: int
    backward_error: float
    certified: bool


@dataclasses.dataclass
class SymbolicExport:
    n: int
    perm: np.ndarray
    perm_inv: np.ndarray
    sn_first: np.ndarray
    sn_ncol: np.ndarray
    sn_nrow: np.ndarray
    ch_ptr: np.ndarray
    ch_idx: np.ndarray
    pp_colp: np.ndarray
    pp_rowi: np.ndarray

    def sn_parent(self) -> np.ndarray:
        """Supernode tree parent (-1 for roots) from the children lists."""
        par = np.full(len(self.sn_first), -1, dtype=np.int32)
        for s in range(len(self.sn_first)):
            par[self.ch_idx[self.ch_ptr[s]:self.ch_ptr[s + 1]]] = s
        return par


# ---------------------------------------------------------------- symbolic
class LdltSymbolic:
    """Phase-1 analysis reused across numeric factorizations."""

    def __init__(self, handle, n: int, nnz: int, settings: SolverSettings):
        self._h = handle
        self._n = n
        self._nnz = nnz
        self.settings = settings

    @staticmethod
    def analyze(a: CscMatrix) -> "LdltSymbolic":
        return LdltSymbolic.analyze_with(a, SolverSettings())

    @staticmethod
    def analyze_with(a: CscMatrix, opts: SolverSettings) -> "LdltSymbolic":
        h = C.c_void_p()
        s = opts.to_c()
        _check(lib.rsl_analyze(a.n, ptr(a.col_ptr, C.c_int64), ptr(a.row_idx, C.c_int64),
                               C.byref(s), C.byref(h)))
        return LdltSymbolic(h, a.n, a.nnz, opts)

    @staticmethod
    def from_symbolic(a: CscMatrix, perm, sn_first, sn_ncol, sn_parent,
                      opts: Optional[SolverSettings] = None) -> "LdltSymbolic":
        """LdltSymbolic from the caller's own analysis (what factor_numeric
        takes, multifrontal_ldlt.rs:1851-1855): permutation (new -> old), the
        supernode partition of the permuted columns in postorder and its tree
        (parent a later supernode, or -1). Raises InvalidInput when they do not
        form a valid supernodal elimination structure for a's pattern."""
        opts = opts or SolverSettings()
        ar = lambda v: np.ascontiguousarray(v, dtype=np.int32)  # noqa: E731
        perm, sn_first, sn_ncol, sn_parent = ar(perm), ar(sn_first), ar(sn_ncol), ar(sn_parent)
        if not (sn_first.size == sn_ncol.size == sn_parent.size):
            raise InvalidInput("supernode arrays must have equal length")
        if perm.size != a.n:
            raise InvalidInput(f"perm must hold {a.n} entries, got {perm.size}")
        h = C.c_void_p()
        s = opts.to_c()
        _check(lib.rsl_plan_create(a.n, ptr(a.col_ptr, C.c_int64), ptr(a.row_idx, C.c_int64),
                                   ptr(perm, C.c_int32), int(sn_first.size), ptr(sn_first, C.c_int32),
                                   ptr(sn_ncol, C.c_int32), ptr(sn_parent, C.c_int32), C.byref(s), C.byref(h)))
        return LdltSymbolic(h, a.n, a.nnz, opts)

    def replan(self, opts: SolverSettings) -> "LdltSymbolic":
        """This analysis with a device plan built for `opts`' plan settings
        (panel_nb, wide_flops, solve_wide); the ordering is not repeated."""
        h = C.c_void_p()
        s = opts.to_c()
        _check(lib.rsl_symbolic_replan(self._h, C.byref(s), C.byref(h)))
        out = LdltSymbolic(h, self._n, self._nnz, opts)
        if hasattr(self, "batch"):
            out.batch = self.batch
        return out

    @staticmethod
    def analyze_batch(mats, opts: Optional[SolverSettings] = None) -> "LdltSymbolic":
        """Independent systems analysed one by one and laid out as one
        block-diagonal system; factor / solve / refine then take the
        concatenated values and right-hand sides (`CscMatrix.block_diag`)."""
        opts = opts or SolverSettings()
        cnt = len(mats)
        ns = np.array([m.n for m in mats], dtype=np.int64)
        keep = [(np.ascontiguousarray(m.col_ptr, dtype=np.int64), np.ascontiguousarray(m.row_idx, dtype=np.int64))
                for m in mats]
        cps = (C.POINTER(C.c_int64) * cnt)(*[ptr(k[0], C.c_int64) for k in keep])
        ris = (C.POINTER(C.c_int64) * cnt)(*[ptr(k[1], C.c_int64) for k in keep])
        h = C.c_void_p()
        s = opts.to_c()
        _check(lib.rsl_analyze_batch(cnt, ptr(ns, C.c_int64), cps, ris, C.byref(s), C.byref(h)))
        sym = LdltSymbolic(h, int(ns.sum()), int(sum(m.nnz for m in mats)), opts)
        sym.batch = cnt
        return sym

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.rsl_free_symbolic(h)
            self._h = None

    def n(self) -> int:
        return self._n

    def info(self) -> dict:
        out = np.zeros(10, dtype=np.int64)
        fl = np.zeros(2, dtype=np.float64)
        _check(lib.rsl_symbolic_info(self._h, ptr(out, C.c_int64), ptr(fl, C.c_double)))
        return dict(n=int(out[0]), nsuper=int(out[1]), colcount_nnz=int(out[2]),
                    panel_entries=int(out[3]), levels=int(out[4]), children=int(out[5]),
                    pattern_nnz=int(out[6]), nnzl_padded=int(out[7]), cmod_elems_max=int(out[8]),
                    solve_scratch_rows=int(out[9]), factor_flops=float(fl[0]), geom_flops=float(fl[1]))

    def cmod_elems_max(self) -> int:
        """Product-buffer elements the largest level's two-phase cmod needs."""
        return self.info()["cmod_elems_max"]

    def export(self) -> SymbolicExport:
        inf = self.info()
        n, ns = inf["n"], inf["nsuper"]
        arr = lambda k, dt=np.int32: np.zeros(k, dtype=dt)
        e = SymbolicExport(n, arr(n), arr(n), arr(ns), arr(ns), arr(ns), arr(ns + 1),
                           arr(max(inf["children"], 1)), arr(n + 1, np.int64),
                           arr(max(inf["pattern_nnz"], 1)))
        _check(lib.rsl_symbolic_export(
            self._h, ptr(e.perm, C.c_int32), ptr(e.perm_inv, C.c_int32), ptr(e.sn_first, C.c_int32),
            ptr(e.sn_ncol, C.c_int32), ptr(e.sn_nrow, C.c_int32), ptr(e.ch_ptr, C.c_int32),
            ptr(e.ch_idx, C.c_int32), ptr(e.pp_colp, C.c_int64), ptr(e.pp_rowi, C.c_int32)))
        return e

    def factor(self, a: CscMatrix, opts: Optional[SolverSettings] = None) -> "LdltSolver":
        """LdltSymbolic::factor: equilibrate, then the device numeric factor."""
        opts = opts or self.settings
        if a.n != self._n or a.nnz != self._nnz:
            raise InvalidInput("factor_numeric: matrix does not match the analyzed pattern")
        field = 1 if a.is_complex else 0
        h = C.c_void_p()
        s = opts.to_c()
        vals = a.values_f64()
        _check(lib.rsl_factor(self._h, field, ptr(vals, C.c_double), C.byref(s), C.byref(h)))
        return LdltSolver(h, self, a.n, a.is_complex)

    def load_factor(self, path) -> "LdltSolver":
        """A factor written by LdltSolver.save for a symbolic of the same plan
        (same pattern and analysis), without a numeric factorization."""
        import os
        p = os.fsencode(path)
        try:
            with open(p, "rb") as fh:
                head = fh.read(20)
        except OSError as e:
            raise InvalidInput(f"factor file: {e}") from None
        if len(head) < 20 or head[:8] != b"RSLFACT1":
            raise InvalidInput("factor file: not a factor file")
        cplx = int.from_bytes(head[16:20], "little") == 1
        h = C.c_void_p()
        _check(lib.rsl_factor_load(self._h, p, C.byref(h)))
        return LdltSolver(h, self, self._n, cplx)


# ---------------------------------------------------------------- solver
class LdltSolver:
    """A factored matrix on the GPU (drop-in for LdltSolver<T>)."""

    def __init__(self, handle, sym: LdltSymbolic, n: int, cplx: bool):
        self._h = handle
        self._sym = sym  # keeps the plan alive
        self._n = n
        self._cplx = cplx

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.rsl_free_factor(h)
            self._h = None

    @staticmethod
    def factor_with(a: CscMatrix, opts: Optional[SolverSettings] = None) -> "LdltSolver":
        opts = opts or SolverSettings()
        return LdltSymbolic.analyze_with(a, opts).factor(a, opts)

    def n(self) -> int:
        return self._n

    def save(self, path):
        """Write this factor (everything a solve reads) for LdltSymbolic.load_factor."""
        import os
        _check(lib.rsl_factor_save(self._h, os.fsencode(path)))

    def refactor(self, a: CscMatrix):
        """Same-pattern numeric refactorization (plan + device storage reused)."""
        vals = a.values_f64()
        _check(lib.rsl_refactor(self._h, ptr(vals, C.c_double)))

    def _dtype(self):
        return np.complex128 if self._cplx else np.float64

    def _as_c(self, b):
        b = np.ascontiguousarray(np.asarray(b, dtype=self._dtype()))
        return b, (b.view(np.float64) if self._cplx else b)

    def solve(self, rhs) -> np.ndarray:
        rhs = np.asarray(rhs)
        if rhs.ndim != 1 or rhs.shape[0] != self._n:
            raise DimensionMismatch(f"expected {self._n}, got {rhs.shape}")
        return self.solve_many(rhs.reshape(self._n, 1), 1).reshape(self._n)

    def solve_many(self, b, nrhs: Optional[int] = None) -> np.ndarray:
        """b row-major n x nrhs (solve_many, sparse_solver.rs:302-327)."""
        b = np.asarray(b)
        if b.ndim == 1:
            if nrhs is None or b.shape[0] != self._n * nrhs:
                raise DimensionMismatch(f"expected {self._n * (nrhs or 1)}, got {b.shape[0]}")
            b = b.reshape(self._n, nrhs)
        if b.shape[0] != self._n or (nrhs is not None and b.shape[1] != nrhs) or b.shape[1] == 0:
            raise DimensionMismatch(f"expected {self._n} x {nrhs}, got {b.shape}")
        k = b.shape[1]
        bb, bv = self._as_c(b)
        x = np.zeros_like(bb)
        xv = x.view(np.float64) if self._cplx else x
        _check(lib.rsl_solve(self._h, ptr(bv, C.c_double), k, ptr(xv, C.c_double)))
        return x

    def refactor_device(self, d_values: int, stream: int = 0):
        """Same-pattern refactorization from lower-CSC values already in
        device memory (d_values is their device address; interleaved for the
        complex field), ordered
        after the work enqueued on `stream` (0 = the null stream)."""
        _check(lib.rsl_refactor_device(self._h, C.c_void_p(d_values), C.c_void_p(stream) if stream else None))

    def solve_device(self, d_b: int, nrhs: int, d_x: int, stream: int = 0):
        """Device-resident solve: d_b, d_x are device addresses of row-major
        n x nrhs arrays; enqueued on `stream` (0 = the null stream)."""
        if nrhs <= 0:
            raise DimensionMismatch("nrhs must be positive")
        _check(lib.rsl_solve_device(self._h, C.c_void_p(d_b), int(nrhs), C.c_void_p(d_x),
                                    C.c_void_p(stream) if stream else None))

    def solve_device_noalloc(self, d_b: int, nrhs: int, d_x: int, stream: int = 0):
        """solve_device without allocating: the scratch is the library's fixed
        device array (RSL_SOLVE_NOALLOC_SCRATCH_BYTES per GPU); a solve needing
        more raises InvalidInput. Overlapping calls on one GPU race."""
        if nrhs <= 0:
            raise DimensionMismatch("nrhs must be positive")
        _check(lib.rsl_solve_device_noalloc(self._h, C.c_void_p(d_b), int(nrhs), C.c_void_p(d_x),
                                            C.c_void_p(stream) if stream else None))

    def profile(self, on: bool = True):
        """Enable (and clear) or disable per-kernel event timing."""
        _check(lib.rsl_profile_enable(self._h, 1 if on else 0))

    def profile_read(self) -> dict:
        """{kind: {ms, flops, bytes, launches}} for every launch kind seen."""
        out = {}
        k = 0
        while True:
            name = lib.rsl_profile_kind_name(k)
            if name is None:
                break
            v = np.zeros(4)
            _check(lib.rsl_profile_read(self._h, k, ptr(v, C.c_double)))
            if v[3] > 0:
                out[name.decode()] = dict(ms=float(v[0]), flops=float(v[1]), bytes=float(v[2]),
                                          launches=int(v[3]))
            k += 1
        return out

    def inertia(self) -> Inertia:
        i = np.zeros(3, dtype=np.int64)
        _check(lib.rsl_factor_info(self._h, ptr(i, C.c_int64), None, None))
        return Inertia(int(i[0]), int(i[1]), int(i[2]))

    def inertia_batch(self, count: int):
        """Per-system inertia of a factor of a batch (rsl_analyze_batch)."""
        out = np.zeros(3 * count, dtype=np.int64)
        _check(lib.rsl_factor_inertia_batch(self._h, count, ptr(out, C.c_int64)))
        return [Inertia(*map(int, out[3 * q:3 * q + 3])) for q in range(count)]

    def n_perturbed(self) -> int:
        v = C.c_int64()
        _check(lib.rsl_factor_info(self._h, None, C.byref(v), None))
        return int(v.value)

    def factor_time_ms(self) -> float:
        v = C.c_double()
        _check(lib.rsl_factor_info(self._h, None, None, C.byref(v)))
        return float(v.value)

    def factor_nnz(self) -> int:
        v = C.c_int64()
        _check(lib.rsl_factor_nnz(self._h, C.byref(v)))
        return int(v.value)

    def factors(self) -> LdltFactors:
        nnz = self.factor_nnz()
        n = self._n
        nd = 2 if self._cplx else 1
        colp = np.zeros(n + 1, dtype=np.int64)
        rowi = np.zeros(max(nnz, 1), dtype=np.int32)
        lv = np.zeros(max(nnz, 1) * nd, dtype=np.float64)
        dd = np.zeros(max(n, 1) * nd, dtype=np.float64)
        ds = np.zeros(max(n, 1) * nd, dtype=np.float64)
        tw = np.zeros(max(n, 1), dtype=np.uint8)
        pm = np.zeros(max(n, 1), dtype=np.int32)
        _check(lib.rsl_factor_download(self._h, ptr(colp, C.c_int64), ptr(rowi, C.c_int32),
                                       ptr(lv, C.c_double), ptr(dd, C.c_double), ptr(ds, C.c_double),
                                       ptr(tw, C.c_uint8), ptr(pm, C.c_int32)))
        cv = (lambda a: a.view(np.complex128)) if self._cplx else (lambda a: a)
        return LdltFactors(n, colp, rowi[:nnz], cv(lv)[:nnz], cv(dd)[:n], cv(ds)[:n],
                           tw[:n].astype(bool), pm[:n], self.n_perturbed(), self.inertia())


# ---------------------------------------------------------------- mixed
CERT_FACTOR = 64.0  # mixed.rs:97-105
EPS_HI = float(np.finfo(np.float64).eps)


class MixedLdltSolver:
    """Reduced-precision Schur factor refined to FP64 against the original A."""

    def __init__(self, lo: LdltSolver, a: CscMatrix):
        self.lo = lo
        self._a = a
        self.anorm_inf = a.norm_inf()

    @staticmethod
    def factor_with(a: CscMatrix, opts: Optional[SolverSettings] = None) -> "MixedLdltSolver":
        """Factor with reduced-precision Schur products (FP32 unless `opts`
        names another precision; FP64 gives the plain factor under the same
        ladder)."""
        opts = opts or SolverSettings().with_precision(Precision.FP32_SCHUR)
        return MixedLdltSolver(LdltSolver.factor_with(a, opts), a)

    def solve(self, a: CscMatrix, b):
        return self.solve_to(a, b, CERT_FACTOR * EPS_HI)

    def solve_to(self, a: CscMatrix, b, target_backward_error: float):
        b = np.asarray(b)
        single = b.ndim == 1
        if a.n != self.lo.n() or b.shape[0] != self.lo.n():
            raise DimensionMismatch(f"expected {self.lo.n()}, got {a.n if a.n != self.lo.n() else b.shape[0]}")
        B = b.reshape(a.n, -1)
        k = B.shape[1]
        bb, bv = self.lo._as_c(B)
        x = np.zeros_like(bb)
        xv = x.view(np.float64) if self.lo._cplx else x
        info = np.zeros(4 * k, dtype=np.float64)
        vals = a.values_f64()
        _check(lib.rsl_refine(self.lo._h, ptr(vals, C.c_double), ptr(bv, C.c_double), k,
                              float(target_backward_error), ptr(xv, C.c_double), ptr(info, C.c_double)))
        infos = [MixedInfo(int(info[4 * c]), int(info[4 * c + 1]), float(info[4 * c + 2]),
                           bool(info[4 * c + 3] != 0)) for c in range(k)]
        if single:
            return x.reshape(a.n), infos[0]
        return x, infos

    def solve_to_device(self, d_a_values: int, d_b: int, nrhs: int, d_x: int, target_backward_error: float,
                        stream: int = 0):
        """solve_to with the matrix values, b and x in device memory (device
        addresses, b and x row-major n x nrhs, interleaved for the complex
        field). Returns the per-column MixedInfo list."""
        if nrhs <= 0:
            raise DimensionMismatch("nrhs must be positive")
        info = np.zeros(4 * nrhs, dtype=np.float64)
        _check(lib.rsl_refine_device(self.lo._h, C.c_void_p(d_a_values), C.c_void_p(d_b), int(nrhs),
                                     float(target_backward_error), C.c_void_p(d_x), ptr(info, C.c_double),
                                     C.c_void_p(stream) if stream else None))
        return [MixedInfo(int(info[4 * c]), int(info[4 * c + 1]), float(info[4 * c + 2]), bool(info[4 * c + 3] != 0))
                for c in range(nrhs)]


def amd_order(n, col_ptr, row_idx):
    """Stand-alone AMD permutation (new->old) of a full symmetric pattern, plus stats."""
    col_ptr = np.ascontiguousarray(col_ptr, dtype=np.int64)
    row_idx = np.ascontiguousarray(row_idx, dtype=np.int64)
    perm = np.zeros(max(n, 1), dtype=np.int32)
    st = np.zeros(7, dtype=np.float64)
    _check(lib.rsl_order(0, n, ptr(col_ptr, C.c_int64), ptr(row_idx, C.c_int64), 0,
                         ptr(perm, C.c_int32), ptr(st, C.c_double)))
    keys = ["ncmpa", "ndense", "lnz", "ndiv", "nms_lu", "nms_ldl", "dmax"]
    return perm[:n], dict(zip(keys, st.tolist()))


def nd_order(n, col_ptr, row_idx, seed: int = 1):
    col_ptr = np.ascontiguousarray(col_ptr, dtype=np.int64)
    row_idx = np.ascontiguousarray(row_idx, dtype=np.int64)
    perm = np.zeros(max(n, 1), dtype=np.int32)
    _check(lib.rsl_order(1, n, ptr(col_ptr, C.c_int64), ptr(row_idx, C.c_int64), seed,
                         ptr(perm, C.c_int32), None))
    return perm[:n]
