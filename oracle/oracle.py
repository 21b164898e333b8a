"""Python handle on the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, by __graft_entry__.smoke() as the checker and by bench.py's
cpu_baseline leg. The product never imports this module. The C restatement
and the reference functions it follows are listed in ldlt_oracle.h.

Also holds small numpy restatements used to pin the host symbolic:
brute-force elimination column counts (the O(n^2) simulation of
src/symbolic/column_counts.rs:27-77) and the reference's Inertia closed forms.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def _load():
    # idle OpenMP threads sleep instead of spinning, so the oracle does not
    # starve other work on a shared host (its results do not depend on it)
    os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    return lib


class OraSym(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("nsuper", C.c_int32),
        ("perm", C.POINTER(C.c_int32)), ("perm_inv", C.POINTER(C.c_int32)),
        ("sn_first", C.POINTER(C.c_int32)), ("sn_ncol", C.POINTER(C.c_int32)),
        ("ch_ptr", C.POINTER(C.c_int32)), ("ch_idx", C.POINTER(C.c_int32)),
        ("pp_colp", C.POINTER(C.c_int64)), ("pp_rowi", C.POINTER(C.c_int32)),
    ]


class OraFactors(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("cplx", C.c_int32), ("status", C.c_int32),
        ("nnzl", C.c_int64), ("n_perturbed", C.c_int64), ("inertia", C.c_int64 * 3),
        ("l_col_ptr", C.POINTER(C.c_int64)), ("l_row_idx", C.POINTER(C.c_int32)),
        ("l_values", C.POINTER(C.c_double)), ("d_diag", C.POINTER(C.c_double)),
        ("d_subdiag", C.POINTER(C.c_double)), ("two_by_two", C.POINTER(C.c_uint8)),
        ("perm", C.POINTER(C.c_int32)),
    ]


_lib = _load()
_lib.ora_factor.restype = C.POINTER(OraFactors)
_lib.ora_factor.argtypes = [C.POINTER(OraSym), C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                            C.POINTER(C.c_double), C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]
_lib.ora_round_lowp.restype = C.c_float
_lib.ora_round_lowp.argtypes = [C.c_double, C.c_int]

_lib.ora_set_threads.argtypes = [C.c_int]
_lib.ora_get_threads.restype = C.c_int

PREC_FP64, PREC_FP32, PREC_BF16 = 0, 1, 2


def set_threads(n: int):
    """Threads of the oracle's factor (0 = all available); the factor does
    not depend on the count."""
    _lib.ora_set_threads(int(n))


def get_threads() -> int:
    return int(_lib.ora_get_threads())


def round_lowp(v: float, prec: int) -> float:
    """v rounded as the reduced-precision products round their operands
    (prec 1: to FP32; 2: to BF16 by way of FP32, round to nearest even)."""
    return float(_lib.ora_round_lowp(float(v), int(prec)))
_lib.ora_free.argtypes = [C.POINTER(OraFactors)]
_lib.ora_solve_many.restype = C.c_int
_lib.ora_solve_many.argtypes = [C.POINTER(OraFactors), C.POINTER(C.c_double), C.c_int32,
                                C.POINTER(C.c_double)]
_lib.ora_onepass_scale.argtypes = [C.c_int32, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
_lib.ora_symv.argtypes = [C.c_int32, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                          C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
_lib.ora_sym_norm_inf.restype = C.c_double
_lib.ora_sym_norm_inf.argtypes = [C.c_int32, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_double)]


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _vals(a):
    if np.iscomplexobj(a.values):
        return np.ascontiguousarray(a.values, dtype=np.complex128).view(np.float64), 1
    return np.ascontiguousarray(a.values, dtype=np.float64), 0


@dataclasses.dataclass
class OracleFactors:
    n: int
    l_col_ptr: np.ndarray
    l_row_idx: np.ndarray
    l_values: np.ndarray
    d_diag: np.ndarray
    d_subdiag: np.ndarray
    two_by_two: np.ndarray
    perm: np.ndarray
    n_perturbed: int
    inertia: tuple
    cplx: bool
    _handle: object = None

    def solve_many(self, b):
        """solve_ldlt_many on the (already equilibrated) factor; b row-major n x k."""
        b = np.asarray(b)
        single = b.ndim == 1
        B = np.ascontiguousarray(b.reshape(self.n, -1), dtype=np.complex128 if self.cplx else np.float64)
        X = np.zeros_like(B)
        bv = B.view(np.float64) if self.cplx else B
        xv = X.view(np.float64) if self.cplx else X
        rc = _lib.ora_solve_many(self._handle, _p(bv, C.c_double), B.shape[1], _p(xv, C.c_double))
        if rc:
            raise ZeroDivisionError("singular D block")
        return X.reshape(self.n) if single else X

    def solve_arrays(self, b):
        """solve_ldlt (ldlt_generic.rs:339-424) in numpy on this object's arrays
        (small cases; lets a test plant a change in L or D and observe it)."""
        y = np.asarray(b)[self.perm].astype(self.l_values.dtype)
        n = self.n
        for j in range(n):
            s, e = self.l_col_ptr[j], self.l_col_ptr[j + 1]
            y[self.l_row_idx[s + 1:e]] -= self.l_values[s + 1:e] * y[j]
        k = 0
        while k < n:
            if self.two_by_two[k]:
                d11, d21, d22 = self.d_diag[k], self.d_subdiag[k], self.d_diag[k + 1]
                det = d11 * d22 - d21 * d21
                z0, z1 = y[k], y[k + 1]
                y[k], y[k + 1] = (d22 * z0 - d21 * z1) / det, (d11 * z1 - d21 * z0) / det
                k += 2
            else:
                y[k] = y[k] / self.d_diag[k]
                k += 1
        for j in range(n - 1, -1, -1):
            s, e = self.l_col_ptr[j], self.l_col_ptr[j + 1]
            y[j] -= np.dot(self.l_values[s + 1:e], y[self.l_row_idx[s + 1:e]])
        x = np.empty_like(y)
        x[self.perm] = y
        return x

    def __del__(self):
        if self._handle is not None:
            _lib.ora_free(self._handle)
            self._handle = None


class RankDeficient(Exception):
    pass


def onepass_scale(a) -> np.ndarray:
    v, cp = _vals(a)
    ri = np.ascontiguousarray(a.row_idx, dtype=np.int32)
    cpt = np.ascontiguousarray(a.col_ptr, dtype=np.int64)
    s = np.zeros(max(a.n, 1))
    _lib.ora_onepass_scale(a.n, cp, _p(cpt, C.c_int64), _p(ri, C.c_int32), _p(v, C.c_double),
                           _p(s, C.c_double))
    return s[: a.n]


def symv(a, x):
    v, cp = _vals(a)
    ri = np.ascontiguousarray(a.row_idx, dtype=np.int32)
    cpt = np.ascontiguousarray(a.col_ptr, dtype=np.int64)
    dt = np.complex128 if cp else np.float64
    x = np.ascontiguousarray(x, dtype=dt)
    y = np.zeros_like(x)
    xv = x.view(np.float64) if cp else x
    yv = y.view(np.float64) if cp else y
    _lib.ora_symv(a.n, cp, _p(cpt, C.c_int64), _p(ri, C.c_int32), _p(v, C.c_double),
                  _p(xv, C.c_double), _p(yv, C.c_double))
    return y


def sym_norm_inf(a) -> float:
    v, cp = _vals(a)
    ri = np.ascontiguousarray(a.row_idx, dtype=np.int32)
    cpt = np.ascontiguousarray(a.col_ptr, dtype=np.int64)
    return float(_lib.ora_sym_norm_inf(a.n, cp, _p(cpt, C.c_int64), _p(ri, C.c_int32),
                                       _p(v, C.c_double)))


def factor(sym, a_scaled, panel_nb=64, pivot=0, abs_floor=0.0, drop_tol=-1.0, prec=0) -> OracleFactors:
    """factor_left_looking on an equilibrated matrix given the exported symbolic
    (prec: 0 FP64, 1 FP32 / 2 BF16 Schur-product emulation, real field)."""
    if drop_tol is None:
        drop_tol = -1.0
    keep = dict(
        perm=np.ascontiguousarray(sym.perm, dtype=np.int32),
        perm_inv=np.ascontiguousarray(sym.perm_inv, dtype=np.int32),
        sn_first=np.ascontiguousarray(sym.sn_first, dtype=np.int32),
        sn_ncol=np.ascontiguousarray(sym.sn_ncol, dtype=np.int32),
        ch_ptr=np.ascontiguousarray(sym.ch_ptr, dtype=np.int32),
        ch_idx=np.ascontiguousarray(sym.ch_idx, dtype=np.int32),
        pp_colp=np.ascontiguousarray(sym.pp_colp, dtype=np.int64),
        pp_rowi=np.ascontiguousarray(sym.pp_rowi, dtype=np.int32),
    )
    s = OraSym(sym.n, len(sym.sn_first), _p(keep["perm"], C.c_int32), _p(keep["perm_inv"], C.c_int32),
               _p(keep["sn_first"], C.c_int32), _p(keep["sn_ncol"], C.c_int32),
               _p(keep["ch_ptr"], C.c_int32), _p(keep["ch_idx"], C.c_int32),
               _p(keep["pp_colp"], C.c_int64), _p(keep["pp_rowi"], C.c_int32))
    v, cp = _vals(a_scaled)
    ri = np.ascontiguousarray(a_scaled.row_idx, dtype=np.int32)
    cpt = np.ascontiguousarray(a_scaled.col_ptr, dtype=np.int64)
    h = _lib.ora_factor(C.byref(s), cp, _p(cpt, C.c_int64), _p(ri, C.c_int32), _p(v, C.c_double),
                        panel_nb, pivot, abs_floor, drop_tol, int(prec))
    f = h.contents
    if f.status != 0:
        _lib.ora_free(h)
        raise RankDeficient("numerically rank deficient")
    n, nnz = f.n, f.nnzl
    nd = 2 if cp else 1

    def arr(p, k, dt):
        return np.ctypeslib.as_array(p, shape=(k,)).astype(dt, copy=True)

    lv = arr(f.l_values, nnz * nd, np.float64)
    dd = arr(f.d_diag, n * nd, np.float64)
    ds = arr(f.d_subdiag, n * nd, np.float64)
    if cp:
        lv, dd, ds = lv.view(np.complex128), dd.view(np.complex128), ds.view(np.complex128)
    return OracleFactors(n, arr(f.l_col_ptr, n + 1, np.int64), arr(f.l_row_idx, nnz, np.int32), lv, dd,
                         ds, arr(f.two_by_two, n, np.uint8).astype(bool), arr(f.perm, n, np.int32),
                         int(f.n_perturbed), tuple(int(x) for x in f.inertia), bool(cp), h)


def scaled(a, s):
    """A_hat = D A D on the same pattern (apply_symmetric_scaling, sparse_solver.rs:380-397)."""
    cols = np.repeat(np.arange(a.n), np.diff(a.col_ptr))
    vals = a.values * (s[a.row_idx] * s[cols])
    return type(a)(a.n, a.col_ptr, a.row_idx, vals)


def solve(a, b, sym, panel_nb=64, pivot=0, abs_floor=0.0, scaling=True):
    """LdltSymbolic::factor + LdltSolver::solve_many on the CPU oracle."""
    s = onepass_scale(a) if scaling else np.ones(a.n)
    f = factor(sym, scaled(a, s), panel_nb=panel_nb, pivot=pivot, abs_floor=abs_floor)
    B = np.asarray(b)
    single = B.ndim == 1
    B = B.reshape(a.n, -1)
    X = f.solve_many(B * s[:, None]) * s[:, None]
    return (X.reshape(a.n) if single else X), f


def backward_error(a, x, b, anorm=None):
    """Normwise backward error |r|inf / (|A|inf |x|inf + |b|inf) per column (mixed.rs:151-155)."""
    anorm = sym_norm_inf(a) if anorm is None else anorm
    X = np.asarray(x).reshape(a.n, -1)
    B = np.asarray(b).reshape(a.n, -1)
    out = []
    for c in range(X.shape[1]):
        r = B[:, c] - symv(a, X[:, c])
        den = max(anorm * np.abs(X[:, c]).max() + np.abs(B[:, c]).max(), np.finfo(float).tiny)
        out.append(float(np.abs(r).max() / den))
    return np.array(out)


# ------------------------------------------------------------ refinement ladder
CERT_FACTOR, IR_MAX, IR_CONTRACT, GMRES_RESTART, GMRES_MAX = 64.0, 8, 0.5, 30, 120  # mixed.rs:97-105
_TINY = np.finfo(float).tiny


def _nrm2(v):
    return float(np.sqrt(np.vdot(v, v).real))


def gmres(op, b, pre, tol, max_iter=GMRES_MAX, restart=GMRES_RESTART, x0=None):
    """Flexible restarted GMRES (iterative.rs:822-1021): right preconditioner
    `pre`, modified Gram-Schmidt with DGKS re-orthogonalization (eta = 1/sqrt 2),
    Givens rotations (c, s) = (f, g)/sqrt(|f|^2+|g|^2) applied as
    [[conj c, conj s], [-s, c]], early exit of a cycle on the least-squares
    estimate, the true residual relative to ||b||_2 tested at the top of each
    cycle, the breakdown-guarded back substitution (well_conditioned_dim_flat,
    :646-661) and x += Z y. Returns (x, iterations, converged)."""
    eta = 1.0 / np.sqrt(2.0)
    n = b.shape[0]
    m = max(restart, 1)
    dt = np.result_type(b, np.float64)
    x = np.zeros(n, dtype=dt) if x0 is None else np.array(x0, dtype=dt)
    bnorm = _nrm2(b)
    if bnorm == 0.0:
        return np.zeros(n, dtype=dt), 0, True
    total = 0
    while True:
        r = b - op(x)
        beta = _nrm2(r)
        final = beta / bnorm
        if final <= tol or total >= max_iter:
            break
        V = np.zeros((m + 1, n), dtype=dt)
        Z = np.zeros((m, n), dtype=dt)
        h = np.zeros((m + 1, m), dtype=dt)
        cs = np.zeros(m, dtype=dt)
        sn = np.zeros(m, dtype=dt)
        g = np.zeros(m + 1, dtype=dt)
        V[0] = r / beta
        g[0] = beta
        jdim = 0
        for j in range(m):
            if total >= max_iter:
                break
            Z[j] = pre(V[j])
            w = op(Z[j])
            w0 = _nrm2(w)
            for i in range(j + 1):
                hij = np.vdot(V[i], w)
                h[i, j] = hij
                w = w - hij * V[i]
            hn = _nrm2(w)
            if hn < eta * w0:
                for i in range(j + 1):
                    s_ = np.vdot(V[i], w)
                    h[i, j] += s_
                    w = w - s_ * V[i]
                hn = _nrm2(w)
            h[j + 1, j] = hn
            V[j + 1] = w / hn if hn > 0.0 else 0.0
            for i in range(j):
                t = np.conj(cs[i]) * h[i, j] + np.conj(sn[i]) * h[i + 1, j]
                h[i + 1, j] = -sn[i] * h[i, j] + cs[i] * h[i + 1, j]
                h[i, j] = t
            f_, g_ = h[j, j], h[j + 1, j]
            if g_ == 0:
                c_, s2 = 1.0, 0.0
            elif f_ == 0:
                c_, s2 = 0.0, 1.0
            else:
                rr = np.sqrt(abs(f_) ** 2 + abs(g_) ** 2)
                c_, s2 = f_ / rr, g_ / rr
            cs[j], sn[j] = c_, s2
            h[j, j] = np.conj(c_) * f_ + np.conj(s2) * g_
            h[j + 1, j] = 0.0
            g_next = -s2 * g[j]
            g[j] = np.conj(c_) * g[j]
            g[j + 1] = g_next
            total += 1
            jdim = j + 1
            if abs(g[j + 1]) / bnorm <= tol:
                break
        hmax = max((abs(h[i, i]) for i in range(jdim)), default=0.0)
        jd = jdim
        for i in range(jdim):
            if abs(h[i, i]) <= np.finfo(float).eps * hmax:
                jd = i
                break
        y = np.zeros(m, dtype=dt)
        for i in range(jd - 1, -1, -1):
            s_ = g[i]
            for k in range(i + 1, jd):
                s_ = s_ - h[i, k] * y[k]
            y[i] = s_ / h[i, i]
        for i in range(jd):
            x = x + Z[i] * y[i]
    return x, total, final <= tol


def _stop(converged, iters, max_iter):
    """stop_reason (iterative.rs:351-363)."""
    return "converged" if converged else ("breakdown" if iters < max_iter else "max_iter")


def cocg(op, b, pre, tol, max_iter):
    """Preconditioned COCG (iterative.rs:402-489) for complex-symmetric A,
    x0 = 0, unconjugated products r^T z and p^T A p, breakdown on a zero
    denominator. Returns (x, iters, converged, final_res, stop)."""
    b = np.asarray(b)
    x = np.zeros_like(b)
    r = b.copy()
    bnorm = _nrm2(b)
    if bnorm == 0.0:
        return x, 0, True, 0.0, "converged"
    z = pre(r)
    p = z.copy()
    rho = np.sum(r * z)
    final, conv, it = _nrm2(r) / bnorm, False, 0
    while it < max_iter:
        q = op(p)
        pq = np.sum(p * q)
        if pq == 0:
            break
        alpha = rho / pq
        x = x + alpha * p
        r = r - alpha * q
        it += 1
        final = _nrm2(r) / bnorm
        if final <= tol:
            conv = True
            break
        z = pre(r)
        rho_new = np.sum(r * z)
        if rho == 0:
            break
        beta = rho_new / rho
        p = z + beta * p
        rho = rho_new
    return x, it, conv, final, _stop(conv, it, max_iter)


def cocr(op, b, pre, tol, max_iter):
    """Preconditioned COCR (iterative.rs:495-583), x0 = 0: w = M^-1 A p,
    alpha = gamma / ((A p)^T w), the four coupled updates, beta from
    z^T A z. Returns (x, iters, converged, final_res, stop)."""
    b = np.asarray(b)
    x = np.zeros_like(b)
    r = b.copy()
    bnorm = _nrm2(b)
    if bnorm == 0.0:
        return x, 0, True, 0.0, "converged"
    z = pre(r)
    p = z.copy()
    ap = op(p)
    az = ap.copy()
    gamma = np.sum(z * az)
    final, conv, it = _nrm2(r) / bnorm, False, 0
    while it < max_iter:
        w = pre(ap)
        aw = op(w)
        denom = np.sum(ap * w)
        if denom == 0:
            break
        alpha = gamma / denom
        x = x + alpha * p
        r = r - alpha * ap
        z = z - alpha * w
        az = az - alpha * aw
        it += 1
        final = _nrm2(r) / bnorm
        if final <= tol:
            conv = True
            break
        gamma_new = np.sum(z * az)
        if gamma == 0:
            break
        beta = gamma_new / gamma
        p = z + beta * p
        ap = az + beta * ap
        gamma = gamma_new
    return x, it, conv, final, _stop(conv, it, max_iter)


def refine_ladder(a, fo, s, b, target, anorm=None):
    """refine_ladder (mixed.rs:140-221) for one right-hand side: x0 = the
    factor's solve of b; plain IR with the FP64 residual against the original
    A while the backward error |r|inf / (|A|inf |x|inf + |b|inf) is above
    `target`, at most IR_MAX steps and each shrinking |r|inf by IR_CONTRACT;
    then FGMRES warm-started at the best iterate, preconditioned by the same
    factor; the better of the two kept. `fo` factors the equilibrated matrix
    with scale `s` (LdltSolver::solve applies it). Returns (x, ir_iters,
    gmres_iters, backward_error, certified)."""
    anorm = sym_norm_inf(a) if anorm is None else anorm
    b = np.asarray(b)
    pre = lambda r: fo.solve_many(r * s) * s  # noqa: E731
    op = lambda v: symv(a, v)  # noqa: E731
    bn = float(np.abs(b).max()) if b.size else 0.0

    def berr(x, r):
        return float(np.abs(r).max()) / max(anorm * float(np.abs(x).max()) + bn, _TINY)

    x = pre(b)
    ir = 0
    best_x, best_be, prev_rn = x.copy(), np.inf, np.inf
    while True:
        r = b - op(x)
        be = berr(x, r)
        if be < best_be:
            best_be, best_x = be, x.copy()
        if be <= target:
            return best_x, ir, 0, best_be, True
        rn = float(np.abs(r).max())
        if ir >= IR_MAX or (ir > 0 and rn > prev_rn * IR_CONTRACT):
            break
        prev_rn = rn
        x = x + pre(r)
        ir += 1
    xg, its, _ = gmres(op, b, pre, target, GMRES_MAX, GMRES_RESTART, best_x)
    be2 = berr(xg, b - op(xg))
    if be2 < best_be:
        return xg, ir, its, be2, be2 <= target
    return best_x, ir, its, best_be, best_be <= target


# ------------------------------------------------------------ symbolic pins
def brute_force_col_counts(n, colp, rowi):
    """Exact column counts of L by symbolic elimination on a full pattern
    (the O(n^2) simulation of column_counts.rs:27-77)."""
    cols = [set() for _ in range(n)]
    for j in range(n):
        for k in range(colp[j], colp[j + 1]):
            i = int(rowi[k])
            if i > j:
                cols[j].add(i)
    counts = [1] * n
    for j in range(n):
        rows = sorted(cols[j])
        counts[j] += len(rows)
        if len(rows) > 1:
            cols[rows[0]].update(rows[1:])
    return np.array(counts)


def brute_force_etree(n, colp, rowi):
    """parent[j] = min{ i > j : L(i,j) != 0 } by symbolic elimination."""
    cols = [set() for _ in range(n)]
    for j in range(n):
        for k in range(colp[j], colp[j + 1]):
            i = int(rowi[k])
            if i > j:
                cols[j].add(i)
    parent = [-1] * n
    for j in range(n):
        rows = sorted(cols[j])
        if rows:
            parent[j] = rows[0]
            cols[rows[0]].update(rows[1:])
    return np.array(parent)


def laplacian2d_inertia(m: int, sigma: float):
    """Closed-form inertia of L - sigma I on an m x m Dirichlet grid:
    eigenvalues 4 - 2cos(p pi/(m+1)) - 2cos(q pi/(m+1))."""
    t = 2.0 - 2.0 * np.cos(np.arange(1, m + 1) * np.pi / (m + 1))
    lam = (t[:, None] + t[None, :]).ravel() - sigma
    return int((lam > 0).sum()), int((lam < 0).sum()), int((lam == 0).sum()), float(np.abs(lam).min())


# ------------------------------------------------------------------ LU twin
class OruFactors(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("cplx", C.c_int32), ("status", C.c_int32),
        ("n_perturbed", C.c_int64), ("nnz_l", C.c_int64), ("nnz_u", C.c_int64),
        ("l_col_ptr", C.POINTER(C.c_int64)), ("l_row_idx", C.POINTER(C.c_int32)),
        ("l_values", C.POINTER(C.c_double)),
        ("u_row_ptr", C.POINTER(C.c_int64)), ("u_col_idx", C.POINTER(C.c_int32)),
        ("u_values", C.POINTER(C.c_double)),
        ("perm", C.POINTER(C.c_int32)), ("perm_row", C.POINTER(C.c_int32)),
        ("d_row", C.POINTER(C.c_double)), ("d_col", C.POINTER(C.c_double)),
    ]


_lib.oru_factor.restype = C.POINTER(OruFactors)
_lib.oru_factor.argtypes = [C.POINTER(OraSym), C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                            C.POINTER(C.c_double), C.c_int, C.c_double, C.c_double, C.c_double]
_lib.oru_free.argtypes = [C.POINTER(OruFactors)]
_lib.oru_solve_many.restype = C.c_int
_lib.oru_solve_many.argtypes = [C.POINTER(OruFactors), C.POINTER(C.c_double), C.c_int32,
                                C.POINTER(C.c_double)]


def symmetrized_lower(n, col_ptr, row_idx):
    """Lower triangle of the pattern A + A^T as (col_ptr, row_idx)
    (symmetrized_lower_pattern, multifrontal_lu.rs:364-409)."""
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(np.asarray(col_ptr, dtype=np.int64)))
    rows = np.asarray(row_idx, dtype=np.int64)
    hi, lo = np.maximum(rows, cols), np.minimum(rows, cols)
    key = np.unique(lo * n + hi)
    cp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(cp, key // n + 1, 1)
    return np.cumsum(cp), key % n


def general_matvec(a, x):
    """y = A x for a full (unsymmetric) CSC matrix `a` (GeneralCsc::matvec)."""
    X = np.asarray(x)
    cols = np.repeat(np.arange(a.n), np.diff(a.col_ptr))
    vals = np.asarray(a.values)
    y = np.zeros((a.n,) + X.shape[1:], dtype=np.result_type(vals, X))
    contrib = vals[:, None] * X[cols] if X.ndim == 2 else vals * X[cols]
    np.add.at(y, np.asarray(a.row_idx), contrib)
    return y


def general_norm_inf(a) -> float:
    """max_i sum_j |A_ij| of a full CSC matrix."""
    r = np.zeros(a.n)
    np.add.at(r, np.asarray(a.row_idx), np.abs(np.asarray(a.values)))
    return float(r.max()) if a.n else 0.0


def lu_backward_error(a, x, b):
    """Normwise backward error |b - A x|inf / (|A|inf |x|inf + |b|inf) per column."""
    an = general_norm_inf(a)
    X = np.asarray(x).reshape(a.n, -1)
    B = np.asarray(b).reshape(a.n, -1)
    R = B - general_matvec(a, X)
    out = []
    for c in range(X.shape[1]):
        den = max(an * np.abs(X[:, c]).max() + np.abs(B[:, c]).max(), np.finfo(float).tiny)
        out.append(float(np.abs(R[:, c]).max() / den))
    return np.array(out)


@dataclasses.dataclass
class OracleLu:
    """LuFactors (multifrontal_lu.rs:321-352): L CSC unit lower, U CSR upper,
    perm / perm_row (position -> original column / row), d_row, d_col."""
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
    cplx: bool
    _handle: object = None

    def factor_nnz(self) -> int:
        return int(self.l_values.size + self.u_values.size)

    def solve_many(self, b):
        """solve_lu_block on the unscaled system; b row-major n x k."""
        b = np.asarray(b)
        single = b.ndim == 1
        B = np.ascontiguousarray(b.reshape(self.n, -1), dtype=np.complex128 if self.cplx else np.float64)
        X = np.zeros_like(B)
        bv = B.view(np.float64) if self.cplx else B
        xv = X.view(np.float64) if self.cplx else X
        if _lib.oru_solve_many(self._handle, _p(bv, C.c_double), B.shape[1], _p(xv, C.c_double)):
            raise RankDeficient("numerically rank deficient")
        return X.reshape(self.n) if single else X

    def dense_lu(self):
        """(L, U) as dense matrices in factorization order."""
        n = self.n
        L = np.zeros((n, n), dtype=self.l_values.dtype)
        U = np.zeros((n, n), dtype=self.u_values.dtype)
        for e in range(n):
            s, t = self.l_col_ptr[e], self.l_col_ptr[e + 1]
            L[self.l_row_idx[s:t], e] = self.l_values[s:t]
            s, t = self.u_row_ptr[e], self.u_row_ptr[e + 1]
            U[e, self.u_col_idx[s:t]] = self.u_values[s:t]
        return L, U

    def __del__(self):
        if self._handle is not None:
            _lib.oru_free(self._handle)
            self._handle = None


def lu_factor(sym, a, pivot=0, abs_floor=0.0, pivot_u=0.1, drop_tol=None) -> OracleLu:
    """factor_general_lu_numeric, left-looking path, on the full matrix `a`
    (unscaled) given the export of the symmetrized pattern's analysis."""
    keep = dict(
        perm=np.ascontiguousarray(sym.perm, dtype=np.int32),
        perm_inv=np.ascontiguousarray(sym.perm_inv, dtype=np.int32),
        sn_first=np.ascontiguousarray(sym.sn_first, dtype=np.int32),
        sn_ncol=np.ascontiguousarray(sym.sn_ncol, dtype=np.int32),
        ch_ptr=np.ascontiguousarray(sym.ch_ptr, dtype=np.int32),
        ch_idx=np.ascontiguousarray(sym.ch_idx, dtype=np.int32),
        pp_colp=np.ascontiguousarray(sym.pp_colp, dtype=np.int64),
        pp_rowi=np.ascontiguousarray(sym.pp_rowi, dtype=np.int32),
    )
    s = OraSym(sym.n, len(sym.sn_first), _p(keep["perm"], C.c_int32), _p(keep["perm_inv"], C.c_int32),
               _p(keep["sn_first"], C.c_int32), _p(keep["sn_ncol"], C.c_int32),
               _p(keep["ch_ptr"], C.c_int32), _p(keep["ch_idx"], C.c_int32),
               _p(keep["pp_colp"], C.c_int64), _p(keep["pp_rowi"], C.c_int32))
    v, cp = _vals(a)
    ri = np.ascontiguousarray(a.row_idx, dtype=np.int32)
    cpt = np.ascontiguousarray(a.col_ptr, dtype=np.int64)
    h = _lib.oru_factor(C.byref(s), cp, _p(cpt, C.c_int64), _p(ri, C.c_int32), _p(v, C.c_double),
                        int(pivot), float(abs_floor), float(pivot_u), -1.0 if drop_tol is None else float(drop_tol))
    f = h.contents
    if f.status != 0:
        _lib.oru_free(h)
        raise RankDeficient("numerically rank deficient")
    n, nd = f.n, (2 if cp else 1)

    def arr(p, k, dt):
        return np.ctypeslib.as_array(p, shape=(k,)).astype(dt, copy=True) if k else np.zeros(0, dt)

    lv = arr(f.l_values, f.nnz_l * nd, np.float64)
    uv = arr(f.u_values, f.nnz_u * nd, np.float64)
    if cp:
        lv, uv = lv.view(np.complex128), uv.view(np.complex128)
    return OracleLu(n, arr(f.l_col_ptr, n + 1, np.int64), arr(f.l_row_idx, f.nnz_l, np.int32), lv,
                    arr(f.u_row_ptr, n + 1, np.int64), arr(f.u_col_idx, f.nnz_u, np.int32), uv,
                    arr(f.perm, n, np.int32), arr(f.perm_row, n, np.int32), arr(f.d_row, n, np.float64),
                    arr(f.d_col, n, np.float64), int(f.n_perturbed), bool(cp), h)
