"""Synthetic matrices and right-hand sides of the five benchmark workloads.

Grid Laplacians follow rslab's stencil generator (src/matgen/stencil.rs:40-95):
off-diagonal -1 per grid edge, Dirichlet ghost diagonal 2*ndim, natural node
numbering i = x + nx*(y + ny*z). Random values use rslab's xorshift64* Rng
(src/matgen/mod.rs:63-96) with the fixed seeds listed per workload.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import lib, ptr
from .solver import CscMatrix


def rng_uniform(seed: int, lo: float, hi: float, count: int) -> np.ndarray:
    out = np.empty(max(int(count), 1), dtype=np.float64)
    lib.rsl_rng_uniform(int(seed), float(lo), float(hi), int(count), ptr(out, C.c_double))
    return out[: int(count)]


def grid_laplacian(dims, shift=0.0, diag_real=None, diag_imag=None) -> CscMatrix:
    """Lower CSC of the 5/7-point Laplacian on `dims` plus diagonal terms."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    stride = [1]
    for d in dims[:-1]:
        stride.append(stride[-1] * d)
    idx = np.arange(n, dtype=np.int64)
    diag = np.full(n, 2.0 * len(dims) + shift)
    if diag_real is not None:
        diag = diag + np.asarray(diag_real, dtype=np.float64)
    cplx = diag_imag is not None
    rows = [idx]
    cols = [idx]
    vals = [diag.astype(np.complex128) + 1j * np.asarray(diag_imag) if cplx else diag]
    for dd, st in enumerate(stride):
        coord = (idx // st) % dims[dd]
        m = coord + 1 < dims[dd]
        rows.append(idx[m] + st)
        cols.append(idx[m])
        vals.append(np.full(int(m.sum()), -1.0, dtype=np.complex128 if cplx else np.float64))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    order = np.lexsort((r, c))
    r, c, v = r[order], c[order], v[order]
    col_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(col_ptr, c + 1, 1)
    return CscMatrix(n, np.cumsum(col_ptr), r, v)


# --------------------------------------------------------------- workloads
def config1():
    """Real indefinite 2D: 100x100 grid, A = L - 2 I (n = 10,000), AMD, 1 RHS.
    x_true ~ U[-1,1] from Rng(1); b = A x_true."""
    a = grid_laplacian((100, 100), shift=-2.0)
    x = rng_uniform(1, -1.0, 1.0, a.n)
    return a, a.symv(x), x


def helmholtz_eta(m: int, layer: int = 4, peak: float = 0.5) -> np.ndarray:
    """Absorbing profile: 0 inside, peak*((layer-d)/layer)^2 at face distance d < layer."""
    g = np.arange(m)
    dist1 = np.minimum(g, m - 1 - g)
    z, y, x = np.meshgrid(dist1, dist1, dist1, indexing="ij")
    d = np.minimum(np.minimum(x, y), z).reshape(-1)
    eta = np.where(d < layer, peak * ((layer - d) / layer) ** 2, 0.0)
    return eta


def config2(m: int = 40, nrhs: int = 16):
    """Complex-symmetric 3D Helmholtz: A = L - 1.2 I + i diag(eta), MetisND, 16 RHS.
    RHS re/im ~ U[-1,1] from Rng(2), row-major (re, im) pairs."""
    a = grid_laplacian((m, m, m), shift=-1.2, diag_imag=helmholtz_eta(m))
    raw = rng_uniform(2, -1.0, 1.0, 2 * a.n * nrhs).reshape(a.n, nrhs, 2)
    b = raw[..., 0] + 1j * raw[..., 1]
    return a, b


def config3(m: int = 60, nrhs: int = 8):
    """Real indefinite 3D: A = L - 1.5 I + diag(delta), delta ~ U[-0.01,0.01]
    from Rng(3); 8 RHS ~ U[-1,1] from Rng(4), row-major."""
    n = m ** 3
    delta = rng_uniform(3, -0.01, 0.01, n)
    a = grid_laplacian((m, m, m), shift=-1.5, diag_real=delta)
    b = rng_uniform(4, -1.0, 1.0, n * nrhs).reshape(n, nrhs)
    return a, b


def config4(count: int = 1024):
    """Batched uneven 2D systems: side ~ U{16..96}, sigma ~ U[0.5,3.5] drawn
    alternately from Rng(5); A_i = L_i - sigma_i I; one RHS each from Rng(6)."""
    draws = rng_uniform(5, 0.0, 1.0, 2 * count)
    sides = 16 + np.minimum((draws[0::2] * 81).astype(np.int64), 80)
    sigmas = 0.5 + 3.0 * draws[1::2]
    systems = []
    total = int(np.sum(sides ** 2))
    rhs = rng_uniform(6, -1.0, 1.0, total)
    off = 0
    for s, sg in zip(sides, sigmas):
        a = grid_laplacian((int(s), int(s)), shift=-float(sg))
        systems.append((a, rhs[off:off + a.n].copy(), float(sg)))
        off += a.n
    return systems


def vector_stencil27(m: int = 48, omega2: float = 0.05) -> CscMatrix:
    """3 dof per node on an m^3 grid, 27-point coupling: off-diagonal block for
    neighbour offset d is -(1/26)(I + 0.25 d d^T/|d|^2); diagonal block is minus
    the sum of the row's off-diagonal blocks plus 0.01 I, shifted by -omega2 I."""
    nn = m ** 3
    node = np.arange(nn, dtype=np.int64)
    x = node % m
    y = (node // m) % m
    z = node // (m * m)
    offsets = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
               if (dx, dy, dz) != (0, 0, 0)]

    def block(d):
        d = np.asarray(d, dtype=np.float64)
        return -(np.eye(3) + 0.25 * np.outer(d, d) / d.dot(d)) / 26.0

    diag = np.zeros((nn, 3, 3))
    rows, cols, vals = [], [], []
    for d in offsets:
        dx, dy, dz = d
        ok = (x + dx >= 0) & (x + dx < m) & (y + dy >= 0) & (y + dy < m) & (z + dz >= 0) & (z + dz < m)
        B = block(d)
        diag[ok] -= B
        lin = dx + m * dy + m * m * dz
        if lin > 0:
            p = node[ok]
            q = p + lin
            for a_ in range(3):
                for b_ in range(3):
                    rows.append(3 * q + b_)
                    cols.append(3 * p + a_)
                    vals.append(np.full(p.shape[0], B[b_, a_]))
    diag += (0.01 - omega2) * np.eye(3)
    for a_ in range(3):
        for b_ in range(a_, 3):
            rows.append(3 * node + b_)
            cols.append(3 * node + a_)
            vals.append(diag[:, b_, a_].copy())
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    order = np.lexsort((r, c))
    n = 3 * nn
    col_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(col_ptr, c[order] + 1, 1)
    return CscMatrix(n, np.cumsum(col_ptr), r[order], v[order])


def config5_omegas(count: int = 20):
    return [0.05 + k * 0.45 / (count - 1) for k in range(count)]


def config5_rhs(n: int, k: int, nrhs: int = 32):
    return rng_uniform(7 + k, -1.0, 1.0, n * nrhs).reshape(n, nrhs)


# ------------------------------------------------------- unsymmetric (LU)
def convection_diffusion(dims, velocity=(0.8, 0.4, 0.2), shift=0.0, delta=None, eta=None):
    """Full CSC (GeneralCsc) of the 5/7-point diffusion + central-difference
    convection operator on `dims`: for each grid edge (i, i + e_d) the
    entries A[i + e_d, i] = -1 - v_d/2 and A[i, i + e_d] = -1 + v_d/2 (v_d the
    cell Peclet number of direction d), diagonal 2*ndim + shift (+ delta)
    (+ i*eta for a complex matrix). Unsymmetric values on a symmetric
    pattern, indefinite for negative shifts."""
    from .lu import GeneralCsc
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    stride = [1]
    for d in dims[:-1]:
        stride.append(stride[-1] * d)
    idx = np.arange(n, dtype=np.int64)
    diag = np.full(n, 2.0 * len(dims) + shift)
    if delta is not None:
        diag = diag + np.asarray(delta, dtype=np.float64)
    cplx = eta is not None
    dt = np.complex128 if cplx else np.float64
    rows, cols = [idx], [idx]
    vals = [diag.astype(dt) + (1j * np.asarray(eta) if cplx else 0.0)]
    for dd, st in enumerate(stride):
        v = float(velocity[dd]) if dd < len(velocity) else 0.0
        coord = (idx // st) % dims[dd]
        m = coord + 1 < dims[dd]
        lo = idx[m]
        rows += [lo + st, lo]
        cols += [lo, lo + st]
        vals += [np.full(lo.size, -1.0 - 0.5 * v, dtype=dt), np.full(lo.size, -1.0 + 0.5 * v, dtype=dt)]
    return GeneralCsc.from_triplets(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def lu_workload(m: int = 60, nrhs: int = 8, cplx: bool = False):
    """Unsymmetric workload of the LU twin: convection_diffusion on m^3 with
    velocity (0.8, 0.4, 0.2), shift -1.5 and delta ~ U[-0.01, 0.01] from
    Rng(3) (the grid, shift and perturbation of config 3), plus i*0.1 on the
    diagonal when complex; B ~ U[-1, 1] from Rng(4), n x nrhs row-major."""
    n = m ** 3
    delta = rng_uniform(3, -0.01, 0.01, n)
    a = convection_diffusion((m, m, m), shift=-1.5, delta=delta, eta=np.full(n, 0.1) if cplx else None)
    B = rng_uniform(4, -1.0, 1.0, n * nrhs).reshape(n, nrhs)
    return a, (B.astype(np.complex128) if cplx else B)
