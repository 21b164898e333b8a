// Multi-RHS triangular solves on the supernodal panels (gfx950).
//
// Computes x = D_s P L^-T D_bk^-1 L^-1 P^T (D_s b) with b, x row-major n x nrhs,
// as LdltSolver::solve_many (src/numeric/sparse_solver.rs:302-327) around
// solve_ldlt_block (src/dense/ldlt_generic.rs:618-714). The sweeps run level by
// level over the supernode tree directly on the factored panels:
//   forward  : pull form -- a supernode first gathers the updates of its
//              descendants (ascending updater order), then solves its own
//              unit-lower block (the 2x2 coupling row is D, not L);
//   D block  : 1x1 / 2x2 inverse per pivot block;
//   backward : a supernode gathers its off-diagonal rows from already solved
//              ancestors, then back-substitutes its own block.
// Every output entry is written by one thread in a fixed order: no atomics,
// run-to-run identical.
#include <hip/hip_runtime.h>

#include "devtypes.hpp"
#include "prof.hpp"
#include "scalar.hpp"

namespace rsl {

constexpr int SNT = 256;
constexpr int RCH = 8;  // right-hand sides per accumulator chunk

template <class T>
__global__ void k_gather_rhs(int32_t n, int32_t nrhs, const int32_t* __restrict__ perm_e,
                             const double* __restrict__ scale, const T* __restrict__ b,
                             T* __restrict__ y) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * nrhs) return;
  int64_t e = idx / nrhs, c = idx % nrhs;
  int32_t u = perm_e[e];
  y[idx] = F<T>::scale(b[(int64_t)u * nrhs + c], scale[u]);
}

template <class T>
__global__ void k_scatter_sol(int32_t n, int32_t nrhs, const int32_t* __restrict__ perm_e,
                              const double* __restrict__ scale, const T* __restrict__ y,
                              T* __restrict__ x) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * nrhs) return;
  int64_t e = idx / nrhs, c = idx % nrhs;
  int32_t u = perm_e[e];
  x[(int64_t)u * nrhs + c] = F<T>::scale(y[idx], scale[u]);
}

template <class T>
__device__ inline T shfl_t(T v, int src) {
  if constexpr (sizeof(T) == sizeof(double)) {
    return __shfl(v, src, 64);
  } else {
    cplx o;
    o.re = __shfl(v.re, src, 64);
    o.im = __shfl(v.im, src, 64);
    return o;
  }
}

template <class T>
__device__ inline T shfl_xor_t(T v, int m) {
  if constexpr (sizeof(T) == sizeof(double)) {
    return __shfl_xor(v, m, 64);
  } else {
    cplx o;
    o.re = __shfl_xor(v.re, m, 64);
    o.im = __shfl_xor(v.im, m, 64);
    return o;
  }
}

// L entry of the own block, with the 2x2 coupling row (held in D) read as zero.
template <class T>
__device__ inline T l_own(const T* P, int64_t h, const uint8_t* two, int i, int p) {
  if (i == p + 1 && two[p]) return F<T>::zero();
  return P[i + (int64_t)p * h];
}

constexpr int TB = 64;  // triangular block

// Forward, gather: own rows of a level's supernodes take the contributions of
// their descendants (ascending updater order, precomputed); 64-row chunks.
template <class T>
__global__ __launch_bounds__(SNT) void k_fgather(DevSym S, DevFac<T> Fc, const int2* __restrict__ items,
                                                 int32_t nrhs, T* __restrict__ y, const T* __restrict__ u) {
  const int2 it = items[blockIdx.x];
  const int s = it.x, r0 = it.y;
  const int w = S.sn_ncol[s], first = S.sn_first[s];
  const int nr = min(64, w - r0);
  for (int t = threadIdx.x; t < nr * nrhs; t += SNT) {
    const int gl = r0 + t % nr, c = t / nr;
    const int g = first + gl;
    const int64_t e = Fc.e_of_g[g];
    T acc = y[e * nrhs + c];
    for (int64_t q = S.gat_ptr[g]; q < S.gat_ptr[g + 1]; ++q) acc = acc - u[S.gat_idx[q] * nrhs + c];
    y[e * nrhs + c] = acc;
  }
}

// Solve the unit-lower 64-block at j0 in place for all right-hand sides: one
// wave per right-hand side, the block's strict lower part staged in LDS.
template <class T>
__device__ void fwd_diag_block(const T* P, int64_t h, const uint8_t* two, int eo, int j0, int nbk, int32_t nrhs,
                               T* y, T* Ld) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int t = threadIdx.x; t < TB * TB; t += SNT) {
    const int i = t % TB, p = t / TB;
    Ld[t] = (p < nbk && i < nbk && i > p) ? l_own(P, h, two, j0 + i, j0 + p) : F<T>::zero();
  }
  __syncthreads();
  for (int c = wv; c < nrhs; c += SNT / 64) {
    T yi = lane < nbk ? y[(int64_t)(eo + j0 + lane) * nrhs + c] : F<T>::zero();
    for (int p = 0; p < nbk; ++p) {
      const T yp = shfl_t(yi, p);
      yi = yi - Ld[p * TB + lane] * yp;
    }
    if (lane < nbk) y[(int64_t)(eo + j0 + lane) * nrhs + c] = yi;
  }
  __syncthreads();
}

// y_i -= sum_{p<nbk} L[i][j0+p] y_{j0+p} for own rows i in [r0, r1): thread per
// row, right-hand sides in chunks of RCH, the solved block's y staged in LDS.
template <class T>
__device__ void fwd_rows_update(const T* P, int64_t h, const uint8_t* two, int eo, int j0, int nbk, int r0,
                                int r1, int32_t nrhs, T* y, T* yb) {
  for (int c0 = 0; c0 < nrhs; c0 += RCH) {
    const int nc = min(RCH, nrhs - c0);
    for (int t = threadIdx.x; t < nbk * RCH; t += SNT) {
      const int p = t / RCH, k = t % RCH;
      yb[t] = k < nc ? y[(int64_t)(eo + j0 + p) * nrhs + c0 + k] : F<T>::zero();
    }
    __syncthreads();
    for (int i = r0 + threadIdx.x; i < r1; i += SNT) {
      T acc[RCH];
#pragma unroll
      for (int k = 0; k < RCH; ++k) acc[k] = F<T>::zero();
      for (int p = 0; p < nbk; ++p) {
        const T l = l_own(P, h, two, i, j0 + p);
#pragma unroll
        for (int k = 0; k < RCH; ++k) acc[k] = acc[k] + l * yb[p * RCH + k];
      }
      T* yi = y + (int64_t)(eo + i) * nrhs + c0;
      for (int k = 0; k < nc; ++k) yi[k] = yi[k] - acc[k];
    }
    __syncthreads();
  }
}

// Forward own-block solve of a narrow supernode (w <= 128): one workgroup.
template <class T>
__global__ __launch_bounds__(SNT) void k_fwd_tri(DevSym S, DevFac<T> Fc,
                                                 const int32_t* __restrict__ list, int32_t nrhs,
                                                 T* __restrict__ y) {
  __shared__ T Ld[TB * TB];
  __shared__ T yb[TB * RCH];
  const int s = list[blockIdx.x];
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const uint8_t* two = Fc.two + eo;
  for (int j0 = 0; j0 < w; j0 += TB) {
    const int nbk = min(TB, w - j0);
    fwd_diag_block(P, h, two, eo, j0, nbk, nrhs, y, Ld);
    if (j0 + nbk < w) fwd_rows_update(P, h, two, eo, j0, nbk, j0 + nbk, w, nrhs, y, yb);
  }
}

// Wide supernodes, forward step: the diagonal block j0 of each (one workgroup each)
template <class T>
__global__ __launch_bounds__(SNT) void k_ftri_diag(DevSym S, DevFac<T> Fc, const SItem* __restrict__ items,
                                                   int32_t nrhs, T* __restrict__ y) {
  __shared__ T Ld[TB * TB];
  const SItem it = items[blockIdx.x];
  const int s = it.s, j0 = it.j0;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  fwd_diag_block(Fc.pool + S.panel_off[s], (int64_t)S.sn_nrow[s], Fc.two + eo, eo, j0, min(TB, w - j0), nrhs,
                 y, Ld);
}

// ... then the 64-row chunks below it inside the own block
template <class T>
__global__ __launch_bounds__(SNT) void k_ftri_upd(DevSym S, DevFac<T> Fc, const SItem* __restrict__ items,
                                                  int32_t nrhs, T* __restrict__ y) {
  __shared__ T yb[TB * RCH];
  const SItem it = items[blockIdx.x];
  const int s = it.s, j0 = it.j0;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  fwd_rows_update(Fc.pool + S.panel_off[s], (int64_t)S.sn_nrow[s], Fc.two + eo, eo, j0, min(TB, w - j0), it.r0,
                  min(it.r0 + 64, w), nrhs, y, yb);
}

// Forward, part 2: u_s = L21 y_own for 64-row chunks of the off-diagonal rows
// (items are (supernode, row offset) pairs of one level).
template <class T>
__global__ __launch_bounds__(SNT) void k_fwd_update(DevSym S, DevFac<T> Fc,
                                                    const int2* __restrict__ items, int32_t nrhs,
                                                    const T* __restrict__ y, T* __restrict__ u) {
  __shared__ T part[4][64][8];
  const int2 it = items[blockIdx.x];
  const int s = it.x, r0 = it.y;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int64_t row = w + r0 + r;
  const bool live = row < h;
  const int64_t uo = S.u_off[s] + r0 + r;
  for (int c0 = 0; c0 < nrhs; c0 += 8) {
    const int nc = min(8, nrhs - c0);
    T acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = F<T>::zero();
    if (live) {
      for (int p = q; p < w; p += 4) {
        const T l = P[row + (int64_t)p * h];
        const T* yp = y + (int64_t)(eo + p) * nrhs + c0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (k < nc) acc[k] = acc[k] + l * yp[k];
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) part[q][r][k] = acc[k];
    __syncthreads();
    if (q == 0 && live) {
      for (int k = 0; k < nc; ++k) {
        T v = part[0][r][k] + part[1][r][k];
        v = v + part[2][r][k];
        v = v + part[3][r][k];
        u[uo * nrhs + c0 + k] = v;
      }
    }
    __syncthreads();
  }
}

// Backward, part 1: y_own[p] -= sum_{i>=w} L[i][p] x[rs[i]] for 16-column
// chunks (items (supernode, column offset)); one wave per column.
template <class T>
__global__ __launch_bounds__(SNT) void k_bwd_update(DevSym S, DevFac<T> Fc,
                                                    const int2* __restrict__ items, int32_t nrhs,
                                                    T* __restrict__ y) {
  const int2 it = items[blockIdx.x];
  const int s = it.x, p0 = it.y;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const int32_t* re = Fc.rs_e + S.rs_off[s];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int p = p0 + wv; p < min(p0 + 16, w); p += SNT / 64) {
    for (int c0 = 0; c0 < nrhs; c0 += RCH) {
      const int nc = min(RCH, nrhs - c0);
      T acc[RCH];
#pragma unroll
      for (int k = 0; k < RCH; ++k) acc[k] = F<T>::zero();
      for (int64_t i = w + lane; i < h; i += 64) {
        const T l = P[i + (int64_t)p * h];
        const T* xs = y + (int64_t)re[i] * nrhs + c0;
#pragma unroll
        for (int k = 0; k < RCH; ++k)
          if (k < nc) acc[k] = acc[k] + l * xs[k];
      }
#pragma unroll
      for (int k = 0; k < RCH; ++k) {
        T v = acc[k];
        for (int o = 32; o > 0; o >>= 1) v = v + shfl_xor_t(v, o);
        if (lane == 0 && k < nc) {
          const int64_t o2 = (int64_t)(eo + p) * nrhs + c0 + k;
          y[o2] = y[o2] - v;
        }
      }
    }
  }
}

// Back-substitute the transposed unit-lower 64-block at j0 in place, one
// wave per right-hand side (L^T of the block staged in LDS).
template <class T>
__device__ void bwd_diag_block(const T* P, int64_t h, const uint8_t* two, int eo, int j0, int nbk, int32_t nrhs,
                               T* y, T* Lt) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int t = threadIdx.x; t < TB * TB; t += SNT) {
    const int i = t % TB, p = t / TB;
    Lt[t] = (p < nbk && i < nbk && p > i) ? l_own(P, h, two, j0 + p, j0 + i) : F<T>::zero();
  }
  __syncthreads();
  for (int c = wv; c < nrhs; c += SNT / 64) {
    T yi = lane < nbk ? y[(int64_t)(eo + j0 + lane) * nrhs + c] : F<T>::zero();
    for (int p = nbk - 1; p >= 0; --p) {
      const T xp = shfl_t(yi, p);
      yi = yi - Lt[p * TB + lane] * xp;
    }
    if (lane < nbk) y[(int64_t)(eo + j0 + lane) * nrhs + c] = yi;
  }
  __syncthreads();
}

// Backward own-block solve of a narrow supernode (w <= 128): one workgroup
// back-substitutes L11^T x = y from the last 64-block up.
template <class T>
__global__ __launch_bounds__(SNT) void k_bwd_tri(DevSym S, DevFac<T> Fc,
                                                 const int32_t* __restrict__ list, int32_t nrhs,
                                                 T* __restrict__ y) {
  __shared__ T Lt[TB * TB];
  const int s = list[blockIdx.x];
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const uint8_t* two = Fc.two + eo;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nblk = (w + TB - 1) / TB;
  for (int b = nblk - 1; b >= 0; --b) {
    const int j0 = b * TB, nbk = min(TB, w - j0), jend = j0 + nbk;
    // contributions of the already solved own rows below this block
    if (jend < w) {
      for (int p = j0 + wv; p < jend; p += SNT / 64) {
        for (int c0 = 0; c0 < nrhs; c0 += RCH) {
          const int nc = min(RCH, nrhs - c0);
          T acc[RCH];
#pragma unroll
          for (int k = 0; k < RCH; ++k) acc[k] = F<T>::zero();
          for (int i = jend + lane; i < w; i += 64) {
            const T l = l_own(P, h, two, i, p);
            const T* xs = y + (int64_t)(eo + i) * nrhs + c0;
#pragma unroll
            for (int k = 0; k < RCH; ++k)
              if (k < nc) acc[k] = acc[k] + l * xs[k];
          }
#pragma unroll
          for (int k = 0; k < RCH; ++k) {
            T v = acc[k];
            for (int o = 32; o > 0; o >>= 1) v = v + shfl_xor_t(v, o);
            if (lane == 0 && k < nc) {
              const int64_t o2 = (int64_t)(eo + p) * nrhs + c0 + k;
              y[o2] = y[o2] - v;
            }
          }
        }
      }
      __syncthreads();
    }
    bwd_diag_block(P, h, two, eo, j0, nbk, nrhs, y, Lt);
  }
}

// Wide supernodes, backward step j0: partial sums over one 64-row chunk of the
// solved own rows below block j0, part[p][c] = sum_i L[i][j0+p] x_i,c.
template <class T>
__global__ __launch_bounds__(SNT) void k_btri_part(DevSym S, DevFac<T> Fc, const SItem* __restrict__ items,
                                                   int32_t nrhs, const T* __restrict__ y, T* __restrict__ part) {
  __shared__ T Ls[TB * (TB + 1)];
  __shared__ T xs[TB * RCH];
  const SItem it = items[blockIdx.x];
  const int s = it.s, j0 = it.j0;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const uint8_t* two = Fc.two + eo;
  const int nbk = min(TB, w - j0), r0 = it.r0, nr = min(64, w - r0);
  for (int t = threadIdx.x; t < TB * TB; t += SNT) {
    const int i = t % TB, p = t / TB;
    Ls[i * (TB + 1) + p] = (p < nbk && i < nr) ? l_own(P, h, two, r0 + i, j0 + p) : F<T>::zero();
  }
  T* out = part + (int64_t)blockIdx.x * TB * nrhs;
  for (int c0 = 0; c0 < nrhs; c0 += RCH) {
    const int nc = min(RCH, nrhs - c0);
    for (int t = threadIdx.x; t < TB * RCH; t += SNT) {
      const int i = t / RCH, k = t % RCH;
      xs[t] = (i < nr && k < nc) ? y[(int64_t)(eo + r0 + i) * nrhs + c0 + k] : F<T>::zero();
    }
    __syncthreads();
    for (int t = threadIdx.x; t < TB * RCH; t += SNT) {
      const int p = t / RCH, k = t % RCH;
      T acc = F<T>::zero();
      for (int i = 0; i < nr; ++i) acc = acc + Ls[i * (TB + 1) + p] * xs[i * RCH + k];
      if (k < nc) out[(int64_t)p * nrhs + c0 + k] = acc;
    }
    __syncthreads();
  }
}

// ... then each wide supernode's block j0 takes its partial sums (fixed
// order) and is back-substituted.
template <class T>
__global__ __launch_bounds__(SNT) void k_btri_diag(DevSym S, DevFac<T> Fc, const SItem* __restrict__ items,
                                                   int32_t nrhs, T* __restrict__ y, const T* __restrict__ part) {
  __shared__ T Lt[TB * TB];
  const SItem it = items[blockIdx.x];
  const int s = it.s, j0 = it.j0;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int nbk = min(TB, w - j0);
  for (int t = threadIdx.x; t < nbk * nrhs; t += SNT) {
    const int p = t / nrhs, c = t % nrhs;
    T acc = F<T>::zero();
    for (int q = 0; q < it.r0; ++q) acc = acc + part[((int64_t)(it.x + q) * TB + p) * nrhs + c];
    const int64_t o = (int64_t)(eo + j0 + p) * nrhs + c;
    y[o] = y[o] - acc;
  }
  __syncthreads();
  bwd_diag_block(Fc.pool + S.panel_off[s], (int64_t)S.sn_nrow[s], Fc.two + eo, eo, j0, nbk, nrhs, y, Lt);
}

template <class T>
__global__ void k_dsolve(int32_t n, int32_t nrhs, const T* __restrict__ d,
                         const T* __restrict__ ds, const uint8_t* __restrict__ two,
                         T* __restrict__ y) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * nrhs) return;
  int32_t e = (int32_t)(idx / nrhs), c = (int32_t)(idx % nrhs);
  if (two[e] && e + 1 < n) {
    T d11 = d[e], d21 = ds[e], d22 = d[e + 1];
    T detinv = F<T>::recip(d11 * d22 - d21 * d21);
    T z0 = y[(int64_t)e * nrhs + c], z1 = y[(int64_t)(e + 1) * nrhs + c];
    y[(int64_t)e * nrhs + c] = (d22 * z0 - d21 * z1) * detinv;
    y[(int64_t)(e + 1) * nrhs + c] = (d11 * z1 - d21 * z0) * detinv;
  } else if (!(e > 0 && two[e - 1])) {
    y[idx] = y[idx] * F<T>::recip(d[e]);
  }
}

// ---- launchers ----
template <class T>
void launch_gather_rhs(int32_t n, int32_t nrhs, const int32_t* perm_e, const double* scale,
                       const T* b, T* y, hipStream_t st, LaunchProfiler* pf) {
  int64_t tot = (int64_t)n * nrhs;
  if (tot <= 0) return;
  ProfScope ps(pf, LK_SOLVE_GATHER, 0.0, 2.0 * sizeof(T) * tot, st);
  hipLaunchKernelGGL(k_gather_rhs<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n,
                     nrhs, perm_e, scale, b, y);
}
template <class T>
void launch_scatter_sol(int32_t n, int32_t nrhs, const int32_t* perm_e, const double* scale,
                        const T* y, T* x, hipStream_t st, LaunchProfiler* pf) {
  int64_t tot = (int64_t)n * nrhs;
  if (tot <= 0) return;
  ProfScope ps(pf, LK_SOLVE_SCATTER, 0.0, 2.0 * sizeof(T) * tot, st);
  hipLaunchKernelGGL(k_scatter_sol<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n,
                     nrhs, perm_e, scale, y, x);
}
template <class T>
void launch_fgather(const DevSym& S, const DevFac<T>& Fc, const int2* items, int count, int32_t nrhs, T* y,
                    const T* u, hipStream_t st, LaunchProfiler* pf) {
  if (count <= 0) return;
  ProfScope ps(pf, LK_SOLVE_GATHER, 0.0, 0.0, st);
  hipLaunchKernelGGL(k_fgather<T>, dim3(count), dim3(SNT), 0, st, S, Fc, items, nrhs, y, u);
}
template <class T>
void launch_fwd_tri(const DevSym& S, const DevFac<T>& Fc, const int32_t* list, int count,
                    int32_t nrhs, T* y, double bytes, hipStream_t st, LaunchProfiler* pf) {
  if (count <= 0) return;
  ProfScope ps(pf, LK_FWD_TRI, 0.0, bytes, st);
  hipLaunchKernelGGL(k_fwd_tri<T>, dim3(count), dim3(SNT), 0, st, S, Fc, list, nrhs, y);
}
// one 64-column step of the wide supernodes: diagonal blocks, then the rows below
template <class T>
void launch_ftri_step(const DevSym& S, const DevFac<T>& Fc, const SItem* items, const SolveStep& stp,
                      int32_t nrhs, T* y, hipStream_t st, LaunchProfiler* pf) {
  ProfScope ps(pf, LK_FWD_TRI, 0.0, 0.0, st);
  if (stp.d_end > stp.d_beg)
    hipLaunchKernelGGL(k_ftri_diag<T>, dim3((unsigned)(stp.d_end - stp.d_beg)), dim3(SNT), 0, st, S, Fc,
                       items + stp.d_beg, nrhs, y);
  if (stp.u_end > stp.u_beg)
    hipLaunchKernelGGL(k_ftri_upd<T>, dim3((unsigned)(stp.u_end - stp.u_beg)), dim3(SNT), 0, st, S, Fc,
                       items + stp.u_beg, nrhs, y);
}
template <class T>
void launch_btri_step(const DevSym& S, const DevFac<T>& Fc, const SItem* items, const SolveStep& stp,
                      int32_t nrhs, T* y, T* part, hipStream_t st, LaunchProfiler* pf) {
  ProfScope ps(pf, LK_BWD_TRI, 0.0, 0.0, st);
  if (stp.u_end > stp.u_beg)
    hipLaunchKernelGGL(k_btri_part<T>, dim3((unsigned)(stp.u_end - stp.u_beg)), dim3(SNT), 0, st, S, Fc,
                       items + stp.u_beg, nrhs, y, part);
  if (stp.d_end > stp.d_beg)
    hipLaunchKernelGGL(k_btri_diag<T>, dim3((unsigned)(stp.d_end - stp.d_beg)), dim3(SNT), 0, st, S, Fc,
                       items + stp.d_beg, nrhs, y, part);
}
template <class T>
void launch_fwd_update(const DevSym& S, const DevFac<T>& Fc, const int2* items, int count,
                       int32_t nrhs, const T* y, T* u, double bytes, hipStream_t st, LaunchProfiler* pf) {
  if (count <= 0) return;
  ProfScope ps(pf, LK_FWD_UPDATE, 0.0, bytes, st);
  hipLaunchKernelGGL(k_fwd_update<T>, dim3(count), dim3(SNT), 0, st, S, Fc, items, nrhs, y, u);
}
template <class T>
void launch_dsolve(int32_t n, int32_t nrhs, const T* d, const T* ds, const uint8_t* two, T* y,
                   hipStream_t st, LaunchProfiler* pf) {
  int64_t tot = (int64_t)n * nrhs;
  if (tot <= 0) return;
  ProfScope ps(pf, LK_DSOLVE, 0.0, 2.0 * sizeof(T) * tot, st);
  hipLaunchKernelGGL(k_dsolve<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, nrhs,
                     d, ds, two, y);
}
template <class T>
void launch_bwd_update(const DevSym& S, const DevFac<T>& Fc, const int2* items, int count,
                       int32_t nrhs, T* y, double bytes, hipStream_t st, LaunchProfiler* pf) {
  if (count <= 0) return;
  ProfScope ps(pf, LK_BWD_UPDATE, 0.0, bytes, st);
  hipLaunchKernelGGL(k_bwd_update<T>, dim3(count), dim3(SNT), 0, st, S, Fc, items, nrhs, y);
}
template <class T>
void launch_bwd_tri(const DevSym& S, const DevFac<T>& Fc, const int32_t* list, int count,
                    int32_t nrhs, T* y, double bytes, hipStream_t st, LaunchProfiler* pf) {
  if (count <= 0) return;
  ProfScope ps(pf, LK_BWD_TRI, 0.0, bytes, st);
  hipLaunchKernelGGL(k_bwd_tri<T>, dim3(count), dim3(SNT), 0, st, S, Fc, list, nrhs, y);
}

#define RSL_SINST(T)                                                                           \
  template void launch_gather_rhs<T>(int32_t, int32_t, const int32_t*, const double*, const T*, \
                                     T*, hipStream_t, LaunchProfiler*);                        \
  template void launch_scatter_sol<T>(int32_t, int32_t, const int32_t*, const double*,          \
                                      const T*, T*, hipStream_t, LaunchProfiler*);             \
  template void launch_fgather<T>(const DevSym&, const DevFac<T>&, const int2*, int, int32_t, T*, \
                                  const T*, hipStream_t, LaunchProfiler*);                     \
  template void launch_fwd_tri<T>(const DevSym&, const DevFac<T>&, const int32_t*, int, int32_t, \
                                  T*, double, hipStream_t, LaunchProfiler*);                   \
  template void launch_ftri_step<T>(const DevSym&, const DevFac<T>&, const SItem*, const SolveStep&, \
                                    int32_t, T*, hipStream_t, LaunchProfiler*);                \
  template void launch_btri_step<T>(const DevSym&, const DevFac<T>&, const SItem*, const SolveStep&, \
                                    int32_t, T*, T*, hipStream_t, LaunchProfiler*);            \
  template void launch_fwd_update<T>(const DevSym&, const DevFac<T>&, const int2*, int, int32_t, \
                                     const T*, T*, double, hipStream_t, LaunchProfiler*);      \
  template void launch_dsolve<T>(int32_t, int32_t, const T*, const T*, const uint8_t*, T*,      \
                                 hipStream_t, LaunchProfiler*);                                \
  template void launch_bwd_update<T>(const DevSym&, const DevFac<T>&, const int2*, int, int32_t, \
                                     T*, double, hipStream_t, LaunchProfiler*);                \
  template void launch_bwd_tri<T>(const DevSym&, const DevFac<T>&, const int32_t*, int, int32_t, \
                                  T*, double, hipStream_t, LaunchProfiler*);

RSL_SINST(double)
RSL_SINST(cplx)

}  // namespace rsl
