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

#include <stdexcept>
#include <type_traits>

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

// ---------------------------------------------------------------------------
// Own-block and update building blocks. Every loop that reads global memory
// issues its loads as one batch into registers before using them, so a
// workgroup keeps many loads in flight instead of waiting on one at a time.
// ---------------------------------------------------------------------------
constexpr int TB = 64;   // triangular block
constexpr int NWV = SNT / 64;  // waves per workgroup

// lane p's value, made wave-uniform (p wave-uniform)
__device__ inline double lane_bcast(double v, int p) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), p);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), p);
  return __hiloint2double(hi, lo);
}
__device__ inline cplx lane_bcast(cplx v, int p) { return {lane_bcast(v.re, p), lane_bcast(v.im, p)}; }

// sum over the 64 lanes (DPP partial sums inside rows of 16, then across the
// rows); the total lands in lane 63. Fixed order: run-to-run identical.
template <int CTRL, int RM>
__device__ inline double dpp_take(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, RM, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, RM, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double wave_sum_to63(double v) {
  v = v + dpp_take<0x111, 0xF>(v);
  v = v + dpp_take<0x112, 0xF>(v);
  v = v + dpp_take<0x114, 0xF>(v);
  v = v + dpp_take<0x118, 0xF>(v);
  v = v + dpp_take<0x142, 0xA>(v);
  v = v + dpp_take<0x143, 0xC>(v);
  return v;
}
__device__ inline cplx wave_sum_to63(cplx v) { return {wave_sum_to63(v.re), wave_sum_to63(v.im)}; }

// Stage the strictly lower part of the own block's 64-block at j0 in LDS:
// Ld[p * TB + i] = L[j0+i][j0+p] for i > p, zero elsewhere and on the 2x2
// coupling entries (held in D, not in L).
template <class T>
__device__ void stage_block_lower(const T* P, int64_t h, const uint8_t* two, int j0, int nbk, T* Ld) {
  constexpr int PER = TB * TB / SNT;
  T v[PER];
  uint8_t tw[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * SNT, i = t % TB, p = t / TB;
    const bool in = p < nbk && i < nbk && i > p;
    v[u] = in ? P[(int64_t)(j0 + i) + (int64_t)(j0 + p) * h] : F<T>::zero();
    tw[u] = (in && i == p + 1) ? two[j0 + p] : 0;
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) Ld[threadIdx.x + u * SNT] = tw[u] ? F<T>::zero() : v[u];
  __syncthreads();
}

// Transposed: Lt[p * (TB+1) + i] = L[j0+p][j0+i] for p > i (read along the
// panel's columns, padded against LDS bank conflicts).
template <class T>
__device__ void stage_block_upper(const T* P, int64_t h, const uint8_t* two, int j0, int nbk, T* Lt) {
  constexpr int PER = TB * TB / SNT;
  T v[PER];
  uint8_t tw[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * SNT, p = t % TB, i = t / TB;
    const bool in = p < nbk && i < nbk && p > i;
    v[u] = in ? P[(int64_t)(j0 + p) + (int64_t)(j0 + i) * h] : F<T>::zero();
    tw[u] = (in && p == i + 1) ? two[j0 + i] : 0;
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * SNT, p = t % TB, i = t / TB;
    Lt[p * (TB + 1) + i] = tw[u] ? F<T>::zero() : v[u];
  }
  __syncthreads();
}

// Forward substitution with the staged block: one wave per right-hand side,
// two right-hand sides interleaved per wave; lane i holds y_i and takes
// y_p from lane p once it is final.
template <class T>
__device__ void fwd_block_solve(const T* Ld, int eo, int j0, int nbk, int32_t nrhs, T* y) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t row = eo + j0 + lane;
  for (int c = wv; c < nrhs; c += 2 * NWV) {
    const int c2 = c + NWV;
    const bool two_rhs = c2 < nrhs;
    T y0 = lane < nbk ? y[row * nrhs + c] : F<T>::zero();
    T y1 = (two_rhs && lane < nbk) ? y[row * nrhs + c2] : F<T>::zero();
#pragma unroll 8
    for (int p = 0; p + 1 < nbk; ++p) {
      const T l = Ld[p * TB + lane];
      const T a0 = lane_bcast(y0, p), a1 = lane_bcast(y1, p);
      y0 = y0 - l * a0;
      y1 = y1 - l * a1;
    }
    if (lane < nbk) {
      y[row * nrhs + c] = y0;
      if (two_rhs) y[row * nrhs + c2] = y1;
    }
  }
}

// Back substitution with L^T of the staged block (Lt from stage_block_upper).
template <class T>
__device__ void bwd_block_solve(const T* Lt, int eo, int j0, int nbk, int32_t nrhs, T* y) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t row = eo + j0 + lane;
  for (int c = wv; c < nrhs; c += 2 * NWV) {
    const int c2 = c + NWV;
    const bool two_rhs = c2 < nrhs;
    T y0 = lane < nbk ? y[row * nrhs + c] : F<T>::zero();
    T y1 = (two_rhs && lane < nbk) ? y[row * nrhs + c2] : F<T>::zero();
#pragma unroll 8
    for (int p = nbk - 1; p > 0; --p) {
      const T l = Lt[p * (TB + 1) + lane];
      const T a0 = lane_bcast(y0, p), a1 = lane_bcast(y1, p);
      y0 = y0 - l * a0;
      y1 = y1 - l * a1;
    }
    if (lane < nbk) {
      y[row * nrhs + c] = y0;
      if (two_rhs) y[row * nrhs + c2] = y1;
    }
  }
}

// red[r][k] = sum_{p in [c_lo, c_hi)} L[row0 + r][p] * y[(eo + p) * nrhs + c0 + k]
// for r < nr (<= 64), k < nc (<= RCH). Thread (r, q): columns c_lo + q + 4m,
// 8 of them per batch of 32; the batch's y rows are staged in LDS. zr/zc name one
// entry (relative row, absolute column) read as zero: a 2x2 coupling entry
// that falls outside the diagonal block, or -1.
template <class T>
__device__ void rows_times_y(const T* P, int64_t h, int64_t row0, int nr, int c_lo, int c_hi, int eo,
                             int32_t nrhs, int c0, int nc, const T* __restrict__ y, int zr, int zc, T* ys,
                             T* red) {
  const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
  constexpr int CB = 32;  // columns per batch
  T acc[RCH];
#pragma unroll
  for (int k = 0; k < RCH; ++k) acc[k] = F<T>::zero();
  for (int pb = c_lo; pb < c_hi; pb += CB) {
    const int ncb = min(CB, c_hi - pb);
    // y rows of the batch: CB x RCH values, one per thread
    {
      const int t = threadIdx.x, p = t / RCH, k = t % RCH;
      ys[t] = (p < ncb && k < nc) ? y[(int64_t)(eo + pb + p) * nrhs + c0 + k] : F<T>::zero();
    }
    T l[CB / NWV];
#pragma unroll
    for (int m = 0; m < CB / NWV; ++m) {
      const int p = q + NWV * m;
      l[m] = (r < nr && p < ncb) ? P[(row0 + r) + (int64_t)(pb + p) * h] : F<T>::zero();
      if (r == zr && pb + p == zc) l[m] = F<T>::zero();
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < CB / NWV; ++m) {
      const T* yp = ys + (q + NWV * m) * RCH;
#pragma unroll
      for (int k = 0; k < RCH; ++k) acc[k] = acc[k] + l[m] * yp[k];
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < RCH; ++k) red[(q * 64 + r) * RCH + k] = acc[k];
  __syncthreads();
  // fixed-order sum over the four column groups into group 0's slots
  for (int t = threadIdx.x; t < 64 * RCH; t += SNT) {
    T v = red[t];
    v = v + red[64 * RCH + t];
    v = v + red[2 * 64 * RCH + t];
    v = v + red[3 * 64 * RCH + t];
    red[t] = v;
  }
  __syncthreads();
}

// part[p][k] = sum_{i < nr} L[r0 + i][j0 + p] * x[(eo + r0 + i) * nrhs + c0 + k]
// (transposed product of a 64-row chunk of the own block), p < nbk, k < nc:
// L chunk and x rows staged in LDS, thread (p, k-pair) sums 64 terms.
template <class T>
__device__ void chunk_t_times_x(const T* P, int64_t h, const uint8_t* two, int eo, int j0, int nbk, int r0, int nr,
                                int32_t nrhs, int c0, int nc, const T* __restrict__ y, T* Ls, T* xs, T* out) {
  constexpr int PER = TB * TB / SNT;
  T v[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * SNT, i = t % TB, p = t / TB;
    v[u] = (p < nbk && i < nr) ? P[(int64_t)(r0 + i) + (int64_t)(j0 + p) * h] : F<T>::zero();
  }
  T xv[TB * RCH / SNT];
#pragma unroll
  for (int u = 0; u < TB * RCH / SNT; ++u) {
    const int t = threadIdx.x + u * SNT, i = t / RCH, k = t % RCH;
    xv[u] = (i < nr && k < nc) ? y[(int64_t)(eo + r0 + i) * nrhs + c0 + k] : F<T>::zero();
  }
  // a 2x2 coupling entry (j0+p+1, j0+p) can only sit in this chunk at its
  // first row, when the chunk starts right below the block
  const bool zc = r0 == j0 + nbk && nbk > 0 && two[j0 + nbk - 1];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * SNT, i = t % TB, p = t / TB;
    Ls[i * (TB + 1) + p] = (zc && i == 0 && p == nbk - 1) ? F<T>::zero() : v[u];
  }
#pragma unroll
  for (int u = 0; u < TB * RCH / SNT; ++u) xs[threadIdx.x + u * SNT] = xv[u];
  __syncthreads();
  const int p = threadIdx.x & 63, kq = threadIdx.x >> 6;
  T a0 = F<T>::zero(), a1 = F<T>::zero();
#pragma unroll 8
  for (int i = 0; i < TB; ++i) {
    const T l = Ls[i * (TB + 1) + p];
    a0 = a0 + l * xs[i * RCH + kq];
    a1 = a1 + l * xs[i * RCH + kq + NWV];
  }
  out[p * RCH + kq] = a0;
  out[p * RCH + kq + NWV] = a1;
  __syncthreads();
}

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
    int64_t q = S.gat_ptr[g];
    const int64_t qe = S.gat_ptr[g + 1];
    for (; q + 4 <= qe; q += 4) {
      int64_t ix[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) ix[m] = S.gat_idx[q + m];
      T v[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) v[m] = u[ix[m] * nrhs + c];
#pragma unroll
      for (int m = 0; m < 4; ++m) acc = acc - v[m];
    }
    for (; q < qe; ++q) acc = acc - u[S.gat_idx[q] * nrhs + c];
    y[e * nrhs + c] = acc;
  }
}

// Forward own-block solve of a narrow supernode (w <= 32): one workgroup.
template <class T>
__global__ __launch_bounds__(SNT) void k_fwd_tri(DevSym S, DevFac<T> Fc,
                                                 const int32_t* __restrict__ list, int32_t nrhs,
                                                 T* __restrict__ y) {
  __shared__ T Ld[TB * TB];
  __shared__ T ys[64 * RCH];
  __shared__ T red[4 * 64 * RCH];
  const int s = list[blockIdx.x];
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const uint8_t* two = Fc.two + eo;
  for (int j0 = 0; j0 < w; j0 += TB) {
    const int nbk = min(TB, w - j0);
    stage_block_lower(P, h, two, j0, nbk, Ld);
    fwd_block_solve(Ld, eo, j0, nbk, nrhs, y);
    __syncthreads();
    const int rb = j0 + nbk;
    for (int r0 = rb; r0 < w; r0 += 64) {
      const int nr = min(64, w - r0);
      const int zr = (r0 == rb && two[rb - 1]) ? 0 : -1;
      for (int c0 = 0; c0 < nrhs; c0 += RCH) {
        const int nc = min(RCH, nrhs - c0);
        rows_times_y(P, h, r0, nr, j0, rb, eo, nrhs, c0, nc, y, zr, rb - 1, ys, red);
        for (int t = threadIdx.x; t < nr * RCH; t += SNT) {
          const int rr = t / RCH, k = t % RCH;
          if (k < nc) {
            const int64_t o = (int64_t)(eo + r0 + rr) * nrhs + c0 + k;
            y[o] = y[o] - red[rr * RCH + k];
          }
        }
        __syncthreads();
      }
    }
  }
}

// Forward, part 2: u_s = L21 y_own for 64-row chunks of the off-diagonal rows
// (items are (supernode, row offset) pairs of one level).
template <class T>
__global__ __launch_bounds__(SNT) void k_fwd_update(DevSym S, DevFac<T> Fc,
                                                    const int2* __restrict__ items, int32_t nrhs,
                                                    const T* __restrict__ y, T* __restrict__ u) {
  __shared__ T ys[64 * RCH];
  __shared__ T red[4 * 64 * RCH];
  const int2 it = items[blockIdx.x];
  const int s = it.x, r0 = it.y;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const int nr = (int)min<int64_t>(64, h - w - r0);
  const int64_t uo = S.u_off[s] + r0;
  for (int c0 = 0; c0 < nrhs; c0 += RCH) {
    const int nc = min(RCH, nrhs - c0);
    rows_times_y(Fc.pool + S.panel_off[s], h, (int64_t)w + r0, nr, 0, w, eo, nrhs, c0, nc, y, -1, -1, ys, red);
    for (int t = threadIdx.x; t < nr * RCH; t += SNT) {
      const int rr = t / RCH, k = t % RCH;
      if (k < nc) u[(uo + rr) * nrhs + c0 + k] = red[rr * RCH + k];
    }
    __syncthreads();
  }
}

// y_own[p] -= sum_{i in [ib0, ie)} L[i][p] x[rs[i]] for the 16 columns p0..
// of supernode s (w <= ib0 <= ie <= h; k_bwd_update's unit takes every
// off-diagonal row); with part set the sums are stored to part[p * nrhs + k]
// instead. xs holds RB * RCH values of LDS, is RB row indices. A pass's x
// rows are gathered eight lanes to a row (its RCH values, one 64-byte line),
// their indices shared through is.
constexpr int BW_RB = 512, BW_CW = 4;  // rows per pass, columns per wave
static_assert(RCH == 8, "a gathered x row is one lane octet");
template <class T>
__device__ void bwd_update_cols(const DevSym& S, const DevFac<T>& Fc, int s, int p0, int32_t nrhs,
                                T* __restrict__ y, T* xs, int32_t* is, int64_t ib0, int64_t ie,
                                T* __restrict__ part) {
  constexpr int RB = BW_RB, CW = BW_CW;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const int32_t* re = Fc.rs_e + S.rs_off[s];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int pw = p0 + wv * CW;  // this wave's first column
  for (int c0 = 0; c0 < nrhs; c0 += RCH) {
    const int nc = min(RCH, nrhs - c0);
    T acc[CW][RCH];
#pragma unroll
    for (int a = 0; a < CW; ++a)
#pragma unroll
      for (int k = 0; k < RCH; ++k) acc[a][k] = F<T>::zero();
    // row indices of the next pass, loaded one pass ahead (real field; the
    // complex field's registers hold no second set): a pass then waits on
    // one round of loads (its x rows and L entries together)
    constexpr bool AHEAD = sizeof(T) == sizeof(double);
    int32_t en[RB / SNT];
    auto load_idx = [&](int64_t ib) {
#pragma unroll
      for (int q = 0; q < RB / SNT; ++q) {
        const int64_t r = ib + threadIdx.x + q * SNT;
        en[q] = r < ie ? re[r] : -1;
      }
    };
    if (AHEAD) load_idx(ib0);
    for (int64_t ib = ib0; ib < ie; ib += RB) {
      const int nr = (int)min<int64_t>(RB, ie - ib);
      if (!AHEAD) load_idx(ib);
#pragma unroll
      for (int q = 0; q < RB / SNT; ++q) is[threadIdx.x + q * SNT] = en[q];
      __syncthreads();
      // x rows of the pass: thread t takes row t / 8 + 32 q, value t % 8
      constexpr int XR = RB * RCH / SNT;
      const int xk = threadIdx.x & 7, xr0 = threadIdx.x >> 3;
      T xv[XR];
#pragma unroll
      for (int q = 0; q < XR; ++q) {
        const int32_t ee = is[xr0 + 32 * q];
        xv[q] = (ee >= 0 && xk < nc) ? y[(int64_t)ee * nrhs + c0 + xk] : F<T>::zero();
      }
      auto stage_x = [&]() {
#pragma unroll
        for (int q = 0; q < XR; ++q) xs[xk * RB + xr0 + 32 * q] = xv[q];  // [k][row]
      };
      // complex field: x staged before the L loads (registers)
      if (!AHEAD) stage_x();
      T l[RB / 64][CW];
#pragma unroll
      for (int sb = 0; sb < RB / 64; ++sb)
#pragma unroll
        for (int a = 0; a < CW; ++a) {
          const int ir = sb * 64 + lane, p = pw + a;
          l[sb][a] = (ir < nr && p < w) ? P[(ib + ir) + (int64_t)p * h] : F<T>::zero();
        }
      if (AHEAD && ib + RB < ie) load_idx(ib + RB);
      if (AHEAD) stage_x();
      __syncthreads();
#pragma unroll
      for (int sb = 0; sb < RB / 64; ++sb) {
        T xr[RCH];
#pragma unroll
        for (int k = 0; k < RCH; ++k) xr[k] = xs[k * RB + sb * 64 + lane];
#pragma unroll
        for (int a = 0; a < CW; ++a)
#pragma unroll
          for (int k = 0; k < RCH; ++k) acc[a][k] = acc[a][k] + l[sb][a] * xr[k];
      }
      __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < CW; ++a)
#pragma unroll
      for (int k = 0; k < RCH; ++k) {
        const T v = wave_sum_to63(acc[a][k]);
        if (lane == 63 && k < nc && pw + a < w && pw + a < p0 + 16) {
          if (part) {
            part[(int64_t)(pw + a) * nrhs + c0 + k] = v;
          } else {
            const int64_t o2 = (int64_t)(eo + pw + a) * nrhs + c0 + k;
            y[o2] = y[o2] - v;
          }
        }
      }
  }
}

// Backward, part 1: y_own[p] -= sum_{i>=w} L[i][p] x[rs[i]] for 16-column
// chunks (items (supernode, column offset)). Lanes run over rows, 512 rows
// per pass (their x rows gathered into LDS; each pass costs a dependent
// gather, so few large passes), each wave owns four columns; the lane
// partial sums (lane l: rows = l mod 64, ascending) are reduced once at the
// end.
template <class T>
__global__ __launch_bounds__(SNT) void k_bwd_update(DevSym S, DevFac<T> Fc,
                                                    const int2* __restrict__ items, int32_t nrhs,
                                                    T* __restrict__ y) {
  __shared__ T xs[BW_RB * RCH];
  __shared__ int32_t is[BW_RB];
  const int2 it = items[blockIdx.x];
  bwd_update_cols(S, Fc, it.x, it.y, nrhs, y, xs, is, S.sn_ncol[it.x], S.sn_nrow[it.x], (T*)nullptr);
}

// Backward own-block solve of a narrow supernode (w <= 32): one workgroup
// back-substitutes L11^T x = y from the last 64-block up.
template <class T>
__global__ __launch_bounds__(SNT) void k_bwd_tri(DevSym S, DevFac<T> Fc,
                                                 const int32_t* __restrict__ list, int32_t nrhs,
                                                 T* __restrict__ y) {
  __shared__ T Lt[TB * (TB + 1)];
  __shared__ T xs[TB * RCH];
  __shared__ T out[TB * RCH];
  const int s = list[blockIdx.x];
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const uint8_t* two = Fc.two + eo;
  const int nblk = (w + TB - 1) / TB;
  for (int b = nblk - 1; b >= 0; --b) {
    const int j0 = b * TB, nbk = min(TB, w - j0), jend = j0 + nbk;
    // contributions of the already solved own rows below this block
    for (int r0 = jend; r0 < w; r0 += 64) {
      const int nr = min(64, w - r0);
      for (int c0 = 0; c0 < nrhs; c0 += RCH) {
        const int nc = min(RCH, nrhs - c0);
        chunk_t_times_x(P, h, two, eo, j0, nbk, r0, nr, nrhs, c0, nc, y, Lt, xs, out);
        for (int t = threadIdx.x; t < nbk * RCH; t += SNT) {
          const int p = t / RCH, k = t % RCH;
          if (k < nc) {
            const int64_t o = (int64_t)(eo + j0 + p) * nrhs + c0 + k;
            y[o] = y[o] - out[t];
          }
        }
        __syncthreads();
      }
    }
    stage_block_upper(P, h, two, j0, nbk, Lt);
    bwd_block_solve(Lt, eo, j0, nbk, nrhs, y);
    __syncthreads();
  }
}

// ---- wide supernodes: one launch per 64-column block step ----
// Every unit of a step solves the step's diagonal block itself (its y rows
// are read only within the launch) into LDS; the diagonal unit stores it in z,
// and the other units apply it: forward to 64 own rows below the block,
// backward (push form) to an earlier own block. z is copied back to y once
// the level's steps are done.

// A step issues every global load of a right-hand-side chunk (the diagonal
// block, its y entries, the unit's 64 x 64 L operand and the y entries it
// updates) before using any of them, so one memory latency covers the step
// instead of one per phase. Arithmetic per unit: the block solve is
// fwd/bwd_block_solve's (lane i holds y_i, one pair of right-hand sides per
// wave); a forward row unit sums L[r][j0 + p] * y_p over 32-column batches
// (thread (r, q) on columns q + 4m, then a fixed-order sum over q), as
// rows_times_y; a backward push unit sums L[j0 + i][i0 + p] * x_i over the
// block's 64 rows in order, as chunk_t_times_x.
static_assert(RCH == 2 * NWV, "one pair of right-hand sides per wave and chunk");
constexpr int TPER = TB * TB / SNT;  // 64 x 64 block entries per thread

// forward step: units [0, nu) are row units (s, j0, r0), the rest diagonal units
template <class T>
__global__ __launch_bounds__(SNT) void k_ftri_step(DevSym S, DevFac<T> Fc, const SItem* __restrict__ items, int nu,
                                                   int32_t nrhs, T* __restrict__ y, T* __restrict__ z) {
  __shared__ T Ld[TB * TB];
  __shared__ T sb[TB * RCH];
  __shared__ T red[4 * 64 * RCH];
  const SItem it = items[blockIdx.x];
  const bool diag = (int)blockIdx.x >= nu;
  const int s = it.s, j0 = it.j0;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const uint8_t* two = Fc.two + eo;
  const int nbk = min(TB, w - j0), rb = j0 + nbk;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // the block's strict lower part (stage_block_lower's entries)
  T vst[TPER];
#pragma unroll
  for (int u = 0; u < TPER; ++u) {
    const int t = tid + u * SNT, i = t % TB, p = t / TB;
    const bool in = p < nbk && i < nbk && i > p && !(i == p + 1 && two[j0 + p]);
    vst[u] = in ? P[(int64_t)(j0 + i) + (int64_t)(j0 + p) * h] : F<T>::zero();
  }
  // row unit: L[r0 + r][j0 + p] for r = lane, p = pb + wv + NWV m (rows_times_blk's operand)
  const int r0 = diag ? 0 : it.r0, nr = diag ? 0 : min(64, w - r0);
  const int zr = (!diag && r0 == rb && two[rb - 1]) ? 0 : -1;
  T lr[TPER];
#pragma unroll
  for (int u = 0; u < TPER; ++u) {
    const int p = (u / 8) * 32 + wv + NWV * (u % 8);
    lr[u] = (lane < nr && p < nbk) ? P[(int64_t)(r0 + lane) + (int64_t)(j0 + p) * h] : F<T>::zero();
    if (lane == zr && j0 + p == rb - 1) lr[u] = F<T>::zero();
  }
  for (int c0 = 0; c0 < nrhs; c0 += RCH) {
    const int nc = min(RCH, nrhs - c0);
    const int k = wv, k2 = wv + NWV;
    const int64_t brow = eo + j0 + lane;
    T y0 = (k < nc && lane < nbk) ? y[brow * nrhs + c0 + k] : F<T>::zero();
    T y1 = (k2 < nc && lane < nbk) ? y[brow * nrhs + c0 + k2] : F<T>::zero();
    T yr[RCH * 64 / SNT];
#pragma unroll
    for (int u = 0; u < RCH * 64 / SNT; ++u) {
      const int t = tid + u * SNT, rr = t / RCH, kk = t % RCH;
      yr[u] = (rr < nr && kk < nc) ? y[(int64_t)(eo + r0 + rr) * nrhs + c0 + kk] : F<T>::zero();
    }
    if (c0 == 0) {
#pragma unroll
      for (int u = 0; u < TPER; ++u) Ld[tid + u * SNT] = vst[u];
    }
    for (int t = tid; t < TB * RCH; t += SNT) sb[t] = F<T>::zero();
    __syncthreads();
    if (k < nc) {
#pragma unroll 8
      for (int p = 0; p + 1 < nbk; ++p) {
        const T l = Ld[p * TB + lane];
        const T a0 = lane_bcast(y0, p), a1 = lane_bcast(y1, p);
        y0 = y0 - l * a0;
        y1 = y1 - l * a1;
      }
      sb[lane * RCH + k] = y0;
      if (k2 < nc) sb[lane * RCH + k2] = y1;
    }
    __syncthreads();
    if (diag) {
      for (int t = tid; t < nbk * RCH; t += SNT) {
        const int p = t / RCH, kk = t % RCH;
        if (kk < nc) z[(int64_t)(eo + j0 + p) * nrhs + c0 + kk] = sb[t];
      }
    } else {
      T acc[RCH];
#pragma unroll
      for (int kk = 0; kk < RCH; ++kk) acc[kk] = F<T>::zero();
#pragma unroll
      for (int u = 0; u < TPER; ++u) {
        const int pb = (u / 8) * 32, p = pb + wv + NWV * (u % 8);
        if (pb < nbk && p < nbk) {
          const T* yp = sb + p * RCH;
#pragma unroll
          for (int kk = 0; kk < RCH; ++kk) acc[kk] = acc[kk] + lr[u] * yp[kk];
        }
      }
#pragma unroll
      for (int kk = 0; kk < RCH; ++kk) red[(wv * 64 + lane) * RCH + kk] = acc[kk];
      __syncthreads();
      for (int t = tid; t < 64 * RCH; t += SNT) {
        T v = red[t];
        v = v + red[64 * RCH + t];
        v = v + red[2 * 64 * RCH + t];
        v = v + red[3 * 64 * RCH + t];
        red[t] = v;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < RCH * 64 / SNT; ++u) {
        const int t = tid + u * SNT, rr = t / RCH, kk = t % RCH;
        if (rr < nr && kk < nc) y[(int64_t)(eo + r0 + rr) * nrhs + c0 + kk] = yr[u] - red[rr * RCH + kk];
      }
    }
    __syncthreads();
  }
}

// backward step (last block first): units [0, nd) are diagonal units, the
// rest push units (s, j0, i0)
template <class T>
__global__ __launch_bounds__(SNT) void k_btri_step(DevSym S, DevFac<T> Fc, const SItem* __restrict__ items, int nd,
                                                   int32_t nrhs, T* __restrict__ y, T* __restrict__ z) {
  __shared__ T Lt[TB * (TB + 1)];
  __shared__ T sb[TB * RCH];
  __shared__ T out[TB * RCH];
  const SItem it = items[blockIdx.x];
  const bool diag = (int)blockIdx.x < nd;
  const int s = it.s, j0 = it.j0;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const uint8_t* two = Fc.two + eo;
  const int nbk = min(TB, w - j0);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // the block's transposed strict lower part (stage_block_upper's entries)
  T vst[TPER];
#pragma unroll
  for (int u = 0; u < TPER; ++u) {
    const int t = tid + u * SNT, p = t % TB, i = t / TB;
    const bool in = p < nbk && i < nbk && p > i && !(p == i + 1 && two[j0 + i]);
    vst[u] = in ? P[(int64_t)(j0 + p) + (int64_t)(j0 + i) * h] : F<T>::zero();
  }
  // push unit: L[j0 + i][i0 + p] (blk_t_times_sb's operand, rows j0.. of block i0's columns)
  const int i0 = diag ? 0 : it.r0;
  const bool zc = !diag && j0 == i0 + TB && two[i0 + TB - 1];
  T vp[TPER];
#pragma unroll
  for (int u = 0; u < TPER; ++u) {
    const int t = tid + u * SNT, i = t % TB, p = t / TB;
    vp[u] = (!diag && i < nbk) ? P[(int64_t)(j0 + i) + (int64_t)(i0 + p) * h] : F<T>::zero();
    if (zc && i == 0 && p == TB - 1) vp[u] = F<T>::zero();
  }
  for (int c0 = 0; c0 < nrhs; c0 += RCH) {
    const int nc = min(RCH, nrhs - c0);
    const int k = wv, k2 = wv + NWV;
    const int64_t brow = eo + j0 + lane;
    T y0 = (k < nc && lane < nbk) ? y[brow * nrhs + c0 + k] : F<T>::zero();
    T y1 = (k2 < nc && lane < nbk) ? y[brow * nrhs + c0 + k2] : F<T>::zero();
    T yt[RCH * TB / SNT];
#pragma unroll
    for (int u = 0; u < RCH * TB / SNT; ++u) {
      const int t = tid + u * SNT, p = t / RCH, kk = t % RCH;
      yt[u] = (!diag && kk < nc) ? y[(int64_t)(eo + i0 + p) * nrhs + c0 + kk] : F<T>::zero();
    }
#pragma unroll
    for (int u = 0; u < TPER; ++u) {
      const int t = tid + u * SNT, p = t % TB, i = t / TB;
      Lt[p * (TB + 1) + i] = vst[u];
    }
    for (int t = tid; t < TB * RCH; t += SNT) sb[t] = F<T>::zero();
    __syncthreads();
    if (k < nc) {
#pragma unroll 8
      for (int p = nbk - 1; p > 0; --p) {
        const T l = Lt[p * (TB + 1) + lane];
        const T a0 = lane_bcast(y0, p), a1 = lane_bcast(y1, p);
        y0 = y0 - l * a0;
        y1 = y1 - l * a1;
      }
      sb[lane * RCH + k] = y0;
      if (k2 < nc) sb[lane * RCH + k2] = y1;
    }
    __syncthreads();
    if (diag) {
      for (int t = tid; t < nbk * RCH; t += SNT) {
        const int p = t / RCH, kk = t % RCH;
        if (kk < nc) z[(int64_t)(eo + j0 + p) * nrhs + c0 + kk] = sb[t];
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int u = 0; u < TPER; ++u) {
        const int t = tid + u * SNT, i = t % TB, p = t / TB;
        Lt[i * (TB + 1) + p] = vp[u];
      }
      __syncthreads();
      const int p = lane, kq = wv;
      T a0 = F<T>::zero(), a1 = F<T>::zero();
#pragma unroll 8
      for (int i = 0; i < TB; ++i) {
        const T l = Lt[i * (TB + 1) + p];
        a0 = a0 + l * sb[i * RCH + kq];
        a1 = a1 + l * sb[i * RCH + kq + NWV];
      }
      out[p * RCH + kq] = a0;
      out[p * RCH + kq + NWV] = a1;
      __syncthreads();
#pragma unroll
      for (int u = 0; u < RCH * TB / SNT; ++u) {
        const int t = tid + u * SNT, pp = t / RCH, kk = t % RCH;
        if (kk < nc) y[(int64_t)(eo + i0 + pp) * nrhs + c0 + kk] = yt[u] - out[t];
      }
      __syncthreads();
    }
  }
}

// y := z on the own blocks of a level's wide supernodes
template <class T>
__global__ __launch_bounds__(SNT) void k_zcopy(DevSym S, const SItem* __restrict__ items, int32_t nrhs,
                                               const T* __restrict__ z, T* __restrict__ y) {
  const SItem it = items[blockIdx.x];
  const int w = S.sn_ncol[it.s], eo = S.e_off[it.s];
  const int nbk = min(TB, w - it.j0);
  const int64_t base = (int64_t)(eo + it.j0) * nrhs, n = (int64_t)nbk * nrhs;
  for (int64_t t = threadIdx.x; t < n; t += SNT) y[base + t] = z[base + t];
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

// ---------------------------------------------------------------------------
// Dependency-driven sweeps: one launch per sweep, one workgroup per unit.
// A unit's workgroup takes a ticket (an atomic counter, so units start in
// plan order whatever order the hardware dispatches workgroups in) and waits
// only on units with earlier tickets, which belong to workgroups that are
// already running or finished: the launch makes progress whatever the
// residency. A producer publishes a counter with an agent-scope release after
// its stores; a consumer polls it and acquires before reading.
//   forward unit (s, c), c < nblk: own 64-row block c -- gather the
//     descendants' products (waits for every updater's off-diagonal chunks),
//     subtract L[c, j] y_j for the earlier blocks j (waiting for each), solve
//     the unit-lower diagonal block, publish its block counter = c + 1;
//   forward unit (s, c), c >= nblk: 64 off-diagonal rows of s, u = L21 y_own
//     over every own block, publish by its chunk counter += 1;
//   backward off unit (s, -1 - q): once every ancestor owning the
//     off-diagonal rows is solved, subtract L21^T x from the own entries of
//     the 16 columns 16q.. (one writer per entry), publish by its counter += 1;
//   backward own unit (s, c): own 64-column block c -- once the off units of
//     s are done, subtract L[j, c]^T x_j for the later own blocks j (last
//     first, waiting for each), back-substitute with the diagonal block,
//     publish its backward counter = nblk - c.
// Sums run in a fixed order per output entry: run-to-run identical.
// ---------------------------------------------------------------------------
constexpr uint64_t kSpinTicks = 200000000ull;  // 2 s of the 100 MHz constant clock
// counters of supernode s: cnt[s * CSTRIDE + kind], one 128-byte line per supernode
constexpr int CSTRIDE = 32;
constexpr int CF_BLK = 0, CF_U = 1, CB_BLK = 2, CB_OFF = 3, CF_H = 4;

__device__ inline int32_t* counter(const SolveSync& sy, int s, int kind) {
  return sy.cnt + (int64_t)s * CSTRIDE + kind;
}

// Poll *p until it reaches target and return the value read. Returns target
// at once when the solve is aborted: this wait or an earlier one of the same
// solve ran past kSpinTicks. A poll is one load; a counter more than one
// step short of its target is polled about every 0.8 us (hundreds of
// waiting workgroups polling at full rate slow every poll down), one step
// short about every 30 ns. The abort word and the clock are read every 32nd
// poll.
__device__ inline int32_t wait_ge(int32_t* p, int32_t target, uint32_t* ctl) {
  int32_t v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (v >= target) return v;
  const uint64_t t0 = (uint64_t)wall_clock64();
  for (uint32_t it = 1;; ++it) {
    if (target - v > 1) __builtin_amdgcn_s_sleep(31);
    else __builtin_amdgcn_s_sleep(1);
    v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v >= target) return v;
    if ((it & 31u) == 0) {
      if (__hip_atomic_load(ctl + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return target;
      if ((uint64_t)wall_clock64() - t0 > kSpinTicks) {
        __hip_atomic_store(ctl + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(ctl + 3, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return target;
      }
    }
  }
}

// The workgroup's stores, then the counter. Every wave waits for its own
// stores to complete before the barrier (the barrier does not wait for
// them); then thread 0 writes the L2 back at agent scope and waits for that
// write-back (an explicit wait: the fence alone is not followed by one when
// an atomic comes next) before it stores / adds the counter. s_waitcnt
// immediate: vmcnt 0, expcnt and lgkmcnt not waited on.
__device__ inline void wg_publish(int32_t* p, int32_t v, bool add) {
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_s_waitcnt(0x0F70);
    if (add) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__device__ inline int take_ticket(uint32_t* ctr, int* slot) {
  if (threadIdx.x == 0) *slot = (int)__hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return *slot;
}

// Counters at the start of a solve: zero, except that the off-diagonal chunk
// counter of a supernode solved before the dependency-driven launch (below the
// split level) already holds its full count; control words zero (the sticky
// timeout count, ctl[3], is kept).
__global__ void k_sync_init(int32_t* __restrict__ cnt, const int32_t* __restrict__ fu_init, int32_t ns) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)ns * CSTRIDE;
  if (i < n) cnt[i] = (i % CSTRIDE == CF_U) ? fu_init[i / CSTRIDE] : 0;
  else if (i < n + 3) cnt[i] = 0;
}

// ---------------------------------------------------------------------------
// Forward chain unit (real field; the dependency-driven solve runs only for
// up to RCH right-hand sides): one workgroup solves every own 64-row block of
// supernode s in order, so consecutive blocks pass no counter between
// workgroups. Block c takes t_c = b_c - sum_{j < c-FKC} L[c, j] y_j from its
// helper unit (z rows of the block, published in block order through CF_H),
// subtracts the near terms L[c, j] y_j for j = c-FKC .. c-1 (tiles in
// registers, each wave holding a quarter of the columns for every right-hand
// side, y_j from the LDS ring of the last FKC + 1 solved blocks; the four
// waves' partial sums added in wave order) and solves the unit-lower diagonal
// block. Block 0 gathers its own b.
// The next block's tiles and diagonal block are loaded while the current one
// is solved. No wait may fall between those loads and their use: the load
// counter retires in order, so waiting on any later load would wait on them
// too. Hence the 2x2 coupling flags are loaded beside them and applied when
// the block is used, and the chain acquires the helper counter only when it
// runs out of published helpers (keeping the highest count seen), loading a
// block's t one block ahead otherwise.
// ---------------------------------------------------------------------------
constexpr int FKC = 2;
constexpr int FRING = (FKC + 1) * TB * RCH;
template <class T>
__device__ void fwd_chain(const DevSym& S, const DevFac<T>& Fc, const SolveSync& sy, int32_t nrhs,
                          T* __restrict__ y, const T* __restrict__ u, const T* __restrict__ z, int s, T* Ls,
                          T* ring, T* red, int* s_flag) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const uint8_t* two = Fc.two + eo;
  const int nblk = (w + TB - 1) / TB;
  const int nc = min(RCH, (int)nrhs);
  const int k0 = wv, k1 = wv + NWV;
  T lt[FKC][TPER];  // lt[q][m] = L[row0 + lane][(c - FKC + q) * TB + wv + NWV m]
  T dv[TPER];       // the diagonal block's strict lower part (Ld[p * TB + i] order)
  uint8_t cpl = 0;  // block c starts inside a 2x2 pivot: L[c TB][c TB - 1] is held in D
  uint8_t dz = 0;   // this thread's diagonal entry on a 2x2 pivot's coupling (zero it)
  // the one diagonal entry (i, p) of this thread with i == p + 1: p = wv + NWV qz
  const int qz = (lane - wv - 1) % NWV == 0 && lane > wv ? (lane - wv - 1) / NWV : -1;
  auto load_tiles = [&](int c) {
    const int64_t row0 = (int64_t)c * TB;
    const int nr = min(TB, w - c * TB);
#pragma unroll
    for (int q = 0; q < FKC; ++q) {
      const int j = c - FKC + q;
#pragma unroll
      for (int m = 0; m < TPER; ++m) {
        const int p = wv + NWV * m;
        lt[q][m] = (j >= 0 && lane < nr) ? P[(row0 + lane) + (int64_t)(j * TB + p) * h] : F<T>::zero();
      }
    }
    cpl = (c > 0 && lane == 0 && wv == NWV - 1) ? two[c * TB - 1] : 0;
  };
  auto load_diag = [&](int c) {
    const int64_t row0 = (int64_t)c * TB;
    const int nr = min(TB, w - c * TB);
#pragma unroll
    for (int q = 0; q < TPER; ++q) {
      const int t = tid + q * SNT, i = t % TB, p = t / TB;
      dv[q] = (p < nr && i < nr && i > p) ? P[(row0 + i) + (row0 + p) * h] : F<T>::zero();
    }
    dz = (qz >= 0 && wv + NWV * qz + 1 < nr) ? two[row0 + wv + NWV * qz] : 0;
  };
  // block 0: its rows' gather lists (static), then the updaters' chunks
  constexpr int GN = 4;
  const int nr0 = min(TB, w);
  int64_t gq = 0, gqe = 0, ix[GN];
  if (lane < nr0) {
    const int g = S.sn_first[s] + Fc.lperm[eo + lane];
    gq = S.gat_ptr[g];
    gqe = S.gat_ptr[g + 1];
#pragma unroll
    for (int m = 0; m < GN; ++m) ix[m] = gq + m < gqe ? S.gat_idx[gq + m] : 0;
  }
  load_diag(0);
  if (nblk > 1) load_tiles(1);
  if (wv == 0) {
    for (int q = S.upd_ptr[s] + lane; q < S.upd_ptr[s + 1]; q += 64) {
      const int k = S.upd_idx[q];
      (void)wait_ge(counter(sy, k, CF_U), (S.sn_nrow[k] - S.sn_ncol[k] + TB - 1) / TB, sy.ctl);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  T t0 = F<T>::zero(), t1 = F<T>::zero();
  if (lane < nr0) {
    const int64_t e = eo + lane;
    if (k0 < nc) t0 = y[e * nrhs + k0];
    if (k1 < nc) t1 = y[e * nrhs + k1];
    for (int64_t q = gq; q < gqe; q += GN) {
      T v0[GN], v1[GN];
#pragma unroll
      for (int m = 0; m < GN; ++m) {
        const bool in = q + m < gqe;
        v0[m] = (in && k0 < nc) ? u[ix[m] * nrhs + k0] : F<T>::zero();
        v1[m] = (in && k1 < nc) ? u[ix[m] * nrhs + k1] : F<T>::zero();
      }
#pragma unroll
      for (int m = 0; m < GN; ++m) ix[m] = q + GN + m < gqe ? S.gat_idx[q + GN + m] : 0;
#pragma unroll
      for (int m = 0; m < GN; ++m) {
        if (q + m < gqe) {
          t0 = t0 - v0[m];
          t1 = t1 - v1[m];
        }
      }
    }
  }
  int hseen = 0;  // helpers 1 .. hseen published and acquired
  for (int c = 0; c < nblk; ++c) {
    const int nr = min(TB, w - c * TB);
    const int64_t e = (int64_t)eo + c * TB + lane;
    const int nrn = c + 1 < nblk ? min(TB, w - (c + 1) * TB) : 0;
    // the next block's t, when its helper is already known published
    const bool tpre = c + 1 < nblk && c + 1 <= hseen;
    T tn0 = F<T>::zero(), tn1 = F<T>::zero();
    if (tpre && lane < nrn) {
      if (k0 < nc) tn0 = z[(e + TB) * nrhs + k0];
      if (k1 < nc) tn1 = z[(e + TB) * nrhs + k1];
    }
    // near terms: this wave's columns of the (up to) FKC preceding blocks
    T part[RCH];
#pragma unroll
    for (int k = 0; k < RCH; ++k) part[k] = F<T>::zero();
    if (c > 0) {
      if (cpl) lt[FKC - 1][TPER - 1] = F<T>::zero();
#pragma unroll
      for (int q = 0; q < FKC; ++q) {
        const int j = c - FKC + q;
        if (j < 0) continue;
        const T* yr = ring + (j % (FKC + 1)) * TB * RCH;
#pragma unroll
        for (int m = 0; m < TPER; ++m) {
          const T* yp = yr + (wv + NWV * m) * RCH;
#pragma unroll
          for (int k = 0; k < RCH; ++k) part[k] = part[k] + lt[q][m] * yp[k];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < RCH; ++k) red[(wv * TB + lane) * RCH + k] = part[k];
#pragma unroll
    for (int q = 0; q < TPER; ++q) Ls[tid + q * SNT] = (dz && q == qz) ? F<T>::zero() : dv[q];  // Ld[p * TB + i]
    if (c > 0 && c + 1 < nblk) load_tiles(c + 1);  // block 1's went out before the gather
    if (c + 1 < nblk) load_diag(c + 1);
    __syncthreads();
    T a0 = F<T>::zero(), a1 = F<T>::zero();
#pragma unroll
    for (int q = 0; q < NWV; ++q) {
      a0 = a0 + red[(q * TB + lane) * RCH + k0];
      a1 = a1 + red[(q * TB + lane) * RCH + k1];
    }
    T y0 = t0 - a0, y1 = t1 - a1;
#pragma unroll 8
    for (int p = 0; p + 1 < nr; ++p) {
      const T l = Ls[p * TB + lane];
      const T q0 = lane_bcast(y0, p), q1 = lane_bcast(y1, p);
      y0 = y0 - l * q0;
      y1 = y1 - l * q1;
    }
    T* yr = ring + (c % (FKC + 1)) * TB * RCH;
    const bool in = lane < nr;
    yr[lane * RCH + k0] = (in && k0 < nc) ? y0 : F<T>::zero();
    yr[lane * RCH + k1] = (in && k1 < nc) ? y1 : F<T>::zero();
    if (in) {
      if (k0 < nc) y[e * nrhs + k0] = y0;
      if (k1 < nc) y[e * nrhs + k1] = y1;
    }
    wg_publish(counter(sy, s, CF_BLK), c + 1, false);
    if (c + 1 < nblk) {
      if (!tpre) {
        // out of published helpers: wait for the next one, keep the count
        if (tid == 0) {
          *s_flag = wait_ge(counter(sy, s, CF_H), c + 1, sy.ctl);
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        hseen = *s_flag;
        if (lane < nrn) {
          if (k0 < nc) tn0 = z[(e + TB) * nrhs + k0];
          if (k1 < nc) tn1 = z[(e + TB) * nrhs + k1];
        }
      }
      t0 = tn0;
      t1 = tn1;
    }
  }
}

template <class T>
__global__ __launch_bounds__(SNT, 2) void k_fwd_dep(DevSym S, DevFac<T> Fc, const int2* __restrict__ items,
                                                 SolveSync sy, int32_t nrhs, T* __restrict__ y,
                                                 T* __restrict__ u, T* __restrict__ z, int chain) {
  // chain mode (real field): (s, -1) is the chain unit of s, (s, c) for
  // 1 <= c < nblk its block-c helper (gather and the terms of blocks
  // < c - FKC into z), (s, c >= nblk) the off-diagonal chunks as otherwise
  constexpr bool CH = std::is_same<T, double>::value;
  __shared__ T Ls[TB * (TB + 1)];  // Ls[p * (TB+1) + r] = L[row0 + r][pj + p]; then the diagonal block
  __shared__ T vs[TB * RCH];       // vs[p * RCH + k] = y[pj + p][c0 + k]
  __shared__ T ring[CH ? FRING : 1];
  __shared__ T red[CH ? NWV * TB * RCH : 1];
  __shared__ int s_t, s_ready, s_flag;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int2 it = items[take_ticket(sy.ctl, &s_t)];
  const int s = it.x, c = it.y;
  if constexpr (CH) {
    if (c < 0) {
      fwd_chain<T>(S, Fc, sy, nrhs, y, u, z, s, Ls, ring, red, &s_flag);
      return;
    }
  }
  const bool helper = CH && chain != 0;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const uint8_t* two = Fc.two + eo;
  const int nblk = (w + TB - 1) / TB;
  const bool own = c < nblk;
  const int64_t row0 = own ? (int64_t)c * TB : w + (int64_t)(c - nblk) * TB;
  const int nr = (int)min<int64_t>(TB, (own ? (int64_t)w : h) - row0);
  const int jn = own ? (helper ? max(0, c - FKC) : c) : nblk;  // own blocks feeding this unit
  const int k0 = wv, k1 = wv + NWV;
  // own block: the strict lower part of its diagonal block (stage_block_lower's
  // entries), loaded while the last product runs
  T dv[TPER];
  auto load_diag = [&]() {
#pragma unroll
    for (int q = 0; q < TPER; ++q) {
      const int t = tid + q * SNT, i = t % TB, p = t / TB;
      const bool in = p < nr && i < nr && i > p && !(i == p + 1 && two[row0 + p]);
      dv[q] = in ? P[(row0 + i) + (row0 + p) * h] : F<T>::zero();
    }
  };
  // own unit: the static part of its rows' gather (the list bounds and its
  // first GN scratch-row indices), loaded before the waits so that after
  // them only the scratch values are fetched
  constexpr int GN = 4;
  int64_t gq = 0, gqe = 0, gix[GN];
  if (own && lane < nr) {
    const int g = S.sn_first[s] + Fc.lperm[eo + row0 + lane];
    gq = S.gat_ptr[g];
    gqe = S.gat_ptr[g + 1];
#pragma unroll
    for (int m = 0; m < GN; ++m) gix[m] = gq + m < gqe ? S.gat_idx[gq + m] : 0;
  }
  // the diagonal block's loads go out before the unit's last wait: here for
  // a first block, before the last earlier block's wait otherwise (real
  // field; the complex field's registers load it after that wait)
  constexpr bool DEARLY = sizeof(T) == sizeof(double);
  bool dpre = helper;  // a helper solves no diagonal block
  if (DEARLY && own && jn == 0 && !dpre) {
    load_diag();
    dpre = true;
  }
  if (own) {
    if (wv == 0) {
      for (int q = S.upd_ptr[s] + lane; q < S.upd_ptr[s + 1]; q += 64) {
        const int k = S.upd_idx[q];
        (void)wait_ge(counter(sy, k, CF_U), (S.sn_nrow[k] - S.sn_ncol[k] + TB - 1) / TB, sy.ctl);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
  }
  T lv[TPER];
  auto load_tile = [&](int j) {
    const int pj = j * TB, nbj = min(TB, w - pj);
#pragma unroll
    for (int m = 0; m < TPER; ++m) {
      const int p = wv + NWV * m;
      lv[m] = (lane < nr && p < nbj) ? P[(row0 + lane) + (int64_t)(pj + p) * h] : F<T>::zero();
    }
    // the 2x2 coupling entry (pj + 64, pj + 63) is held in D, not in L
    if (own && j == c - 1 && lane == 0 && wv == NWV - 1 && two[pj + TB - 1]) lv[TPER - 1] = F<T>::zero();
  };
  int ready = 0;  // own blocks of s seen solved (and acquired)
  const int64_t e = eo + row0 + lane;  // own unit: this lane's row (elimination order)
  for (int c0 = 0; c0 < nrhs; c0 += RCH) {
    const int nc = min(RCH, nrhs - c0);
    T a0 = F<T>::zero(), a1 = F<T>::zero();
    // block j's y entries, fetched with its tile once the block is known solved
    T yv[TB * RCH / SNT];
    bool ypre = false;
    auto load_y = [&](int j) {
      const int pj = j * TB, nbj = min(TB, w - pj);
#pragma unroll
      for (int v = 0; v < TB * RCH / SNT; ++v) {
        const int t = tid + v * SNT, p = t / RCH, k = t % RCH;
        yv[v] = (p < nbj && k < nc) ? y[(int64_t)(eo + pj + p) * nrhs + c0 + k] : F<T>::zero();
      }
    };
    if (jn > 0) {
      load_tile(0);
      if (ready > 0) {
        load_y(0);
        ypre = true;
      }
    } else if (own && !dpre) {
      load_diag();
      dpre = true;
    }
    // own unit: b minus the descendants' products (ascending list order),
    // before waiting on the earlier blocks: GN terms per round, the first
    // round's indices prefetched, each later round's loaded with the values
    // of the round before
    T b0 = F<T>::zero(), b1 = F<T>::zero();
    if (own && lane < nr) {
      if (k0 < nc) b0 = y[e * nrhs + c0 + k0];
      if (k1 < nc) b1 = y[e * nrhs + c0 + k1];
      int64_t ix[GN];
#pragma unroll
      for (int m = 0; m < GN; ++m) ix[m] = gix[m];
      for (int64_t q = gq; q < gqe; q += GN) {
        T v0[GN], v1[GN];
#pragma unroll
        for (int m = 0; m < GN; ++m) {
          const bool in = q + m < gqe;
          v0[m] = (in && k0 < nc) ? u[ix[m] * nrhs + c0 + k0] : F<T>::zero();
          v1[m] = (in && k1 < nc) ? u[ix[m] * nrhs + c0 + k1] : F<T>::zero();
        }
#pragma unroll
        for (int m = 0; m < GN; ++m) ix[m] = q + GN + m < gqe ? S.gat_idx[q + GN + m] : 0;
#pragma unroll
        for (int m = 0; m < GN; ++m) {
          if (q + m < gqe) {
            b0 = b0 - v0[m];
            b1 = b1 - v1[m];
          }
        }
      }
    }
    for (int j = 0; j < jn; ++j) {
      const bool must = j >= ready;
      if (DEARLY && own && j + 1 == jn && !dpre) {
        load_diag();
        dpre = true;
      }
      if (must && tid == 0) {
        s_ready = wait_ge(counter(sy, s, CF_BLK), j + 1, sy.ctl);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      __syncthreads();  // also: the previous products are done with Ls / vs
      if (must) ready = s_ready;
#pragma unroll
      for (int m = 0; m < TPER; ++m) Ls[(wv + NWV * m) * (TB + 1) + lane] = lv[m];
      if (!ypre) load_y(j);
#pragma unroll
      for (int v = 0; v < TB * RCH / SNT; ++v) vs[tid + v * SNT] = yv[v];
      __syncthreads();
      ypre = false;
      if (j + 1 < jn) {
        load_tile(j + 1);
        if (j + 1 < ready) {
          load_y(j + 1);
          ypre = true;
        }
      } else if (own && !dpre) {
        load_diag();
        dpre = true;
      }
#pragma unroll 8
      for (int p = 0; p < TB; ++p) {
        const T l = Ls[p * (TB + 1) + lane];
        a0 = a0 + l * vs[p * RCH + k0];
        a1 = a1 + l * vs[p * RCH + k1];
      }
    }
    if (own && helper) {
      if (lane < nr) {
        if (k0 < nc) z[e * nrhs + c0 + k0] = b0 - a0;
        if (k1 < nc) z[e * nrhs + c0 + k1] = b1 - a1;
      }
    } else if (own) {
      T y0 = b0 - a0, y1 = b1 - a1;
      __syncthreads();  // the products are done with Ls
#pragma unroll
      for (int q = 0; q < TPER; ++q) Ls[tid + q * SNT] = dv[q];  // Ld[p * TB + i]
      dpre = false;
      __syncthreads();
#pragma unroll 8
      for (int p = 0; p + 1 < nr; ++p) {
        const T l = Ls[p * TB + lane];
        const T q0 = lane_bcast(y0, p), q1 = lane_bcast(y1, p);
        y0 = y0 - l * q0;
        y1 = y1 - l * q1;
      }
      if (lane < nr) {
        if (k0 < nc) y[e * nrhs + c0 + k0] = y0;
        if (k1 < nc) y[e * nrhs + c0 + k1] = y1;
      }
    } else if (lane < nr) {
      const int64_t ur = (S.u_off[s] + (row0 - w) + lane) * nrhs + c0;
      if (k0 < nc) u[ur + k0] = a0;
      if (k1 < nc) u[ur + k1] = a1;
    }
    __syncthreads();  // before the next right-hand-side chunk reuses Ls / vs
  }
  if (own && helper) {
    // helpers publish in block order (the chain waits for CF_H >= c)
    if (tid == 0 && c > 1) (void)wait_ge(counter(sy, s, CF_H), c - 1, sy.ctl);
    wg_publish(counter(sy, s, CF_H), c, false);
  } else if (own) {
    wg_publish(counter(sy, s, CF_BLK), c + 1, false);
  } else {
    wg_publish(counter(sy, s, CF_U), 1, true);
  }
}

template <class T>
__global__ __launch_bounds__(SNT, 2) void k_bwd_dep(DevSym S, DevFac<T> Fc, const int2* __restrict__ items,
                                                 SolveSync sy, int32_t nrhs, T* __restrict__ y,
                                                 T* __restrict__ u) {
  // own unit: Ls[i * (TB+1) + p] = L[rb + i][j0 + p], then the diagonal block,
  // and xs[i * RCH + k] = x[rb + i][c0 + k]; off unit: its pass's x rows
  constexpr int NOWN = TB * (TB + 1) + TB * RCH, NOFF = BW_RB * RCH + (BW_RB * 4 + sizeof(T) - 1) / sizeof(T);
  __shared__ T sm[NOWN > NOFF ? NOWN : NOFF];
  T* Ls = sm;
  T* xs = sm + TB * (TB + 1);
  __shared__ int s_t, s_ready;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int2 it = items[take_ticket(sy.ctl + 1, &s_t)];
  const int s = it.x, c = it.y;
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const int nblk = (w + TB - 1) / TB;
  const int ng = (w + 15) / 16, nch = bwd_row_chunks(w, h - w);
  if (c < 0) {
    // off unit: the terms of 16 own columns from one chunk of the off-
    // diagonal rows, once every ancestor holding those rows is solved
    if (wv == 0) {
      for (int q = S.anc_ptr[s] + lane; q < S.anc_ptr[s + 1]; q += 64) {
        const int a = S.anc_idx[q];
        (void)wait_ge(counter(sy, a, CB_BLK), (S.sn_ncol[a] + TB - 1) / TB, sy.ctl);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    const int g = (-1 - c) % ng, r = (-1 - c) / ng;
    const int64_t ib0 = w + (int64_t)r * BWD_CHUNK;
    bwd_update_cols(S, Fc, s, g * 16, nrhs, y, sm, reinterpret_cast<int32_t*>(sm + BW_RB * RCH), ib0,
                    nch > 1 ? min<int64_t>(h, ib0 + BWD_CHUNK) : h,
                    nch > 1 ? u + (S.u_off[s] + (int64_t)r * w) * nrhs : (T*)nullptr);
    wg_publish(counter(sy, s, CB_OFF), 1, true);
    return;
  }
  const T* P = Fc.pool + S.panel_off[s];
  const uint8_t* two = Fc.two + eo;
  const int j0 = c * TB, nbk = min(TB, w - j0);
  const int nm = nblk - 1 - c;  // later own blocks nblk-1 .. c+1
  const int k0 = wv, k1 = wv + NWV;
  // the transposed strict lower part of the diagonal block (stage_block_upper's
  // entries), loaded while the last product runs
  T dv[TPER];
  auto load_diag = [&]() {
#pragma unroll
    for (int q = 0; q < TPER; ++q) {
      const int t = tid + q * SNT, p = t % TB, i = t / TB;
      const bool in = p < nbk && i < nbk && p > i && !(p == i + 1 && two[j0 + i]);
      dv[q] = in ? P[(int64_t)(j0 + p) + (int64_t)(j0 + i) * h] : F<T>::zero();
    }
  };
  // the diagonal block's loads go out before the unit's last wait: here for
  // the last block, before the last later block's wait otherwise (real
  // field; the complex field's registers load it after that wait)
  constexpr bool DEARLY = sizeof(T) == sizeof(double);
  bool dpre = false;
  if (DEARLY && nm == 0) {
    load_diag();
    dpre = true;
  }
  if (h > w) {
    // the off units of s have applied the off-diagonal rows to every own entry
    if (tid == 0) {
      (void)wait_ge(counter(sy, s, CB_OFF), ng * nch, sy.ctl);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
  }
  T lv[TPER];
  auto load_tile = [&](int m) {
    const int j = nblk - 1 - m;
    const int64_t rb = (int64_t)j * TB;
    const int nrm = min(TB, w - j * TB);
#pragma unroll
    for (int v = 0; v < TPER; ++v) {
      const int p = wv + NWV * v;
      lv[v] = (lane < nrm && p < nbk) ? P[(rb + lane) + (int64_t)(j0 + p) * h] : F<T>::zero();
    }
    // the 2x2 coupling entry (j0 + 64, j0 + 63) is held in D, not in L
    if (j == c + 1 && lane == 0 && wv == NWV - 1 && two[j0 + TB - 1]) lv[TPER - 1] = F<T>::zero();
  };
  int ready = 0;  // own blocks of s seen solved, counted from the last (and acquired)
  for (int c0 = 0; c0 < nrhs; c0 += RCH) {
    const int nc = min(RCH, nrhs - c0);
    T a0 = F<T>::zero(), a1 = F<T>::zero();
    // block j's x entries, fetched with its tile once the block is known solved
    T xv[TB * RCH / SNT];
    bool xpre = false;
    auto load_x = [&](int m) {
      const int j = nblk - 1 - m;
      const int64_t rb = (int64_t)j * TB;
      const int nrm = min(TB, w - j * TB);
#pragma unroll
      for (int v = 0; v < TB * RCH / SNT; ++v) {
        const int t = tid + v * SNT, i = t / RCH, k = t % RCH;
        xv[v] = (i < nrm && k < nc) ? y[((int64_t)eo + rb + i) * nrhs + c0 + k] : F<T>::zero();
      }
    };
    if (nm > 0) {
      load_tile(0);
      if (ready >= 1) {
        load_x(0);
        xpre = true;
      }
    } else if (!dpre) {
      load_diag();
      dpre = true;
    }
    // this block's rows of y, final once the off units are done: minus the
    // row chunks' partial sums (ascending chunks, summed first) when they
    // were stored apart
    const int64_t e = (int64_t)eo + j0 + lane;
    T yb0 = F<T>::zero(), yb1 = F<T>::zero();
    if (lane < nbk) {
      if (k0 < nc) yb0 = y[e * nrhs + c0 + k0];
      if (k1 < nc) yb1 = y[e * nrhs + c0 + k1];
      if (nch > 1) {
        T q0 = F<T>::zero(), q1 = F<T>::zero();
        for (int r = 0; r < nch; ++r) {
          const int64_t o = (S.u_off[s] + (int64_t)r * w + j0 + lane) * nrhs + c0;
          if (k0 < nc) q0 = q0 + u[o + k0];
          if (k1 < nc) q1 = q1 + u[o + k1];
        }
        yb0 = yb0 - q0;
        yb1 = yb1 - q1;
      }
    }
    for (int m = 0; m < nm; ++m) {
      const int j = nblk - 1 - m;
      const int need = nblk - j;
      const bool must = need > ready;
      if (DEARLY && m + 1 == nm && !dpre) {
        load_diag();
        dpre = true;
      }
      if (must && tid == 0) {
        s_ready = wait_ge(counter(sy, s, CB_BLK), need, sy.ctl);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      __syncthreads();  // also: the previous products are done with Ls / xs
      if (must) ready = s_ready;
#pragma unroll
      for (int v = 0; v < TPER; ++v) Ls[lane * (TB + 1) + wv + NWV * v] = lv[v];
      if (!xpre) load_x(m);
#pragma unroll
      for (int v = 0; v < TB * RCH / SNT; ++v) xs[tid + v * SNT] = xv[v];
      __syncthreads();
      xpre = false;
      if (m + 1 < nm) {
        load_tile(m + 1);
        if (nblk - (j - 1) <= ready) {
          load_x(m + 1);
          xpre = true;
        }
      } else if (!dpre) {
        load_diag();
        dpre = true;
      }
#pragma unroll 8
      for (int i = 0; i < TB; ++i) {
        const T l = Ls[i * (TB + 1) + lane];
        a0 = a0 + l * xs[i * RCH + k0];
        a1 = a1 + l * xs[i * RCH + k1];
      }
    }
    T y0 = F<T>::zero(), y1 = F<T>::zero();
    if (lane < nbk) {
      if (k0 < nc) y0 = yb0 - a0;
      if (k1 < nc) y1 = yb1 - a1;
    }
    __syncthreads();  // the products are done with Ls
#pragma unroll
    for (int q = 0; q < TPER; ++q) {
      const int t = tid + q * SNT, p = t % TB, i = t / TB;
      Ls[p * (TB + 1) + i] = dv[q];  // Lt[p * (TB+1) + i]
    }
    dpre = false;
    __syncthreads();
#pragma unroll 8
    for (int p = nbk - 1; p > 0; --p) {
      const T l = Ls[p * (TB + 1) + lane];
      const T q0 = lane_bcast(y0, p), q1 = lane_bcast(y1, p);
      y0 = y0 - l * q0;
      y1 = y1 - l * q1;
    }
    if (lane < nbk) {
      if (k0 < nc) y[e * nrhs + c0 + k0] = y0;
      if (k1 < nc) y[e * nrhs + c0 + k1] = y1;
    }
    __syncthreads();  // before the next right-hand-side chunk reuses Ls / xs
  }
  wg_publish(counter(sy, s, CB_BLK), nblk - c, false);
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
// one 64-column block step of the wide supernodes (one launch)
template <class T>
void launch_ftri_step(const DevSym& S, const DevFac<T>& Fc, const SItem* items, const SolveStep& stp,
                      int32_t nrhs, T* y, T* z, hipStream_t st, LaunchProfiler* pf) {
  if (stp.d_end <= stp.u_beg) return;
  ProfScope ps(pf, LK_FWD_TRI, 0.0, 0.0, st);
  hipLaunchKernelGGL(k_ftri_step<T>, dim3((unsigned)(stp.d_end - stp.u_beg)), dim3(SNT), 0, st, S, Fc,
                     items + stp.u_beg, (int)(stp.u_end - stp.u_beg), nrhs, y, z);
}
template <class T>
void launch_btri_step(const DevSym& S, const DevFac<T>& Fc, const SItem* items, const SolveStep& stp,
                      int32_t nrhs, T* y, T* z, hipStream_t st, LaunchProfiler* pf) {
  if (stp.u_end <= stp.d_beg) return;
  ProfScope ps(pf, LK_BWD_TRI, 0.0, 0.0, st);
  hipLaunchKernelGGL(k_btri_step<T>, dim3((unsigned)(stp.u_end - stp.d_beg)), dim3(SNT), 0, st, S, Fc,
                     items + stp.d_beg, (int)(stp.d_end - stp.d_beg), nrhs, y, z);
}
template <class T>
void launch_zcopy(const DevSym& S, const SItem* items, int count, int32_t nrhs, const T* z, T* y, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_zcopy<T>, dim3((unsigned)count), dim3(SNT), 0, st, S, items, nrhs, z, y);
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

void launch_sync_init(int32_t* cnt, const int32_t* fu_init, int32_t ns, hipStream_t st) {
  const int64_t n = (int64_t)ns * CSTRIDE + 3;
  hipLaunchKernelGGL(k_sync_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cnt, fu_init, ns);
}
template <class T>
void launch_fwd_dep(const DevSym& S, const DevFac<T>& Fc, const int2* items, int count, const SolveSync& sy,
                    int32_t nrhs, T* y, T* u, T* z, int chain, double bytes, hipStream_t st, LaunchProfiler* pf) {
  if (count <= 0) return;
  if (chain && !std::is_same<T, double>::value) throw std::runtime_error("solve: chain units are real-field only");
  if (chain && nrhs > RCH) throw std::runtime_error("solve: chain units take at most 8 right-hand sides");
  ProfScope ps(pf, LK_FWD_TRI, 0.0, bytes, st);
  hipLaunchKernelGGL(k_fwd_dep<T>, dim3((unsigned)count), dim3(SNT), 0, st, S, Fc, items, sy, nrhs, y, u, z, chain);
}
template <class T>
void launch_bwd_dep(const DevSym& S, const DevFac<T>& Fc, const int2* items, int count, const SolveSync& sy,
                    int32_t nrhs, T* y, T* u, double bytes, hipStream_t st, LaunchProfiler* pf) {
  if (count <= 0) return;
  ProfScope ps(pf, LK_BWD_TRI, 0.0, bytes, st);
  hipLaunchKernelGGL(k_bwd_dep<T>, dim3((unsigned)count), dim3(SNT), 0, st, S, Fc, items, sy, nrhs, y, u);
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
                                    int32_t, T*, T*, hipStream_t, LaunchProfiler*);            \
  template void launch_btri_step<T>(const DevSym&, const DevFac<T>&, const SItem*, const SolveStep&, \
                                    int32_t, T*, T*, hipStream_t, LaunchProfiler*);            \
  template void launch_zcopy<T>(const DevSym&, const SItem*, int, int32_t, const T*, T*, hipStream_t); \
  template void launch_fwd_update<T>(const DevSym&, const DevFac<T>&, const int2*, int, int32_t, \
                                     const T*, T*, double, hipStream_t, LaunchProfiler*);      \
  template void launch_dsolve<T>(int32_t, int32_t, const T*, const T*, const uint8_t*, T*,      \
                                 hipStream_t, LaunchProfiler*);                                \
  template void launch_bwd_update<T>(const DevSym&, const DevFac<T>&, const int2*, int, int32_t, \
                                     T*, double, hipStream_t, LaunchProfiler*);                \
  template void launch_bwd_tri<T>(const DevSym&, const DevFac<T>&, const int32_t*, int, int32_t, \
                                  T*, double, hipStream_t, LaunchProfiler*);                   \
  template void launch_fwd_dep<T>(const DevSym&, const DevFac<T>&, const int2*, int, const SolveSync&, \
                                  int32_t, T*, T*, T*, int, double, hipStream_t, LaunchProfiler*);      \
  template void launch_bwd_dep<T>(const DevSym&, const DevFac<T>&, const int2*, int, const SolveSync&, \
                                  int32_t, T*, T*, double, hipStream_t, LaunchProfiler*);

RSL_SINST(double)
RSL_SINST(cplx)

}  // namespace rsl
