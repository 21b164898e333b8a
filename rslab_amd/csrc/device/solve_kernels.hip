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
__device__ inline int lower_bound_i32(const int32_t* a, int n, int32_t v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Forward sweep for the supernodes of one level.
template <class T>
__global__ __launch_bounds__(SNT) void k_forward(DevSym S, DevFac<T> Fc,
                                                 const int32_t* __restrict__ list, int32_t nrhs,
                                                 T* __restrict__ y) {
  const int s = list[blockIdx.x];
  const int w = S.sn_ncol[s], first = S.sn_first[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  // (a) pull the updates of every descendant updater, ascending
  for (int u = S.upd_ptr[s]; u < S.upd_ptr[s + 1]; ++u) {
    const int k = S.upd_idx[u], nck = S.sn_ncol[k], eok = S.e_off[k];
    const int64_t hk = S.sn_nrow[k];
    const int32_t* okg = S.rs_idx + S.rs_off[k] + nck;
    const int nok = (int)(hk - nck);
    const int p0 = lower_bound_i32<T>(okg, nok, first);
    const int p1 = lower_bound_i32<T>(okg, nok, first + w);
    const T* Pk = Fc.pool + S.panel_off[k];
    const int np = p1 - p0;
    for (int64_t t = threadIdx.x; t < (int64_t)np * nrhs; t += SNT) {
      const int r = (int)(t % np), c = (int)(t / np);
      T acc = F<T>::zero();
      for (int j = 0; j < nck; ++j)
        acc = acc + Pk[(nck + p0 + r) + (int64_t)j * hk] * y[(int64_t)(eok + j) * nrhs + c];
      const int e = Fc.e_of_g[okg[p0 + r]];
      y[(int64_t)e * nrhs + c] = y[(int64_t)e * nrhs + c] - acc;
    }
    __syncthreads();
  }
  // (b) own unit-lower block in pivoted order
  for (int p = 0; p < w; ++p) {
    const int i0 = Fc.two[eo + p] ? p + 2 : p + 1;
    const int nr = w - i0;
    for (int t = threadIdx.x; t < nr * nrhs; t += SNT) {
      const int i = i0 + t % nr, c = t / nr;
      y[(int64_t)(eo + i) * nrhs + c] =
          y[(int64_t)(eo + i) * nrhs + c] - P[i + (int64_t)p * h] * y[(int64_t)(eo + p) * nrhs + c];
    }
    __syncthreads();
  }
}

template <class T>
__global__ void k_dsolve(int32_t n, int32_t nrhs, const T* __restrict__ d,
                         const T* __restrict__ ds, const uint8_t* __restrict__ two,
                         T* __restrict__ y) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * nrhs) return;
  int32_t e = (int32_t)(idx / nrhs), c = (int32_t)(idx % nrhs);
  if (two[e]) {
    T d11 = d[e], d21 = ds[e], d22 = d[e + 1];
    T detinv = F<T>::recip(d11 * d22 - d21 * d21);
    T z0 = y[(int64_t)e * nrhs + c], z1 = y[(int64_t)(e + 1) * nrhs + c];
    y[(int64_t)e * nrhs + c] = (d22 * z0 - d21 * z1) * detinv;
    y[(int64_t)(e + 1) * nrhs + c] = (d11 * z1 - d21 * z0) * detinv;
  } else if (!(e > 0 && two[e - 1])) {
    y[idx] = y[idx] * F<T>::recip(d[e]);
  }
}

// Backward sweep for the supernodes of one level (ancestors already solved).
template <class T>
__global__ __launch_bounds__(SNT) void k_backward(DevSym S, DevFac<T> Fc,
                                                  const int32_t* __restrict__ list,
                                                  int32_t nrhs, T* __restrict__ y) {
  __shared__ T red[SNT / 64][RCH];
  const int s = list[blockIdx.x];
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  const int64_t h = S.sn_nrow[s];
  const T* P = Fc.pool + S.panel_off[s];
  const int32_t* rs = S.rs_idx + S.rs_off[s];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // (a) off-diagonal rows: y_own[p] -= sum_{i>=w} L[i][p] x[rs[i]]  (wave per p)
  for (int c0 = 0; c0 < nrhs; c0 += RCH) {
    const int nc = min(RCH, nrhs - c0);
    for (int p = wv; p < w; p += SNT / 64) {
      T acc[RCH];
#pragma unroll
      for (int q = 0; q < RCH; ++q) acc[q] = F<T>::zero();
      for (int64_t i = w + lane; i < h; i += 64) {
        const T l = P[i + (int64_t)p * h];
        const int64_t e = Fc.e_of_g[rs[i]];
#pragma unroll
        for (int q = 0; q < RCH; ++q)
          if (q < nc) acc[q] = acc[q] + l * y[e * nrhs + c0 + q];
      }
#pragma unroll
      for (int q = 0; q < RCH; ++q) {
        T v = acc[q];
        for (int o = 32; o > 0; o >>= 1) {
          if constexpr (sizeof(T) == sizeof(double)) {
            v = v + __shfl_xor(v, o, 64);
          } else {
            cplx ov;
            ov.re = __shfl_xor(v.re, o, 64);
            ov.im = __shfl_xor(v.im, o, 64);
            v = v + ov;
          }
        }
        if (lane == 0 && q < nc)
          y[(int64_t)(eo + p) * nrhs + c0 + q] = y[(int64_t)(eo + p) * nrhs + c0 + q] - v;
      }
    }
  }
  __syncthreads();
  // (b) own block: x_p = y_p - sum_{i>p, not the 2x2 coupling row} L[i][p] x_i
  for (int p = w - 1; p >= 0; --p) {
    const int i0 = Fc.two[eo + p] ? p + 2 : p + 1;
    for (int c0 = 0; c0 < nrhs; c0 += RCH) {
      const int nc = min(RCH, nrhs - c0);
      T acc[RCH];
#pragma unroll
      for (int q = 0; q < RCH; ++q) acc[q] = F<T>::zero();
      for (int i = i0 + threadIdx.x; i < w; i += SNT) {
        const T l = P[i + (int64_t)p * h];
#pragma unroll
        for (int q = 0; q < RCH; ++q)
          if (q < nc) acc[q] = acc[q] + l * y[(int64_t)(eo + i) * nrhs + c0 + q];
      }
#pragma unroll
      for (int q = 0; q < RCH; ++q) {
        T v = acc[q];
        for (int o = 32; o > 0; o >>= 1) {
          if constexpr (sizeof(T) == sizeof(double)) {
            v = v + __shfl_xor(v, o, 64);
          } else {
            cplx ov;
            ov.re = __shfl_xor(v.re, o, 64);
            ov.im = __shfl_xor(v.im, o, 64);
            v = v + ov;
          }
        }
        if (lane == 0) red[wv][q] = v;
      }
      __syncthreads();
      if (threadIdx.x < nc) {
        T v = red[0][threadIdx.x];
        for (int q = 1; q < SNT / 64; ++q) v = v + red[q][threadIdx.x];
        const int64_t o = (int64_t)(eo + p) * nrhs + c0 + threadIdx.x;
        y[o] = y[o] - v;
      }
      __syncthreads();
    }
  }
}

// ---- launchers ----
template <class T>
void launch_gather_rhs(int32_t n, int32_t nrhs, const int32_t* perm_e, const double* scale,
                       const T* b, T* y, hipStream_t st) {
  int64_t tot = (int64_t)n * nrhs;
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_gather_rhs<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n,
                     nrhs, perm_e, scale, b, y);
}
template <class T>
void launch_scatter_sol(int32_t n, int32_t nrhs, const int32_t* perm_e, const double* scale,
                        const T* y, T* x, hipStream_t st) {
  int64_t tot = (int64_t)n * nrhs;
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_scatter_sol<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n,
                     nrhs, perm_e, scale, y, x);
}
template <class T>
void launch_forward(const DevSym& S, const DevFac<T>& Fc, const int32_t* list, int count,
                    int32_t nrhs, T* y, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_forward<T>, dim3(count), dim3(SNT), 0, st, S, Fc, list, nrhs, y);
}
template <class T>
void launch_dsolve(int32_t n, int32_t nrhs, const T* d, const T* ds, const uint8_t* two, T* y,
                   hipStream_t st) {
  int64_t tot = (int64_t)n * nrhs;
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_dsolve<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, nrhs,
                     d, ds, two, y);
}
template <class T>
void launch_backward(const DevSym& S, const DevFac<T>& Fc, const int32_t* list, int count,
                     int32_t nrhs, T* y, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_backward<T>, dim3(count), dim3(SNT), 0, st, S, Fc, list, nrhs, y);
}

#define RSL_SINST(T)                                                                           \
  template void launch_gather_rhs<T>(int32_t, int32_t, const int32_t*, const double*, const T*, \
                                     T*, hipStream_t);                                         \
  template void launch_scatter_sol<T>(int32_t, int32_t, const int32_t*, const double*,          \
                                      const T*, T*, hipStream_t);                              \
  template void launch_forward<T>(const DevSym&, const DevFac<T>&, const int32_t*, int, int32_t, \
                                  T*, hipStream_t);                                            \
  template void launch_dsolve<T>(int32_t, int32_t, const T*, const T*, const uint8_t*, T*,      \
                                 hipStream_t);                                                 \
  template void launch_backward<T>(const DevSym&, const DevFac<T>&, const int32_t*, int,        \
                                   int32_t, T*, hipStream_t);

RSL_SINST(double)
RSL_SINST(cplx)

}  // namespace rsl
