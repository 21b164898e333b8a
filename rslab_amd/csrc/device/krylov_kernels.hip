// Column-batched vector kernels for Krylov solvers that advance k right-hand
// sides in lockstep (each column keeps its own Krylov space and scalars; the
// columns share only the launches): the FGMRES of src/numeric/iterative.rs:
// 822-1021 (and its block form, :1881), COCG (:402) and COCR (:495).
//
// Vectors are row-major n x k blocks (the solve layout): entry (i, c) at
// i * k + c. Per-column scalars live in device arrays of k entries so a
// Gram-Schmidt sweep runs without host round trips. Reductions use a fixed
// grid and a fixed order of summation, so results do not depend on
// scheduling: run-to-run identical.
#include <hip/hip_runtime.h>

#include "devtypes.hpp"
#include "scalar.hpp"

namespace rsl {

constexpr int KT = 256;        // threads per block
constexpr int KRED = 256;      // partial sums per column

template <class T>
__device__ inline T kconj(T v);
template <>
__device__ inline double kconj<double>(double v) { return v; }
template <>
__device__ inline cplx kconj<cplx>(cplx v) { return {v.re, -v.im}; }

// partial[b * k + c] = sum over block b's rows of op(x[i, c]) * y[i, c], with
// op = conj for the Hermitian product (GMRES, norms) and the identity for the
// bilinear one (COCG / COCR). The grid-stride step is a multiple of k, so a
// thread always meets the same column: its column is tid % k.
template <class T, bool CONJ>
__global__ __launch_bounds__(KT) void k_dot_cols(int32_t n, int32_t k, const T* __restrict__ x,
                                                 const T* __restrict__ y, T* __restrict__ partial) {
  const int64_t tot = (int64_t)gridDim.x * KT;
  const int64_t step = (tot / k) * k;  // >= k: launch keeps KT >= k
  const int64_t g = (int64_t)blockIdx.x * KT + threadIdx.x;
  T acc = F<T>::zero();
  if (g < step) {
    const int64_t m = (int64_t)n * k;
    for (int64_t e = g; e < m; e += step) {
      const T a = CONJ ? kconj(x[e]) : x[e];
      acc = acc + a * y[e];
    }
  }
  __shared__ T sm[KT];
  sm[threadIdx.x] = acc;
  __syncthreads();
  // column c of this block: the threads t with (blockIdx.x * KT + t) % k == c,
  // summed in ascending t
  if (threadIdx.x < k) {
    const int c = threadIdx.x;
    const int t0 = (int)((c - ((int64_t)blockIdx.x * KT) % k + k) % k);
    T s = F<T>::zero();
    for (int t = t0; t < KT; t += k) s = s + sm[t];
    partial[(int64_t)blockIdx.x * k + c] = s;
  }
}

// out[c] = sum over the KRED partials of column c, in a fixed tree order
template <class T>
__global__ __launch_bounds__(KT) void k_sum_cols(int32_t k, const T* __restrict__ partial, T* __restrict__ out) {
  const int c = blockIdx.x;
  T acc = F<T>::zero();
  for (int b = threadIdx.x; b < KRED; b += KT) acc = acc + partial[(int64_t)b * k + c];
  __shared__ T sm[KT];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = KT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sm[0];
}

// y[:, c] -= a[c] * x[:, c] for active columns (a in device memory)
template <class T>
__global__ void k_axpy_cols(int32_t n, int32_t k, const T* __restrict__ a, const uint8_t* __restrict__ act,
                            const T* __restrict__ x, T* __restrict__ y) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * k) return;
  const int c = (int)(e % k);
  if (act[c]) y[e] = y[e] - a[c] * x[e];
}

// y[:, c] = s[c] * x[:, c] for active columns (s real, device memory); y may
// be x. s[c] == 0 writes zeros without reading x (clearing a block must not
// keep a NaN or Inf of memory not yet written)
template <class T>
__global__ void k_scal_cols(int32_t n, int32_t k, const double* __restrict__ s, const uint8_t* __restrict__ act,
                            const T* x, T* y) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * k) return;
  const int c = (int)(e % k);
  if (act[c]) y[e] = s[c] == 0.0 ? F<T>::zero() : F<T>::scale(x[e], s[c]);
}

// x[:, c] += sum_{q < jd[c]} Z_q[:, c] y[q, c] for active columns; the basis
// blocks Z_q are n x k, y is (max jd) x k; the terms taken in ascending q as
// the reference's x += Z y loop does, and none past the column's own
// dimension (those basis entries are stale)
template <class T>
__global__ void k_zy_cols(int32_t n, int32_t k, const int32_t* __restrict__ jd, const T* __restrict__ Z,
                          const T* __restrict__ yv, const uint8_t* __restrict__ act, T* __restrict__ x) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t m = (int64_t)n * k;
  if (e >= m) return;
  const int c = (int)(e % k);
  if (!act[c]) return;
  T acc = x[e];
  for (int q = 0; q < jd[c]; ++q) acc = acc + Z[(int64_t)q * m + e] * yv[(int64_t)q * k + c];
  x[e] = acc;
}

// x[:, c] += a[c] p[:, c] and r[:, c] -= a[c] q[:, c] for active columns
template <class T>
__global__ void k_step_cols(int32_t n, int32_t k, const T* __restrict__ a, const uint8_t* __restrict__ act,
                            const T* __restrict__ p, const T* __restrict__ q, T* __restrict__ x, T* __restrict__ r) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * k) return;
  const int c = (int)(e % k);
  if (!act[c]) return;
  x[e] = x[e] + a[c] * p[e];
  r[e] = r[e] - a[c] * q[e];
}

// p[:, c] = z[:, c] + b[c] p[:, c] for active columns
template <class T>
__global__ void k_xpby_cols(int32_t n, int32_t k, const T* __restrict__ b, const uint8_t* __restrict__ act,
                            const T* __restrict__ z, T* __restrict__ p) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * k) return;
  const int c = (int)(e % k);
  if (act[c]) p[e] = z[e] + b[c] * p[e];
}

// ------------------------------------------------------------------ launchers
namespace {
inline unsigned kblocks(int64_t m) { return (unsigned)((m + KT - 1) / KT); }
}  // namespace

int krylov_reduction_blocks() { return KRED; }

// out[c] = op(x[:, c]) . y[:, c]; partial holds KRED * k entries
template <class T>
void launch_dot_cols(int32_t n, int32_t k, bool conj, const T* x, const T* y, T* partial, T* out, hipStream_t st) {
  if (conj)
    hipLaunchKernelGGL((k_dot_cols<T, true>), dim3(KRED), dim3(KT), 0, st, n, k, x, y, partial);
  else
    hipLaunchKernelGGL((k_dot_cols<T, false>), dim3(KRED), dim3(KT), 0, st, n, k, x, y, partial);
  hipLaunchKernelGGL(k_sum_cols<T>, dim3((unsigned)k), dim3(KT), 0, st, k, partial, out);
}

template <class T>
void launch_axpy_cols(int32_t n, int32_t k, const T* a, const uint8_t* act, const T* x, T* y, hipStream_t st) {
  if ((int64_t)n * k <= 0) return;
  hipLaunchKernelGGL(k_axpy_cols<T>, dim3(kblocks((int64_t)n * k)), dim3(KT), 0, st, n, k, a, act, x, y);
}

template <class T>
void launch_scal_cols(int32_t n, int32_t k, const double* s, const uint8_t* act, const T* x, T* y, hipStream_t st) {
  if ((int64_t)n * k <= 0) return;
  hipLaunchKernelGGL(k_scal_cols<T>, dim3(kblocks((int64_t)n * k)), dim3(KT), 0, st, n, k, s, act, x, y);
}

template <class T>
void launch_zy_cols(int32_t n, int32_t k, const int32_t* jd, const T* Z, const T* yv, const uint8_t* act, T* x,
                    hipStream_t st) {
  if ((int64_t)n * k <= 0) return;
  hipLaunchKernelGGL(k_zy_cols<T>, dim3(kblocks((int64_t)n * k)), dim3(KT), 0, st, n, k, jd, Z, yv, act, x);
}

template <class T>
void launch_step_cols(int32_t n, int32_t k, const T* a, const uint8_t* act, const T* p, const T* q, T* x, T* r,
                      hipStream_t st) {
  if ((int64_t)n * k <= 0) return;
  hipLaunchKernelGGL(k_step_cols<T>, dim3(kblocks((int64_t)n * k)), dim3(KT), 0, st, n, k, a, act, p, q, x, r);
}

template <class T>
void launch_xpby_cols(int32_t n, int32_t k, const T* b, const uint8_t* act, const T* z, T* p, hipStream_t st) {
  if ((int64_t)n * k <= 0) return;
  hipLaunchKernelGGL(k_xpby_cols<T>, dim3(kblocks((int64_t)n * k)), dim3(KT), 0, st, n, k, b, act, z, p);
}

#define RSL_INST(T)                                                                                        \
  template void launch_dot_cols<T>(int32_t, int32_t, bool, const T*, const T*, T*, T*, hipStream_t);       \
  template void launch_axpy_cols<T>(int32_t, int32_t, const T*, const uint8_t*, const T*, T*, hipStream_t); \
  template void launch_scal_cols<T>(int32_t, int32_t, const double*, const uint8_t*, const T*, T*,          \
                                    hipStream_t);                                                          \
  template void launch_zy_cols<T>(int32_t, int32_t, const int32_t*, const T*, const T*, const uint8_t*, T*, \
                                  hipStream_t);                                                            \
  template void launch_step_cols<T>(int32_t, int32_t, const T*, const uint8_t*, const T*, const T*, T*, T*, \
                                    hipStream_t);                                                          \
  template void launch_xpby_cols<T>(int32_t, int32_t, const T*, const uint8_t*, const T*, T*, hipStream_t);
RSL_INST(double)
RSL_INST(cplx)
#undef RSL_INST

}  // namespace rsl
