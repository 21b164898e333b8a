// FP64 residual and Krylov vector kernels for the refinement ladder
// (MixedLdltSolver::solve_to, src/numeric/mixed.rs:140-221, and the FGMRES of
// src/numeric/iterative.rs:822-1021).
//
// The matrix is applied as a full symmetric CSR whose values are gathered from
// the user's lower-triangle CSC values (CscMatrix::symv, src/sparse/csc.rs:294-309).
// Vectors are row-major n x k blocks (the solve layout) or single columns.
// Reductions are two-stage with a fixed grid, so results do not depend on
// scheduling: run-to-run identical.
#include <hip/hip_runtime.h>

#include "devtypes.hpp"
#include "prof.hpp"
#include "scalar.hpp"

namespace rsl {

constexpr int RT = 256;      // threads per block

// full-CSR value q := user lower value src[q]
template <class T>
__global__ void k_expand(int64_t nnz_full, const int64_t* __restrict__ src, const T* __restrict__ uv,
                         T* __restrict__ fv) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nnz_full) fv[q] = uv[src[q]];
}

// y = b - A x   (b may be null: y = A x), row-major n x k, thread per (row, column)
template <class T>
__global__ __launch_bounds__(RT) void k_residual(int32_t n, int32_t k, const int64_t* __restrict__ rp,
                                                 const int32_t* __restrict__ ci, const T* __restrict__ av,
                                                 const T* __restrict__ x, const T* __restrict__ b,
                                                 T* __restrict__ y) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * k) return;
  const int64_t i = idx / k, c = idx % k;
  T acc = F<T>::zero();
  for (int64_t q = rp[i]; q < rp[i + 1]; ++q) acc = acc + av[q] * x[(int64_t)ci[q] * k + c];
  y[idx] = b ? b[idx] - acc : acc;
}

// per-column max |v| of a row-major n x k block; out[c] holds the bit pattern
// of a non-negative double (ordered like the unsigned integer), pre-zeroed.
template <class T>
__global__ __launch_bounds__(RT) void k_colmax(int32_t n, int32_t k, const T* __restrict__ v,
                                               unsigned long long* __restrict__ out) {
  const int c = blockIdx.y;
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT)
    m = fmax(m, F<T>::mag(v[i * k + c]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  __shared__ double wm[RT / 64];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < RT / 64; ++w) m = fmax(m, wm[w]);
    atomicMax(out + c, (unsigned long long)__double_as_longlong(m));
  }
}

// x[:, c] += dx[:, c] for columns with mask[c]
template <class T>
__global__ void k_update_cols(int32_t n, int32_t k, const uint8_t* __restrict__ mask, T* __restrict__ x,
                              const T* __restrict__ dx) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * k) return;
  if (mask[idx % k]) x[idx] = x[idx] + dx[idx];
}

// dst[:, c] = src[:, c] for columns with mask[c]
template <class T>
__global__ void k_copy_cols(int32_t n, int32_t k, const uint8_t* __restrict__ mask, T* __restrict__ dst,
                            const T* __restrict__ src) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * k) return;
  if (mask[idx % k]) dst[idx] = src[idx];
}

// column cb of the row-major n x kb block B := column ca of the n x ka block A
template <class T>
__global__ void k_copy_col(int32_t n, int32_t ka, int32_t ca, const T* __restrict__ A, int32_t kb, int32_t cb,
                           T* __restrict__ B) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) B[i * kb + cb] = A[i * ka + ca];
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// ------------------------------------------------------------------ launchers
namespace {
inline unsigned blocks_for(int64_t m) { return (unsigned)((m + RT - 1) / RT); }
}  // namespace

template <class T>
void launch_expand(int64_t nnz_full, const int64_t* src, const T* uv, T* fv, hipStream_t st) {
  if (nnz_full <= 0) return;
  hipLaunchKernelGGL(k_expand<T>, dim3(blocks_for(nnz_full)), dim3(RT), 0, st, nnz_full, src, uv, fv);
}

template <class T>
void launch_residual(int32_t n, int32_t k, const int64_t* rp, const int32_t* ci, const T* av, const T* x,
                     const T* b, T* y, hipStream_t st, LaunchProfiler* pf) {
  if ((int64_t)n * k <= 0) return;
  ProfScope ps(pf, LK_RESIDUAL, 0.0, 0.0, st);
  hipLaunchKernelGGL(k_residual<T>, dim3(blocks_for((int64_t)n * k)), dim3(RT), 0, st, n, k, rp, ci, av, x,
                     b, y);
}

template <class T>
void launch_colmax(int32_t n, int32_t k, const T* v, unsigned long long* out, hipStream_t st) {
  if (n <= 0 || k <= 0) return;
  unsigned gx = std::min<unsigned>(blocks_for(n), 256u);
  hipLaunchKernelGGL(k_colmax<T>, dim3(gx, (unsigned)k), dim3(RT), 0, st, n, k, v, out);
}

template <class T>
void launch_update_cols(int32_t n, int32_t k, const uint8_t* mask, T* x, const T* dx, hipStream_t st) {
  if ((int64_t)n * k <= 0) return;
  hipLaunchKernelGGL(k_update_cols<T>, dim3(blocks_for((int64_t)n * k)), dim3(RT), 0, st, n, k, mask, x, dx);
}

template <class T>
void launch_copy_cols(int32_t n, int32_t k, const uint8_t* mask, T* dst, const T* src, hipStream_t st) {
  if ((int64_t)n * k <= 0) return;
  hipLaunchKernelGGL(k_copy_cols<T>, dim3(blocks_for((int64_t)n * k)), dim3(RT), 0, st, n, k, mask, dst, src);
}

template <class T>
void launch_copy_col(int32_t n, int32_t ka, int32_t ca, const T* A, int32_t kb, int32_t cb, T* B, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy_col<T>, dim3(blocks_for(n)), dim3(RT), 0, st, n, ka, ca, A, kb, cb, B);
}

// ---- device-resident inputs: the host computations of the same quantities,
// in an order that gives the same bits (real field; the complex field's
// modulus is the device's hypot, which may differ from the host's in the last
// bit) ----

// |a_q| of entry q: values interleaved (re, im) for the complex field
__device__ inline double entry_mag(const double* __restrict__ v, int64_t q, int cplx) {
  return cplx ? hypot(v[2 * q], v[2 * q + 1]) : fabs(v[q]);
}

// one-pass equilibration (onepass_scale, src/numeric/sparse_solver.rs:399-418):
// row maxima of |a| over the symmetric matrix; maxima are order-free, so the
// atomics give the host's value.
__global__ void k_scale_max(int64_t nnz, const int32_t* __restrict__ a_row, const int32_t* __restrict__ a_col,
                            const double* __restrict__ v, int cplx, unsigned long long* __restrict__ smax) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  const double m = entry_mag(v, q, cplx);
  if (!(m > 0.0)) return;  // zero cannot raise a maximum; NaN never compares greater
  const unsigned long long b = (unsigned long long)__double_as_longlong(m);
  const int32_t i = a_row[q], j = a_col[q];
  atomicMax(smax + i, b);
  if (i != j) atomicMax(smax + j, b);
}

// inv_sqrt_scale_guarded (src/scaling/mod.rs:85-96)
__global__ void k_scale_final(int32_t n, const unsigned long long* __restrict__ smax, double* __restrict__ scale) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = __longlong_as_double((long long)smax[i]);
  double out = 1.0;
  if (m > 0.0) {
    const double cand = 1.0 / sqrt(m);
    if (isfinite(cand) && cand > 0.0) out = cand;
  }
  scale[i] = out;
}

// max |a_ij| s_i s_j over the entries (the ForceAccept pivot floor's |A_hat|max)
__global__ void k_scaled_amax(int64_t nnz, const int32_t* __restrict__ a_row, const int32_t* __restrict__ a_col,
                              const double* __restrict__ v, int cplx, const double* __restrict__ scale,
                              unsigned long long* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  const double m = entry_mag(v, q, cplx) * scale[a_row[q]] * scale[a_col[q]];
  if (m > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// ||A||_inf (sym_norm_inf, src/numeric/mixed.rs:358-371): each row's sum taken
// in the full symmetric CSR's order, which is the host loop's order of
// accumulation for that row; the maximum over rows is order-free.
__global__ void k_row_abs_max(int32_t n, const int64_t* __restrict__ rp, const int64_t* __restrict__ src,
                              const double* __restrict__ v, int cplx, unsigned long long* __restrict__ out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int64_t q = rp[i]; q < rp[i + 1]; ++q) acc += entry_mag(v, src[q], cplx);
  if (acc > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(acc));
}

void launch_device_scale(int32_t n, int64_t nnz, const int32_t* a_row, const int32_t* a_col, const double* v,
                         int cplx, unsigned long long* smax, double* scale, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(std::max<int64_t>(n, 1))), dim3(RT), 0, st, smax, (int64_t)n);
  if (nnz > 0)
    hipLaunchKernelGGL(k_scale_max, dim3(blocks_for(nnz)), dim3(RT), 0, st, nnz, a_row, a_col, v, cplx, smax);
  if (n > 0) hipLaunchKernelGGL(k_scale_final, dim3(blocks_for(n)), dim3(RT), 0, st, n, smax, scale);
}

void launch_scaled_amax(int64_t nnz, const int32_t* a_row, const int32_t* a_col, const double* v, int cplx,
                        const double* scale, unsigned long long* out, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(RT), 0, st, out, (int64_t)1);
  if (nnz > 0)
    hipLaunchKernelGGL(k_scaled_amax, dim3(blocks_for(nnz)), dim3(RT), 0, st, nnz, a_row, a_col, v, cplx, scale,
                       out);
}

void launch_row_abs_max(int32_t n, const int64_t* rp, const int64_t* src, const double* v, int cplx,
                        unsigned long long* out, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(RT), 0, st, out, (int64_t)1);
  if (n > 0) hipLaunchKernelGGL(k_row_abs_max, dim3(blocks_for(n)), dim3(RT), 0, st, n, rp, src, v, cplx, out);
}

#define RSL_INST(T)                                                                                     \
  template void launch_expand<T>(int64_t, const int64_t*, const T*, T*, hipStream_t);                   \
  template void launch_residual<T>(int32_t, int32_t, const int64_t*, const int32_t*, const T*, const T*, \
                                   const T*, T*, hipStream_t, LaunchProfiler*);                        \
  template void launch_colmax<T>(int32_t, int32_t, const T*, unsigned long long*, hipStream_t);         \
  template void launch_update_cols<T>(int32_t, int32_t, const uint8_t*, T*, const T*, hipStream_t);     \
  template void launch_copy_cols<T>(int32_t, int32_t, const uint8_t*, T*, const T*, hipStream_t);       \
  template void launch_copy_col<T>(int32_t, int32_t, int32_t, const T*, int32_t, int32_t, T*, hipStream_t);
RSL_INST(double)
RSL_INST(cplx)
#undef RSL_INST

}  // namespace rsl
