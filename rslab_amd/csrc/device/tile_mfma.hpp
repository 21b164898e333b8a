// Matrix-core (MFMA) tile for the real-field Schur / cmod products
//   acc(i, j) = sum_c L[ar + i, cb + c] * G[br + j, c],   G = L[:, cb..) * D
// on a 64x64 output tile: 4 waves, each a 32x32 quadrant of 2x2 blocks of
// 16x16. The operands are staged through LDS one K-chunk at a time; the next
// chunk's global loads are issued before the current chunk's MFMAs.
//
// Precisions (the panel and D stay FP64; only the product is reduced):
//   TP_F64   v_mfma_f64_16x16x4f64       operands f64, accumulate f64
//   TP_F32   v_mfma_f32_16x16x4f32       operands rounded to f32, accumulate f32
//   TP_BF16  v_mfma_f32_16x16x32_bf16    operands rounded to bf16, accumulate f32
// Lane layouts (checked on gfx950): A lane l holds A[l%16][k-group l/16],
// B lane l holds B[k-group l/16][l%16]; the 16x16 result sits in lane l,
// register r at column l%16 and row (l/16) + 4r for f64, 4(l/16) + r for f32.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rsl {

enum TilePrec : int { TP_F64 = 0, TP_F32 = 1, TP_BF16 = 2 };

typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 mf_bf8 __attribute__((ext_vector_type(8)));

template <int P>
struct TileCfg;
template <>
struct TileCfg<TP_F64> {
  typedef double S;  // LDS operand type
  typedef mf_d4 V;   // accumulator vector
  static constexpr int KC = 16;
  static constexpr int LDA = 64 + 4;  // k-major: [k][row]
};
template <>
struct TileCfg<TP_F32> {
  typedef float S;
  typedef mf_f4 V;
  static constexpr int KC = 16;
  static constexpr int LDA = 64 + 4;
};
template <>
struct TileCfg<TP_BF16> {
  typedef __bf16 S;
  typedef mf_f4 V;
  static constexpr int KC = 32;
  static constexpr int LDA = 32 + 8;  // row-major: [row][k], 8 consecutive k per lane load
};

// Staging for the gathered cmod: 32 columns of K per chunk (64 for bf16).
template <int P>
struct TileCfgW;
template <>
struct TileCfgW<TP_F64> {
  typedef double S;
  static constexpr int KC = 32;
  static constexpr int LDA = 64 + 4;
  static constexpr int ELEMS = KC * LDA;
};
template <>
struct TileCfgW<TP_F32> {
  typedef float S;
  static constexpr int KC = 32;
  static constexpr int LDA = 64 + 4;
  static constexpr int ELEMS = KC * LDA;
};
template <>
struct TileCfgW<TP_BF16> {
  typedef __bf16 S;
  static constexpr int KC = 64;
  static constexpr int LDA = 64 + 8;  // row-major [row][k]
  static constexpr int ELEMS = 64 * LDA;
};

template <int P>
__device__ inline void stage_store_w(typename TileCfgW<P>::S* As, int r, int c, double v) {
  typedef typename TileCfgW<P>::S S;
  if constexpr (P == TP_BF16)
    As[r * TileCfgW<P>::LDA + c] = (S)(float)v;
  else
    As[c * TileCfgW<P>::LDA + r] = (S)v;
}

// acc += A(64 x KC) * B(64 x KC)^T from the staged chunk (all 256 threads,
// each wave its 32x32 quadrant).
template <int P, class Acc>
__device__ inline void tile_mma_w(const typename TileCfgW<P>::S* As, const typename TileCfgW<P>::S* Bs, Acc& acc) {
  typedef TileCfgW<P> C;
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  if constexpr (P == TP_BF16) {
#pragma unroll
    for (int kk = 0; kk < C::KC; kk += 32) {
      mf_bf8 a[2], b[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        a[q] = *reinterpret_cast<const mf_bf8*>(As + (wr + q * 16 + (l & 15)) * C::LDA + kk + 8 * (l >> 4));
        b[q] = *reinterpret_cast<const mf_bf8*>(Bs + (wc + q * 16 + (l & 15)) * C::LDA + kk + 8 * (l >> 4));
      }
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
          acc.v[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[bi], b[bj], acc.v[bi][bj], 0, 0, 0);
    }
  } else {
#pragma unroll 4
    for (int kk = 0; kk < C::KC; kk += 4) {
      typename C::S a[2], b[2];
      const int k = kk + (l >> 4);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        a[q] = As[k * C::LDA + wr + q * 16 + (l & 15)];
        b[q] = Bs[k * C::LDA + wc + q * 16 + (l & 15)];
      }
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          if constexpr (P == TP_F64)
            acc.v[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], b[bj], acc.v[bi][bj], 0, 0, 0);
          else
            acc.v[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[bi], b[bj], acc.v[bi][bj], 0, 0, 0);
        }
    }
  }
}

template <int P>
struct TileAcc {
  typename TileCfg<P>::V v[2][2];
};

// LDS elements for one operand of a chunk
template <int P>
constexpr int tile_stage_elems() {
  return P == TP_BF16 ? 64 * TileCfg<P>::LDA : TileCfg<P>::KC * TileCfg<P>::LDA;
}

template <int P>
__device__ inline int tacc_row(int bi, int l, int r) {
  return P == TP_F64 ? bi * 16 + (l >> 4) + 4 * r : bi * 16 + 4 * (l >> 4) + r;
}

template <int P>
__device__ inline void tile_zero(TileAcc<P>& a) {
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) a.v[bi][bj][r] = 0;
}

// G entry: row of L times the (1x1 / 2x2) D block of column c of the slice
__device__ inline double g_entry_d(const double* __restrict__ L, int64_t ld, int64_t row, int64_t col, int c,
                                   const double* __restrict__ d, const double* __restrict__ ds,
                                   const uint8_t* __restrict__ two) {
  if (two[c]) return d[c] * L[row + col * ld] + ds[c] * L[row + (col + 1) * ld];
  if (c > 0 && two[c - 1]) return ds[c - 1] * L[row + (col - 1) * ld] + d[c] * L[row + col * ld];
  return L[row + col * ld] * d[c];
}

template <int P>
__device__ inline void stage_store(typename TileCfg<P>::S* As, int r, int c, double v) {
  typedef typename TileCfg<P>::S S;
  if constexpr (P == TP_BF16)
    As[r * TileCfg<P>::LDA + c] = (S)(float)v;
  else
    As[c * TileCfg<P>::LDA + r] = (S)v;
}

// 256 threads; acc is per thread (its wave's quadrant). As/Bs hold
// tile_stage_elems<P>() elements each.
template <int P>
__device__ void ldl_tile_mfma(const double* __restrict__ L, int64_t ld, int64_t cb, int K, int64_t ar, int ma,
                              int64_t br, int mb, const double* __restrict__ d, const double* __restrict__ ds,
                              const uint8_t* __restrict__ two, TileAcc<P>& acc, typename TileCfg<P>::S* As,
                              typename TileCfg<P>::S* Bs) {
  typedef TileCfg<P> C;
  constexpr int KC = C::KC, PER = KC * 64 / 256;
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  double va[PER], vb[PER];
  auto fetch = [&](int c0) {
    const int kc = min(KC, K - c0);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + 256 * u, r = e & 63, c = e >> 6;
      va[u] = 0.0;
      vb[u] = 0.0;
      if (c < kc) {
        const int64_t col = cb + c0 + c;
        if (r < ma) va[u] = L[(ar + r) + col * ld];
        if (r < mb) vb[u] = g_entry_d(L, ld, br + r, col, c0 + c, d, ds, two);
      }
    }
  };
  fetch(0);
  for (int c0 = 0; c0 < K; c0 += KC) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + 256 * u, r = e & 63, c = e >> 6;
      stage_store<P>(As, r, c, va[u]);
      stage_store<P>(Bs, r, c, vb[u]);
    }
    __syncthreads();
    if (c0 + KC < K) fetch(c0 + KC);
    if constexpr (P == TP_BF16) {
      mf_bf8 a[2], b[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        a[q] = *reinterpret_cast<const mf_bf8*>(As + (wr + q * 16 + (l & 15)) * C::LDA + 8 * (l >> 4));
        b[q] = *reinterpret_cast<const mf_bf8*>(Bs + (wc + q * 16 + (l & 15)) * C::LDA + 8 * (l >> 4));
      }
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
          acc.v[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[bi], b[bj], acc.v[bi][bj], 0, 0, 0);
    } else {
#pragma unroll
      for (int kk = 0; kk < KC; kk += 4) {
        typename C::S a[2], b[2];
        const int k = kk + (l >> 4);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          a[q] = As[k * C::LDA + wr + q * 16 + (l & 15)];
          b[q] = Bs[k * C::LDA + wc + q * 16 + (l & 15)];
        }
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
          for (int bj = 0; bj < 2; ++bj) {
            if constexpr (P == TP_F64)
              acc.v[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], b[bj], acc.v[bi][bj], 0, 0, 0);
            else
              acc.v[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[bi], b[bj], acc.v[bi][bj], 0, 0, 0);
          }
      }
    }
    __syncthreads();
  }
}

// Visit every (i, j, value) of this thread's accumulator (tile coordinates).
template <int P, class Fn>
__device__ inline void tile_for_each(const TileAcc<P>& acc, Fn&& fn) {
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        fn(wr + tacc_row<P>(bi, l, r), wc + bj * 16 + (l & 15), (double)acc.v[bi][bj][r]);
}

// acc64 (16 doubles per thread, the accumulator's element order) += acc
template <int P>
__device__ inline void fold_into(double (&acc64)[16], const TileAcc<P>& acc) {
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[(bi * 2 + bj) * 4 + r] += (double)acc.v[bi][bj][r];
}

// tile_for_each over an FP64 copy of the accumulator
template <int P, class Fn>
__device__ inline void tile_for_each_d(const double (&acc64)[16], Fn&& fn) {
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        fn(wr + tacc_row<P>(bi, l, r), wc + bj * 16 + (l & 15), acc64[(bi * 2 + bj) * 4 + r]);
}

}  // namespace rsl
