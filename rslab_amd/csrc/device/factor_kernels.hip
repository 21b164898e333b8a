// Numeric supernodal left-looking LDL^T with Bunch-Kaufman pivoting (gfx950).
//
// Device restatement of rslab's factor_left_looking
// (src/numeric/multifrontal_ldlt.rs:3346-3576):
//   assembly    ll_factor_node          :2429-2446  -> k_assemble
//   cmod        ll_factor_node          :2452-2728  -> k_cmod (64x64 target tiles,
//               updaters applied in ascending order, as update_list)
//   cdiv        ll_cdiv_emit            :2759-3258  -> k_cdiv_small (one workgroup
//               per supernode) or k_panel_* (one launch per nb-panel, for wide ones)
//     getf2     :2846-3016  Bunch-Kaufman on the nb x nb diagonal block in LDS
//     deep rows apply_bk_panel_trailing :1347-1416
//     Schur     :3085-3167  lower-trapezoid  L21 D R^T  update
//   emit        :3196-3256  inertia, elimination maps
#include <hip/hip_runtime.h>

#include <type_traits>

#include "devtypes.hpp"
#include "prof.hpp"
#include "scalar.hpp"
#include "tile_mfma.hpp"
#include "wave.hpp"

namespace rsl {

constexpr int NT = 256;  // threads per workgroup (4 waves)
constexpr int TM = 64;   // tile rows
constexpr int TN = 64;   // tile columns
constexpr int KC = 16;   // inner chunk
constexpr int NBMAX = 128;
constexpr int GT = 512;  // threads of the wide-path getf2
constexpr int STG = 1152;  // staging elements per operand (fits every tile core's chunk)
static_assert(STG >= KC * TM && STG >= tile_stage_elems<TP_F64>() && STG * 8 >= tile_stage_elems<TP_BF16>() * 2,
              "staging buffer too small");

// ---------------------------------------------------------------------------
// Tile products acc(i, j) += sum_c  L[ar + i, cb + c] * G[br + j, c]
// with G = L[:, cb..cb+K) * D  (block diagonal 1x1 / 2x2, symmetric: no
// conjugation); g_entry forms one entry of G.
// ---------------------------------------------------------------------------
template <class T>
__device__ inline T g_entry(const T* __restrict__ L, int64_t ld, int64_t row, int64_t col,
                            int c, const T* __restrict__ d, const T* __restrict__ ds,
                            const uint8_t* __restrict__ two) {
  // c is the column relative to the D slice.
  if (two[c]) {
    T a = L[row + col * ld], b = L[row + (col + 1) * ld];
    return d[c] * a + ds[c] * b;
  }
  if (c > 0 && two[c - 1]) {
    T a = L[row + (col - 1) * ld], b = L[row + col * ld];
    return ds[c - 1] * a + d[c] * b;
  }
  return L[row + col * ld] * d[c];
}

// ---------------------------------------------------------------------------
// Complex-symmetric tile on the FP64 matrix cores: the same product as
// ldl_tile_gemm<cplx> (acc(i, j) = sum_c L[ar + i, cb + c] * G[br + j, c],
// no conjugation) as four real products per complex one,
//   re += Ar Br - Ai Bi,   im += Ar Bi + Ai Br,
// on v_mfma_f64_16x16x4f64. Each K-chunk of the operands is staged as four
// LDS planes (real and imaginary parts of A and B, k-major like the real
// FP64 tile); the next chunk's global loads are in flight during the MFMAs.
// Waves own 32x32 quadrants as in the real tile (tile_for_each's layout).
// stage holds 4 * KC * LDA doubles.
// ---------------------------------------------------------------------------
struct TileAccC {
  TileAcc<TP_F64> re, im;
};

__device__ void ldl_tile_mfma_c64(const cplx* __restrict__ L, int64_t ld, int64_t cb, int K, int64_t ar, int ma,
                                  int64_t br, int mb, const cplx* __restrict__ d, const cplx* __restrict__ ds,
                                  const uint8_t* __restrict__ two, TileAccC& acc, double* stage) {
  typedef TileCfg<TP_F64> C;
  constexpr int KCC = C::KC, LDA = C::LDA, PER = KCC * 64 / NT;
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  double* Ar = stage;
  double* Ai = Ar + KCC * LDA;
  double* Br = Ai + KCC * LDA;
  double* Bi = Br + KCC * LDA;
  cplx va[PER], vb[PER];
  auto fetch = [&](int c0) {
    const int kc = min(KCC, K - c0);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + NT * u, r = e & 63, c = e >> 6;
      va[u] = F<cplx>::zero();
      vb[u] = F<cplx>::zero();
      if (c < kc) {
        const int64_t col = cb + c0 + c;
        if (r < ma) va[u] = L[(ar + r) + col * ld];
        if (r < mb) vb[u] = g_entry(L, ld, br + r, col, c0 + c, d, ds, two);
      }
    }
  };
  fetch(0);
  for (int c0 = 0; c0 < K; c0 += KCC) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + NT * u, r = e & 63, c = e >> 6;
      Ar[c * LDA + r] = va[u].re;
      Ai[c * LDA + r] = va[u].im;
      Br[c * LDA + r] = vb[u].re;
      Bi[c * LDA + r] = vb[u].im;
    }
    __syncthreads();
    if (c0 + KCC < K) fetch(c0 + KCC);
#pragma unroll 2
    for (int kk = 0; kk < KCC; kk += 4) {
      const int k = kk + (l >> 4);
      double a_r[2], a_i[2], a_n[2], b_r[2], b_i[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        a_r[q] = Ar[k * LDA + wr + q * 16 + (l & 15)];
        a_i[q] = Ai[k * LDA + wr + q * 16 + (l & 15)];
        a_n[q] = -a_i[q];
        b_r[q] = Br[k * LDA + wc + q * 16 + (l & 15)];
        b_i[q] = Bi[k * LDA + wc + q * 16 + (l & 15)];
      }
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          acc.re.v[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r[bi], b_r[bj], acc.re.v[bi][bj], 0, 0, 0);
          acc.re.v[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_n[bi], b_i[bj], acc.re.v[bi][bj], 0, 0, 0);
          acc.im.v[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r[bi], b_i[bj], acc.im.v[bi][bj], 0, 0, 0);
          acc.im.v[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i[bi], b_r[bj], acc.im.v[bi][bj], 0, 0, 0);
        }
    }
    __syncthreads();
  }
}

// Visit every (i, j, value) of this thread's complex accumulator.
template <class Fn>
__device__ inline void tile_for_each_c64(const TileAccC& acc, Fn&& fn) {
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        fn(wr + tacc_row<TP_F64>(bi, l, r), wc + bj * 16 + (l & 15),
           cplx{acc.re.v[bi][bj][r], acc.im.v[bi][bj][r]});
}

// ---------------------------------------------------------------------------
// Assembly: zero-filled pool, then A's equilibrated entries scattered to their
// panel slots (each lower entry owns exactly one slot).
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_assemble(DevSym S, DevFac<T> Fc, const T* __restrict__ vals,
                           const double* __restrict__ scale) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S.nnz) return;
  double sc = scale[S.a_row[k]] * scale[S.a_col[k]];
  Fc.pool[S.amap[k]] = F<T>::scale(vals[k], sc);
}

// Elimination position of every panel row (the solve's backward gather).
template <class T>
__global__ void k_rows_e(DevSym S, DevFac<T> Fc, int64_t nrows) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nrows) Fc.rs_e[q] = Fc.e_of_g[S.rs_idx[q]];
}

// Factor status at the start of a numeric factor: inertia and perturbation
// counters cleared, no failure recorded (a kernel, not a memset, so the
// launch sequence can be captured and replayed as a graph).
__global__ void k_factor_init(unsigned long long* __restrict__ counters, int32_t* __restrict__ status) {
  const int t = threadIdx.x;
  if (t < 4) counters[t] = 0ull;
  if (t == 4) counters[4] = ~0ull;
  if (t == 5) status[0] = 0;
  if (t == 6) status[1] = INT32_MAX;
}

void launch_factor_init(unsigned long long* counters, int32_t* status, hipStream_t st) {
  hipLaunchKernelGGL(k_factor_init, dim3(1), dim3(64), 0, st, counters, status);
}

template <class T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}

// ---------------------------------------------------------------------------
// cmod: each workgroup owns one 64x64 tile of a target panel (distinct tiles,
// so no two workgroups write the same entry) and subtracts every relevant
// updater's  L_k[a] D_k L_k[b]^T  block in ascending updater order.
// ---------------------------------------------------------------------------
template <class T, int PR>
__global__ __launch_bounds__(NT) void k_cmod(DevSym S, DevFac<T> Fc,
                                             const CItem* __restrict__ items,
                                             const CTask* __restrict__ tasks) {
  __shared__ T acc_s[TM * TN];
  __shared__ __align__(16) T stage[2 * STG];
  __shared__ int32_t trow[TM], rpos[TM], cpos[TN];
  const int tid = threadIdx.x;
  const CItem it = items[blockIdx.x];
  const int s = it.s;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s], first = S.sn_first[s];
  T* P = Fc.pool + S.panel_off[s];
  const int32_t* rs = S.rs_idx + S.rs_off[s];
  const int mr = min<int64_t>(TM, h - it.r0), mc = min(TN, w - it.c0);
  {
    // the target tile, 16 independent loads per thread before the LDS stores
    T v[TM * TN / NT];
#pragma unroll
    for (int u = 0; u < TM * TN / NT; ++u) {
      const int e = tid + u * NT, i = e % TM, j = e / TM;
      v[u] = (i < mr && j < mc) ? P[(it.r0 + i) + (int64_t)(it.c0 + j) * h] : F<T>::zero();
    }
#pragma unroll
    for (int u = 0; u < TM * TN / NT; ++u) acc_s[tid + u * NT] = v[u];
  }
  for (int i = tid; i < mr; i += NT) trow[i] = rs[it.r0 + i];
  __syncthreads();
  for (int t = it.t_beg; t < it.t_end; ++t) {
    const CTask tk = tasks[t];
    const int k = tk.k, nck = S.sn_ncol[k], eo = S.e_off[k];
    const int64_t hk = S.sn_nrow[k];
    const T* Pk = Fc.pool + S.panel_off[k];
    const int32_t* okg = S.rs_idx + S.rs_off[k] + nck;
    const int ma = tk.a1 - tk.a0, mb = tk.b1 - tk.b0;
    for (int i = tid; i < ma; i += NT) {
      int g = okg[tk.a0 + i], lo = 0, hi = mr - 1;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (trow[mid] < g) lo = mid + 1;
        else hi = mid;
      }
      rpos[i] = lo;
    }
    for (int j = tid; j < mb; j += NT) cpos[j] = okg[tk.b0 + j] - first - it.c0;
    if constexpr (std::is_same<T, double>::value) {
      typedef typename TileCfg<PR>::S Sx;
      TileAcc<PR> acc;
      tile_zero(acc);
      ldl_tile_mfma<PR>(Pk, hk, 0, nck, nck + tk.a0, ma, nck + tk.b0, mb, Fc.d + eo, Fc.ds + eo, Fc.two + eo, acc,
                       reinterpret_cast<Sx*>(stage), reinterpret_cast<Sx*>(stage + STG));
      tile_for_each<PR>(acc, [&](int i, int j, double v) {
        if (i < ma && j < mb && tk.a0 + i >= tk.b0 + j) {
          const int e = rpos[i] + cpos[j] * TM;
          acc_s[e] = acc_s[e] - v;
        }
      });
    } else {
      // complex symmetric: four real FP64 matrix-core products per complex one
      TileAccC acc;
      tile_zero(acc.re);
      tile_zero(acc.im);
      ldl_tile_mfma_c64(Pk, hk, 0, nck, nck + tk.a0, ma, nck + tk.b0, mb, Fc.d + eo, Fc.ds + eo, Fc.two + eo, acc,
                        reinterpret_cast<double*>(stage));
      tile_for_each_c64(acc, [&](int i, int j, cplx v) {
        if (i < ma && j < mb && tk.a0 + i >= tk.b0 + j) {
          const int e = rpos[i] + cpos[j] * TM;
          acc_s[e] = acc_s[e] - v;
        }
      });
    }
    __syncthreads();
  }
  for (int e = tid; e < TM * TN; e += NT) {
    int i = e % TM, j = e / TM;
    if (i < mr && j < mc) P[(it.r0 + i) + (int64_t)(it.c0 + j) * h] = acc_s[e];
  }
}

// Real-field cmod on matrix cores. Every updater k of a 64x64 target tile
// contributes L_k[rows] D_k L_k[cols]^T; its rows and columns are gathered
// straight into tile coordinates (the plan gives each task the tile rows /
// columns it hits as bit masks, so a tile row finds its source row by a
// population count), and all updaters' products accumulate in the matrix
// cores' registers over one long K loop, 32 columns per chunk, the next
// chunk (possibly the next updater's first) loading into registers while
// the current one multiplies. The tile is then reduced once, on the entries
// of the panel's lower triangle, the only ones the reference updates.
template <int PR>
__global__ __launch_bounds__(NT) void k_cmod_mfma(DevSym S, DevFac<double> Fc, const CItem* __restrict__ items,
                                                  const CTask* __restrict__ tasks) {
  typedef TileCfgW<PR> C;
  typedef typename C::S Sx;
  constexpr int KCW = C::KC, PER = KCW * 64 / NT;
  __shared__ __align__(16) Sx As[C::ELEMS];
  __shared__ __align__(16) Sx Bs[C::ELEMS];
  const int tid = threadIdx.x;
  const CItem it = items[blockIdx.x];
  double va[PER], vb[PER];
  // chunk [cc, cc + KCW) of task tt gathered into tile rows / columns
  auto fetch = [&](int tt, int cc) {
    const CTask tk = tasks[tt];
    const int k = tk.k, nck = S.sn_ncol[k], eo = S.e_off[k];
    const int64_t hk = S.sn_nrow[k];
    const double* Pk = Fc.pool + S.panel_off[k];
    const int kc = min(KCW, nck - cc);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + NT * u, r = e & 63, c = e >> 6;
      const uint64_t below = (1ull << r) - 1ull;
      va[u] = 0.0;
      vb[u] = 0.0;
      if (c < kc) {
        const int64_t col = cc + c;
        if ((tk.rmask >> r) & 1ull) va[u] = Pk[(nck + tk.a0 + __popcll(tk.rmask & below)) + col * hk];
        if ((tk.cmask >> r) & 1ull)
          vb[u] = g_entry_d(Pk, hk, nck + tk.b0 + __popcll(tk.cmask & below), col, cc + c, Fc.d + eo, Fc.ds + eo,
                            Fc.two + eo);
      }
    }
  };
  const int t_beg = it.t_beg, t_end = it.t_end;
  int t = t_beg, c0 = 0;
  int nck_t = S.sn_ncol[tasks[t].k];
  fetch(t, 0);
  TileAcc<PR> acc;
  tile_zero(acc);
  double acc64[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc64[q] = 0.0;
  for (;;) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + NT * u, r = e & 63, c = e >> 6;
      stage_store_w<PR>(As, r, c, va[u]);
      stage_store_w<PR>(Bs, r, c, vb[u]);
    }
    __syncthreads();
    int tn = t, cn = c0 + KCW, nck_n = nck_t;
    if (cn >= nck_t) {
      tn = t + 1;
      cn = 0;
      if (tn < t_end) nck_n = S.sn_ncol[tasks[tn].k];
    }
    if (tn < t_end) fetch(tn, cn);
    tile_mma_w<PR>(As, Bs, acc);
    if constexpr (PR != TP_F64) {
      // reduced-precision products: each updater's sum is carried into FP64
      if (tn != t) {
        fold_into(acc64, acc);
        tile_zero(acc);
      }
    }
    __syncthreads();
    if (tn >= t_end) break;
    t = tn;
    c0 = cn;
    nck_t = nck_n;
  }
  const int s = it.s;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s], gcol0 = S.sn_first[s] + it.c0;
  double* P = Fc.pool + S.panel_off[s];
  const int32_t* rs = S.rs_idx + S.rs_off[s] + it.r0;
  const int mr = min<int64_t>(TM, h - it.r0), mc = min(TN, w - it.c0);
  // the targets read as one batch, then updated (a store followed by a
  // load of another column would wait for it: the columns' distance is a
  // runtime stride)
  double old[16];
  int n = 0;
  auto get = [&](int i, int j, double) {
    old[n++] = (i < mr && j < mc && rs[i] >= gcol0 + j) ? P[(it.r0 + i) + (int64_t)(it.c0 + j) * h] : 0.0;
  };
  auto put = [&](int i, int j, double v) {
    if (i < mr && j < mc && rs[i] >= gcol0 + j) P[(it.r0 + i) + (int64_t)(it.c0 + j) * h] = old[n] - v;
    ++n;
  };
  if constexpr (PR != TP_F64) {
    tile_for_each_d<PR>(acc64, get);
    n = 0;
    tile_for_each_d<PR>(acc64, put);
  } else {
    tile_for_each<PR>(acc, get);
    n = 0;
    tile_for_each<PR>(acc, put);
  }
}

// Two-phase real-field cmod. Phase one computes every (updater, target)
// product block densely on matrix cores,
//   C = L_k[p0.., :] D_k L_k[p0..p0+N, :]^T,
// one 64x64 tile per workgroup with contiguous operand rows (no gathers, no
// masks), and stores it in the level's product buffer: FP64 for the FP64
// path, FP32 for the reduced-precision ones (their products accumulate in
// FP32 per updater, as in the gathered kernel). Phase two gives each target
// tile one workgroup that sums its updaters' entries in FP64, in ascending
// updater order, and subtracts the sum from the panel's lower triangle:
// every thread owns fixed tile entries, so no two threads touch one entry and
// the result does not depend on scheduling.
//
// Workgroups are dispatched round-robin over the 8 XCDs; the product tiles
// are renumbered so that each XCD runs one contiguous run of them, keeping a
// pair's operand columns within one L2.
__device__ inline int64_t xcd_contiguous(int64_t b, int64_t n) {
  const int64_t x = b & 7, q = n >> 3, r = n & 7;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + (b >> 3);
}

template <int PR>
__global__ __launch_bounds__(NT) void k_cmod_prod(DevSym S, DevFac<double> Fc, const PTile* __restrict__ pt,
                                                  int64_t npt, void* __restrict__ cbuf) {
  typedef TileCfgW<PR> C;
  typedef typename C::S Sx;
  constexpr int KCW = C::KC, PER = KCW * 64 / NT;
  __shared__ __align__(16) Sx As[C::ELEMS];
  __shared__ __align__(16) Sx Bs[C::ELEMS];
  const PTile p = pt[xcd_contiguous(blockIdx.x, npt)];
  const int k = p.k, nck = S.sn_ncol[k], eo = S.e_off[k];
  const int64_t hk = S.sn_nrow[k];
  const double* Pk = Fc.pool + S.panel_off[k];
  const double *d = Fc.d + eo, *ds = Fc.ds + eo;
  const uint8_t* two = Fc.two + eo;
  const int ma = min(TM, p.m - p.mt * TM), mb = min(TN, p.n - p.nt * TN);
  const int64_t ar = nck + p.p0 + p.mt * TM, br = nck + p.p0 + p.nt * TN;
  const int tid = threadIdx.x;
  double va[PER], vb[PER];
  // chunk [c0, c0 + KCW): operand rows straight from the updater's panel
  auto fetch = [&](int c0) {
    const int kc = min(KCW, nck - c0);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + NT * u, r = e & 63, c = e >> 6;
      va[u] = 0.0;
      vb[u] = 0.0;
      if (c < kc) {
        const int64_t col = c0 + c;
        if (r < ma) va[u] = Pk[(ar + r) + col * hk];
        if (r < mb) vb[u] = g_entry_d(Pk, hk, br + r, col, c0 + c, d, ds, two);
      }
    }
  };
  TileAcc<PR> acc;
  tile_zero(acc);
  fetch(0);
  for (int c0 = 0; c0 < nck; c0 += KCW) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + NT * u, r = e & 63, c = e >> 6;
      stage_store_w<PR>(As, r, c, va[u]);
      stage_store_w<PR>(Bs, r, c, vb[u]);
    }
    __syncthreads();
    if (c0 + KCW < nck) fetch(c0 + KCW);
    tile_mma_w<PR>(As, Bs, acc);
    __syncthreads();
  }
  const int64_t ldc = p.n, r0 = (int64_t)p.mt * TM, c0 = (int64_t)p.nt * TN;
  if constexpr (PR == TP_F64) {
    double* Cp = static_cast<double*>(cbuf) + p.c_off;
    tile_for_each<PR>(acc, [&](int i, int j, double v) {
      if (i < ma && j < mb) Cp[(r0 + i) * ldc + c0 + j] = v;
    });
  } else {
    float* Cp = static_cast<float*>(cbuf) + p.c_off;
    tile_for_each<PR>(acc, [&](int i, int j, double v) {
      if (i < ma && j < mb) Cp[(r0 + i) * ldc + c0 + j] = (float)v;
    });
  }
}

// G = L21 D for every supernode of a finished level (real field), kept so the
// cmod products read their second operand with plain loads: block (y, x) of
// the grid covers 16 columns of the y-th supernode of the list.
__global__ __launch_bounds__(NT) void k_make_g(DevSym S, DevFac<double> Fc, const int32_t* __restrict__ list,
                                               double* __restrict__ gpool) {
  const int s = list[blockIdx.y];
  const int w = S.sn_ncol[s], eo = S.e_off[s], c0 = blockIdx.x * 16;
  const int64_t h = S.sn_nrow[s], m = h - w;
  if (c0 >= w || m <= 0) return;
  const double* P = Fc.pool + S.panel_off[s];
  double* G = gpool + S.g_off[s];
  const int nc = min(16, w - c0);
  // D of columns c0-1 .. c0+16 (a 2x2 pair may straddle the chunk's edges)
  __shared__ double dd[18], dsd[18];
  __shared__ int tw[18];
  if (threadIdx.x < 18) {
    const int c = c0 - 1 + (int)threadIdx.x;
    const bool in = c >= 0 && c < w;
    dd[threadIdx.x] = in ? Fc.d[eo + c] : 0.0;
    dsd[threadIdx.x] = in ? Fc.ds[eo + c] : 0.0;
    tw[threadIdx.x] = in ? Fc.two[eo + c] : 0;
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < m; i += NT) {
    double v[18];
#pragma unroll
    for (int t = 0; t < 18; ++t) {
      const int c = c0 - 1 + t;
      v[t] = (c >= 0 && c < w) ? P[(w + i) + (int64_t)c * h] : 0.0;
    }
    // the expressions of g_entry_d
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      if (q < nc) {
        const int t = q + 1;
        double g;
        if (tw[t]) g = dd[t] * v[t] + dsd[t] * v[t + 1];
        else if (c0 + q > 0 && tw[t - 1]) g = dsd[t - 1] * v[t - 1] + dd[t] * v[t];
        else g = v[t] * dd[t];
        G[i + (int64_t)(c0 + q) * m] = g;
      }
    }
  }
}

// FP64 product tiles straight from HBM/L2 into the matrix cores' operand
// registers (no LDS staging, no barriers): each wave owns a 32x32
// quadrant, and for every 4-column step lane l loads A[l%16 (+16)][k + l/16]
// and G[...] in the MFMA operand layout. The next group of GK steps is loaded
// while the current one multiplies. Same products, same order as k_cmod_prod.
template <int PR>
__global__ __launch_bounds__(NT) void k_cmod_prod_reg(DevSym S, DevFac<double> Fc, const double* __restrict__ gpool,
                                                      const PTile* __restrict__ pt, int64_t npt,
                                                      void* __restrict__ cbuf) {
  static_assert(PR == TP_F64, "reduced-precision products use the LDS-staged kernel");
  typedef typename TileCfg<PR>::S Sx;
  constexpr int GK = 2;  // 4-column steps per register group
  const PTile p = pt[xcd_contiguous(blockIdx.x, npt)];
  const int k = p.k, nck = S.sn_ncol[k];
  const int64_t hk = S.sn_nrow[k], ldg = hk - nck;
  const int ma = min(TM, p.m - p.mt * TM), mb = min(TN, p.n - p.nt * TN);
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32, kq = l >> 4;
  const int ra = wr + (l & 15), rb = wc + (l & 15);
  const double* pa = Fc.pool + S.panel_off[k] + nck + p.p0 + p.mt * TM + ra;  // A(r, c) = pa[c * hk]
  const double* pg = gpool + S.g_off[k] + p.p0 + p.nt * TN + rb;              // G(r, c) = pg[c * ldg]
  const bool a0 = ra < ma, a1 = ra + 16 < ma, b0 = rb < mb, b1 = rb + 16 < mb;
  TileAcc<PR> acc;
  tile_zero(acc);
  Sx A[2][GK][2], B[2][GK][2];
  auto load = [&](int buf, int c0) {
#pragma unroll
    for (int st = 0; st < GK; ++st) {
      const int c = c0 + 4 * st + kq;
      const bool ok = c < nck;
      const int64_t oa = (int64_t)c * hk, og = (int64_t)c * ldg;
      A[buf][st][0] = (Sx)((ok && a0) ? pa[oa] : 0.0);
      A[buf][st][1] = (Sx)((ok && a1) ? pa[oa + 16] : 0.0);
      B[buf][st][0] = (Sx)((ok && b0) ? pg[og] : 0.0);
      B[buf][st][1] = (Sx)((ok && b1) ? pg[og + 16] : 0.0);
    }
  };
  auto mma = [&](int buf) {
#pragma unroll
    for (int st = 0; st < GK; ++st)
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          acc.v[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[buf][st][bi], B[buf][st][bj], acc.v[bi][bj], 0, 0, 0);
        }
  };
  constexpr int CG = 4 * GK;  // columns per group
  load(0, 0);
  for (int c0 = 0; c0 < nck; c0 += 2 * CG) {
    if (c0 + CG < nck) load(1, c0 + CG);
    mma(0);
    if (c0 + CG >= nck) break;
    if (c0 + 2 * CG < nck) load(0, c0 + 2 * CG);
    mma(1);
  }
  const int64_t ldc = p.n, r0 = (int64_t)p.mt * TM, cc0 = (int64_t)p.nt * TN;
  double* Cp = static_cast<double*>(cbuf) + p.c_off;
  tile_for_each<PR>(acc, [&](int i, int j, double v) {
    if (i < ma && j < mb) Cp[(r0 + i) * ldc + cc0 + j] = v;
  });
}

template <class CT>
__global__ __launch_bounds__(NT) void k_cmod_sum(DevSym S, DevFac<double> Fc, const CItem* __restrict__ items,
                                                 const CTask* __restrict__ tasks, const CT* __restrict__ cbuf) {
  __shared__ double tile[TN * (TM + 1)];
  const int tid = threadIdx.x, j = tid & 63, i0 = tid >> 6;
  const CItem it = items[blockIdx.x];
  const uint64_t jbelow = (1ull << j) - 1ull;
  // this thread's entries: column j, rows i0 + 4u
  double acc[TM / 4];
#pragma unroll
  for (int u = 0; u < TM / 4; ++u) acc[u] = 0.0;
  for (int t = it.t_beg; t < it.t_end; ++t) {
    const CTask tk = tasks[t];
    if (!((tk.cmask >> j) & 1ull)) continue;
    const CT* cp = cbuf + tk.cbase + tk.b0 + __popcll(tk.cmask & jbelow);
    CT v[TM / 4];
#pragma unroll
    for (int u = 0; u < TM / 4; ++u) {
      const int i = i0 + 4 * u;
      v[u] = 0;
      if ((tk.rmask >> i) & 1ull) v[u] = cp[(int64_t)(tk.a0 + __popcll(tk.rmask & ((1ull << i) - 1ull))) * tk.ldc];
    }
#pragma unroll
    for (int u = 0; u < TM / 4; ++u) acc[u] += (double)v[u];
  }
#pragma unroll
  for (int u = 0; u < TM / 4; ++u) tile[j * (TM + 1) + i0 + 4 * u] = acc[u];
  __syncthreads();
  const int s = it.s;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s], gcol0 = S.sn_first[s] + it.c0;
  double* P = Fc.pool + S.panel_off[s];
  const int32_t* rs = S.rs_idx + S.rs_off[s] + it.r0;
  const int mr = min<int64_t>(TM, h - it.r0), mc = min(TN, w - it.c0);
  // the targets read as one batch, then updated
  constexpr int PER = TM * TN / NT;
  double old[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int e = tid + u * NT, i = e & 63, jj = e >> 6;
    old[u] = (i < mr && jj < mc && rs[i] >= gcol0 + jj) ? P[(it.r0 + i) + (int64_t)(it.c0 + jj) * h] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int e = tid + u * NT, i = e & 63, jj = e >> 6;
    if (i < mr && jj < mc && rs[i] >= gcol0 + jj)
      P[(it.r0 + i) + (int64_t)(it.c0 + jj) * h] = old[u] - tile[jj * (TM + 1) + i];
  }
}

// ---------------------------------------------------------------------------
// cdiv building blocks
// ---------------------------------------------------------------------------
template <class T>
struct SnCtx {
  int s;
  T* P;
  int64_t h;
  int w, first, eo, nb;
  T *d, *ds;
  uint8_t* two;
  int32_t* lperm;
  int32_t* deep;
};

template <class T>
__device__ inline SnCtx<T> make_ctx(const DevSym& S, const DevFac<T>& Fc, int s, int nb) {
  SnCtx<T> c;
  c.s = s;
  c.P = Fc.pool + S.panel_off[s];
  c.h = S.sn_nrow[s];
  c.w = S.sn_ncol[s];
  c.first = S.sn_first[s];
  c.eo = S.e_off[s];
  c.nb = nb;
  c.d = Fc.d + c.eo;
  c.ds = Fc.ds + c.eo;
  c.two = Fc.two + c.eo;
  c.lperm = Fc.lperm + c.eo;
  c.deep = Fc.deep + (int64_t)s * 2 * NBMAX;
  return c;
}

__device__ inline int pk_off(int i, int j, int ld) {
  // column-major packed lower triangle of an ld x ld block
  return j * ld - (j * (j - 1)) / 2 + (i - j);
}

template <class T>
__device__ inline void wave_argmax(double& m, int& idx) {
  for (int o = 32; o > 0; o >>= 1) {
    double om = __shfl_xor(m, o, 64);
    int oi = __shfl_xor(idx, o, 64);
    if (om > m || (om == m && oi < idx)) {
      m = om;
      idx = oi;
    }
  }
}

__device__ inline double wave_max(double m) {
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  return m;
}

template <class T>
__device__ inline T perturb_pivot(T dv, double floor) {
  double mag = F<T>::mag(dv);
  if (mag >= floor) return dv;
  if (mag == 0.0) return F<T>::from_real(floor);
  return dv * F<T>::from_real(floor / mag);
}

// Initialise the per-supernode D / pivot bookkeeping.
template <int NTG, class T>
__device__ void cdiv_init(const SnCtx<T>& c) {
  for (int p = threadIdx.x; p < c.w; p += NTG) {
    c.lperm[p] = p;
    c.two[p] = 0;
    c.ds[p] = F<T>::zero();
    c.d[p] = F<T>::zero();
  }
}

// Column of packed offset e in a pw x pw column-major packed lower triangle:
// the largest j with S(j) = j*pw - j(j-1)/2 <= e, from the quadratic's root
// in float and one integer correction each way (pw <= 128: all exact in float
// up to the square root's last bit).
__device__ inline int packed_col(int e, int pw) {
  const float b = (float)(2 * pw + 1);
  int j = (int)((b - sqrtf(b * b - 8.0f * (float)e)) * 0.5f);
  if (j * pw - (j * (j - 1)) / 2 > e) --j;
  if ((j + 1) * pw - ((j + 1) * j) / 2 <= e) ++j;
  return j;
}

// Column k0 of the trailing block is the next step's pivot column: its new
// entries fall to threads 1..pw-k0-1 in the first pass (thread t holds row
// k0 + t), which leave its squared-magnitude maximum and first row reaching it
// per wave (waves 0 and 1: pw <= 128) in cm_s / ci_s, the values the next
// step's column scan would find.
template <class T>
__device__ inline void next_column_candidates(int k0, int pw, double m, double* cm_s, int* ci_s) {
  const int tid = threadIdx.x, wv = tid >> 6;
  if (wv < 2) {
    const int i0 = k0 + 64 * wv;
    int idx = (m > 0.0) ? tid + k0 : i0 - 1;
    double mm = m > 0.0 ? m : 0.0;
    wave_argmax_dpp(mm, idx, i0);
    if ((tid & 63) == 0) {
      cm_s[wv] = mm;
      ci_s[wv] = idx;
    }
  }
  (void)pw;
}

// Trailing update of one Bunch-Kaufman step on the packed block, element-
// parallel: the columns k0..pw-1 form one contiguous run of the packed
// triangle, swept in storage order (each thread every NTG-th entry, B of them
// in flight), so every thread gets the same share and LDS accesses are
// consecutive. Entry (i, j):
//   1x1:  x -= w1[i] * l1[j]                      (skipped when l1[j] == 0)
//   2x2:  x  = x - w1[i] * l1[j] - w2[i] * l2[j]
template <int NTG, bool TWO, class T>
__device__ inline void packed_trailing_update(T* Ab, int pw, int k0, const T* w1s, const T* w2s, const T* l1s,
                                              const T* l2s, double* cm_s, int* ci_s) {
  constexpr int B = 4;
  const int base = pk_off(k0, k0, pw), tot = pw * (pw + 1) / 2;
  double cand = 0.0;
  for (int e0 = base + (int)threadIdx.x; e0 < tot; e0 += NTG * B) {
    int ii[B], jj[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int e = e0 + NTG * u;
      ii[u] = -1;
      jj[u] = 0;
      if (e < tot) {
        const int j = packed_col(e, pw);
        jj[u] = j;
        ii[u] = j + (e - (j * pw - (j * (j - 1)) / 2));
      }
    }
    T x[B], a[B], m[B], b[B], m2[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      if (ii[u] >= 0) {
        x[u] = Ab[e0 + NTG * u];
        a[u] = w1s[ii[u]];
        m[u] = l1s[jj[u]];
        if constexpr (TWO) {
          b[u] = w2s[ii[u]];
          m2[u] = l2s[jj[u]];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      if (ii[u] >= 0) {
        T v = x[u];
        if constexpr (TWO) {
          v = x[u] - a[u] * m[u] - b[u] * m2[u];
          Ab[e0 + NTG * u] = v;
        } else if (!F<T>::is_zero(m[u])) {
          v = x[u] - a[u] * m[u];
          Ab[e0 + NTG * u] = v;
        }
        if (u == 0 && jj[0] == k0 && ii[0] > k0) cand = F<T>::magsq(v);
      }
    }
  }
  next_column_candidates<T>(k0, pw, cand, cm_s, ci_s);
}

// The 1x1 step's update reading column k in place: entry (i, j) of the
// trailing block takes w = A(i, k) and the multiplier A(j, k) * dinv (the
// same product the staged form stores), skipped when that multiplier is zero.
template <int NTG, class T>
__device__ inline void packed_trailing_update_1x1(T* Ab, int pw, int k, T dinv, double* cm_s, int* ci_s) {
  constexpr int B = 4;
  const int k0 = k + 1, base = pk_off(k0, k0, pw), tot = pw * (pw + 1) / 2;
  const T* colk = Ab + pk_off(k, k, pw) - k;  // colk[i] = A(i, k)
  double cand = 0.0;  // |new A(k0 + tid, k0)|^2 for this thread's candidate row
  for (int e0 = base + (int)threadIdx.x; e0 < tot; e0 += NTG * B) {
    int ii[B], jj[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int e = e0 + NTG * u;
      ii[u] = -1;
      jj[u] = 0;
      if (e < tot) {
        const int j = packed_col(e, pw);
        jj[u] = j;
        ii[u] = j + (e - (j * pw - (j * (j - 1)) / 2));
      }
    }
    T x[B], a[B], m[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      if (ii[u] >= 0) {
        x[u] = Ab[e0 + NTG * u];
        a[u] = colk[ii[u]];
        m[u] = colk[jj[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      if (ii[u] >= 0) {
        const T mu = m[u] * dinv;
        T v = x[u];
        if (!F<T>::is_zero(mu)) {
          v = x[u] - a[u] * mu;
          Ab[e0 + NTG * u] = v;
        }
        if (u == 0 && jj[0] == k0 && ii[0] > k0) cand = F<T>::magsq(v);
      }
    }
  }
  next_column_candidates<T>(k0, pw, cand, cm_s, ci_s);
}

// Static ownership of the packed block for the 1x1 trailing update: the
// strict lower part is cut into row strips (row i, columns 4 jb .. 4 jb + 3,
// i >= 4 jb), ordered by jb then i and dealt to the threads cyclically, so the
// strips still live at step k (row and some column past k) stay spread over
// all threads. A thread computes its strips' packed offsets once per panel
// and per step needs no index arithmetic beyond a few adds. Threads below T0
// own no strips.
template <int NTG, int T0>
struct PanelStrips {
  // at most NBMAX^2 / 8 + NBMAX / 2 strips (pw = NBMAX)
  static constexpr int MAXS = (NBMAX * NBMAX / 8 + NBMAX / 2 + (NTG - T0) - 1) / (NTG - T0);
  int i[MAXS], j0[MAXS], o[MAXS];  // row, first column, packed offset of (i, j0); i = -1: none
  __device__ void init(int pw) {
    // strips of column block jb: rows 4 jb .. pw - 1; S(jb) strips before it
    const int ns = (pw + 3) / 4 * pw - 2 * ((pw + 3) / 4) * ((pw + 3) / 4 - 1);
#pragma unroll
    for (int q = 0; q < MAXS; ++q) {
      const int sidx = ((int)threadIdx.x - T0) + q * (NTG - T0);
      i[q] = -1;
      j0[q] = 0;
      o[q] = 0;
      if ((int)threadIdx.x >= T0 && sidx < ns) {
        // largest jb with S(jb) = jb pw - 2 jb (jb - 1) <= sidx
        const float b = (float)(pw + 2);
        int jb = (int)((b - sqrtf(b * b - 8.0f * (float)sidx)) * 0.25f);
        if (jb * pw - 2 * jb * (jb - 1) > sidx) --jb;
        if ((jb + 1) * pw - 2 * (jb + 1) * jb <= sidx) ++jb;
        const int row = 4 * jb + (sidx - (jb * pw - 2 * jb * (jb - 1)));
        i[q] = row;
        j0[q] = 4 * jb;
        o[q] = pk_off(row, 4 * jb, pw);
      }
    }
  }
};

// The 1x1 step's update (packed_trailing_update_1x1's arithmetic: entry (i,
// j) -= A(i, k) * (A(j, k) * dinv), skipped for a zero multiplier) over the
// thread's strips, columns jmin.. (k + 1, or k + 2 when column k + 1 is
// updated separately).
template <int NTG, int T0, class T>
__device__ inline void strips_update_1x1(T* Ab, int pw, int k, int jmin, T dinv, const PanelStrips<NTG, T0>& st) {
  const T* colk = Ab + pk_off(k, k, pw) - k;  // colk[i] = A(i, k)
#pragma unroll
  for (int q = 0; q < PanelStrips<NTG, T0>::MAXS; ++q) {
    const int i = st.i[q], j0 = st.j0[q];
    if (i >= jmin && j0 + 3 >= jmin) {
      const T a = colk[i];
      int o = st.o[q];
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const int j = j0 + jj;
        if (j >= jmin && j <= i) {
          const T mu = colk[j] * dinv;
          if (!F<T>::is_zero(mu)) Ab[o] = Ab[o] - a * mu;
        }
        o += pw - j - 1;
      }
    }
  }
}

// Wave 0's share of a pipelined 1x1 step (FP64): column k0 = k + 1 brought up
// to date with the same arithmetic (every lane forms the diagonal entry, lane
// 0 stores it; lane l takes rows k0 + 1 + l + 64 q), then the next step's
// pivot search on it — the column scan, its argmax, sqrt and 1 / A(k0, k0) —
// published in dec (absakk, colmax, akk, 1/akk) and deci (imax) for every wave
// to read after the step's closing barrier.
__device__ inline void pivot_column_step(double* Ab, int pw, int k, double dinv, double* dec, int* deci) {
  const int lane = threadIdx.x & 63, k0 = k + 1;
  const double* colk = Ab + pk_off(k, k, pw) - k;
  const double mu = colk[k0] * dinv;
  const bool upd = !F<double>::is_zero(mu);
  const int od = pk_off(k0, k0, pw);
  double akk = Ab[od];
  if (upd) akk = akk - colk[k0] * mu;
  double cm = 0.0;
  int imax = k0;
  for (int i = k0 + 1 + lane; i < pw; i += 64) {
    const int o = od + (i - k0);
    double v = Ab[o];
    if (upd) {
      v = v - colk[i] * mu;
      Ab[o] = v;
    }
    const double m = F<double>::magsq(v);
    if (m > cm) {
      cm = m;
      imax = i;
    }
  }
  wave_argmax_dpp(cm, imax, k0 + 1);
  if (lane == 0) {
    if (upd) Ab[od] = akk;
    dec[0] = F<double>::mag(akk);
    dec[1] = sqrt(cm);
    dec[2] = akk;
    dec[3] = F<double>::recip(akk);
    deci[0] = imax;
  }
}

// Bunch-Kaufman on the diagonal block [kb,ke) held in LDS (packed lower).
// Candidates, interchanges and updates bounded to the panel rows, as
// ll_cdiv_emit's getf2 (multifrontal_ldlt.rs:2846-3016). Returns 1 when the
// pivot policy is Fail and a zero / singular pivot is met (uniform).
template <int NTG, class T>
__device__ int panel_getf2(const SnCtx<T>& c, int kb, int ke, T* Ab, T* l1s, T* l2s, T* w1s, T* w2s,
                           int32_t* deep_s, int32_t* two_s, int32_t* lperm_s, PivotPolicy pol, int* npert) {
  const int tid = threadIdx.x, lane = tid & 63;
  const int pw = ke - kb, ld = pw;
  const double alpha = (1.0 + sqrt(17.0)) / 8.0;
  const double GROWTH_EPS = 1e-14;
#define A_(i, j) Ab[pk_off((i), (j), ld)]
  for (int q = tid; q < pw; q += NTG) {
    deep_s[q] = -1;
    two_s[q] = 0;
    lperm_s[q] = c.lperm[kb + q];
  }
  __shared__ double cand_m[2];
  __shared__ int cand_i[2];
  // Pipelined 1x1 steps (FP64, 512 threads): wave 0 updates the next pivot
  // column and runs the next search while waves 1.. update the strips; the
  // search result is double-buffered by step parity.
  constexpr bool PIPE = NTG >= 512 && std::is_same<T, double>::value;
  constexpr int T0 = PIPE ? 64 : 0;
  __shared__ double dec[2][4];
  __shared__ int deci[2];
  PanelStrips<NTG, T0> strips;
  strips.init(pw);
  __syncthreads();
  int k = 0, par = 0;
  bool have_cand = false;  // the previous step's update left this column's max
  bool have_dec = false;   // the previous step published this step's search
  while (k < pw) {
    // ---- pivot search (every wave computes the same answer) ----
    const bool pre = PIPE && have_dec;
    T akk = F<T>::zero(), dinv_pre = F<T>::zero();
    double absakk, colmax;
    int imax = k;
    if (pre) {
      if constexpr (PIPE) {
        absakk = dec[par][0];
        colmax = dec[par][1];
        akk = dec[par][2];
        dinv_pre = dec[par][3];
        imax = deci[par];
      }
    } else {
    absakk = F<T>::mag(A_(k, k));
    double cm = 0.0;
    if (have_cand) {
      // wave 0 holds the rows below wave 1's: it wins ties
      const double c0 = cand_m[0], c1 = cand_m[1];
      if (c0 >= c1 && c0 > 0.0) {
        cm = c0;
        imax = cand_i[0];
      } else if (c1 > 0.0) {
        cm = c1;
        imax = cand_i[1];
      }
    } else {
      for (int i = k + 1 + lane; i < pw; i += 64) {
        double m = F<T>::magsq(A_(i, k));
        if (m > cm) {
          cm = m;
          imax = i;
        }
      }
      wave_argmax_dpp(cm, imax, k + 1);  // pw <= 128: at most two candidate rows per lane
    }
    colmax = sqrt(cm);
    }
    // decision below reads row imax; every wave must finish reading before
    // any wave starts the interchange / update of this step (a pipelined
    // plain 1x1 step reads nothing the update writes)
    int kstep, kp;
    bool plain = false;
    if (fmax(absakk, colmax) == 0.0) {
      if (!pol.have_floor) return 1 + k;  // nonzero: the failing column, plus one
      kstep = 1;
      kp = k;
      plain = true;
    } else if (absakk >= alpha * colmax) {
      kstep = 1;
      kp = k;
      plain = true;
    } else {
      double rm = 0.0;
      for (int j = k + lane; j < imax; j += 64) rm = fmax(rm, F<T>::magsq(A_(imax, j)));
      for (int i = imax + 1 + lane; i < pw; i += 64) rm = fmax(rm, F<T>::magsq(A_(i, imax)));
      rm = wave_max_dpp(rm);
      const double rowmax = sqrt(rm);
      if (absakk >= alpha * colmax * (colmax / rowmax)) {
        kstep = 1;
        kp = k;
      } else if (F<T>::mag(A_(imax, imax)) >= alpha * rowmax) {
        kstep = 1;
        kp = imax;
      } else {
        kstep = 2;
        kp = imax;
      }
    }
    if (!(pre && plain)) __syncthreads();
    // ---- symmetric interchange of p < q inside the block ----
    const int p = (kstep == 1) ? k : k + 1;
    if (kp != p) {
      const int q = kp;
      for (int t = tid; t < pw; t += NTG) {
        int i = q + 1 + t;
        if (i < pw) {
          T x = A_(i, p);
          A_(i, p) = A_(i, q);
          A_(i, q) = x;
        }
        i = p + 1 + t;
        if (i < q) {
          T x = A_(i, p);
          A_(i, p) = A_(q, i);
          A_(q, i) = x;
        }
        if (t < p) {
          T x = A_(p, t);
          A_(p, t) = A_(q, t);
          A_(q, t) = x;
        }
        if (t == 0) {
          T x = A_(p, p);
          A_(p, p) = A_(q, q);
          A_(q, q) = x;
          int32_t y = lperm_s[p];
          lperm_s[p] = lperm_s[q];
          lperm_s[q] = y;
          deep_s[k] = q;
        }
      }
      __syncthreads();
    }
    if (kstep == 1) {
      // a pipelined step without interchange takes A(k, k) as published
      const bool use_pre = pre && kp == p;
      T dk = use_pre ? akk : A_(k, k);
      bool fresh = !use_pre;
      if (pol.have_floor) {
        if (F<T>::mag(dk) < pol.floor) {
          dk = perturb_pivot(dk, pol.floor);
          fresh = true;
          if (tid == 0) {
            A_(k, k) = dk;
            (*npert)++;
          }
        }
      } else if (F<T>::is_zero(dk)) {
        return 1;
      }
      if (tid == 0) c.d[kb + k] = dk;
      const T dinv = fresh ? F<T>::recip(dk) : dinv_pre;
      // in-block rank-1 update (rows of this panel) straight from column k,
      // which is replaced by its multipliers once every thread is past the
      // update (the next step reads column k again only after its first
      // barrier, and scaling a column commutes with its interchanges)
      if constexpr (PIPE) {
        if (tid < 64) {
          if (k + 1 < pw) pivot_column_step(Ab, pw, k, dinv, dec[par ^ 1], &deci[par ^ 1]);
        } else {
          strips_update_1x1<NTG, T0>(Ab, pw, k, k + 2, dinv, strips);
        }
        have_dec = k + 1 < pw;
      } else {
        strips_update_1x1<NTG, T0>(Ab, pw, k, k + 1, dinv, strips);
      }
      have_cand = false;
      // column k keeps its unscaled entries until the block is written back
      // (panel_store multiplies them by w1s[k] = 1/d_k then; the later
      // interchanges only move entries within finished columns)
      if (tid == 0) w1s[k] = dinv;
      __syncthreads();
      k += 1;
      par ^= 1;
    } else {
      T d11 = A_(k, k), d21 = A_(k + 1, k), d22 = A_(k + 1, k + 1);
      T det = d11 * d22 - d21 * d21;
      double scale = fmax(fmax(F<T>::mag(d11), F<T>::mag(d22)), F<T>::mag(d21));
      double growth_floor = GROWTH_EPS * scale * scale;
      if (pol.have_floor) {
        double fl = fmax(pol.floor * pol.floor, growth_floor);
        if (F<T>::mag(det) < fl) {
          double lift = fmax(pol.floor, scale * sqrt(GROWTH_EPS));
          d11 = d11 + F<T>::from_real(lift);
          d22 = d22 + F<T>::from_real(lift);
          det = d11 * d22 - d21 * d21;
          if (F<T>::mag(det) < fl) det = det + F<T>::from_real(fl);
          if (tid == 0) (*npert)++;
        }
      } else if (F<T>::mag(det) <= growth_floor) {
        return 1;
      }
      const T detinv = F<T>::recip(det);
      if (tid == 0) {
        c.d[kb + k] = d11;
        c.ds[kb + k] = d21;
        c.d[kb + k + 1] = d22;
        c.two[kb + k] = 1;
        two_s[k] = 1;
      }
      for (int i = k + 2 + tid; i < pw; i += NTG) {
        T wik = A_(i, k), wik1 = A_(i, k + 1);
        T a = (d22 * wik - d21 * wik1) * detinv;
        T b = (d11 * wik1 - d21 * wik) * detinv;
        l1s[i] = a;
        l2s[i] = b;
        w1s[i] = wik;
        w2s[i] = wik1;
      }
      __syncthreads();
      packed_trailing_update<NTG, true>(Ab, pw, k + 2, w1s, w2s, l1s, l2s, cand_m, cand_i);
      have_cand = true;
      have_dec = false;
      for (int i = k + 2 + tid; i < pw; i += NTG) {
        A_(i, k) = l1s[i];
        A_(i, k + 1) = l2s[i];
      }
      __syncthreads();
      k += 2;
    }
  }
  __syncthreads();  // the last step's multipliers are in place
#undef A_
  return 0;
}

// Load the lower part of the diagonal block into packed LDS.
template <int NTG, class T>
__device__ void panel_load(const SnCtx<T>& c, int kb, int ke, T* Ab) {
  const int pw = ke - kb, n = pw * pw;
  // batches of 8 independent loads per thread before their LDS stores
  for (int e0 = threadIdx.x; e0 < n; e0 += 8 * NTG) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * NTG;
      const int i = e % pw, j = e / pw;
      v[u] = (e < n && i >= j) ? c.P[(kb + i) + (int64_t)(kb + j) * c.h] : F<T>::zero();
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * NTG;
      const int i = e % pw, j = e / pw;
      if (e < n && i >= j) Ab[pk_off(i, j, pw)] = v[u];
    }
  }
  __syncthreads();
}

// Write the factored block back, apply its row interchanges to the already
// factored columns [0,kb) of the panel rows (staged 64 columns at a time in
// the block's LDS, which is free once written back), and publish the
// interchange partners and the composite permutation (position q holds the
// row / column that sat at position cperm[q] before this panel's steps).
template <int NTG, bool SWAP_COLS, class T>
__device__ void panel_store(const SnCtx<T>& c, int kb, int ke, T* Ab, const int32_t* deep_s,
                            const int32_t* two_s, const int32_t* lperm_s, int32_t* cperm_s, int* any_s,
                            const T* colscale) {
  const int pw = ke - kb;
  // 1x1 columns below the diagonal hold A(i, k); their multipliers are
  // A(i, k) * (1/d_k) (colscale[k]); 2x2 columns hold their multipliers
#pragma unroll 4
  for (int e = threadIdx.x; e < pw * pw; e += NTG) {
    const int i = e % pw, j = e / pw;
    if (i >= j) {
      T v = Ab[pk_off(i, j, pw)];
      if (i > j && !two_s[j] && !(j > 0 && two_s[j - 1])) v = v * colscale[j];
      c.P[(kb + i) + (int64_t)(kb + j) * c.h] = v;
    }
  }
  for (int q = threadIdx.x; q < pw; q += NTG) c.lperm[kb + q] = lperm_s[q];
  if (threadIdx.x == 0) {
    for (int q = 0; q < pw; ++q) cperm_s[q] = q;
    int any = 0;
    for (int k = 0; k < pw;) {
      const int kp = deep_s[k];
      if (kp >= 0) {
        const int src = two_s[k] ? k + 1 : k;
        const int t = cperm_s[src];
        cperm_s[src] = cperm_s[kp];
        cperm_s[kp] = t;
        any = 1;
      }
      k += two_s[k] ? 2 : 1;
    }
    *any_s = any;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < pw; q += NTG) {
    c.deep[q] = deep_s[q];
    c.deep[NBMAX + q] = cperm_s[q];
  }
  if (SWAP_COLS && *any_s && kb > 0) {
    // CB columns at a time: every thread gathers its (up to 8) permuted
    // sources into registers, the workgroup synchronizes, then they are stored
    constexpr int PER = 8;
    constexpr int CB = PER * NTG / NBMAX;
    for (int j0 = 0; j0 < kb; j0 += CB) {
      const int nc = min(CB, kb - j0), n = pw * nc;
      T v[PER];
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int e = threadIdx.x + u * NTG;
        const int i = e % pw, j = e / pw;
        v[u] = e < n ? c.P[(kb + cperm_s[i]) + (int64_t)(j0 + j) * c.h] : F<T>::zero();
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int e = threadIdx.x + u * NTG;
        const int i = e % pw, j = e / pw;
        if (e < n) c.P[(kb + i) + (int64_t)(j0 + j) * c.h] = v[u];
      }
      __syncthreads();
    }
  }
  __syncthreads();
}

// Deferred Schur update of one 64x64 tile of P[ke.., ke..w).
template <int PR, class T>
__device__ void panel_schur_tile(const SnCtx<T>& c, int kb, int ke, int rr0, int cc0, T* stage) {
  const int64_t mt = c.h - ke;
  const int cw = c.w - ke, pw = ke - kb;
  const int ma = (int)min<int64_t>(TM, mt - rr0), mb = min(TN, cw - cc0);
  if constexpr (std::is_same<T, double>::value) {
    typedef typename TileCfg<PR>::S Sx;
    TileAcc<PR> acc;
    tile_zero(acc);
    ldl_tile_mfma<PR>(c.P, c.h, kb, pw, ke + rr0, ma, ke + cc0, mb, c.d + kb, c.ds + kb, c.two + kb, acc,
                     reinterpret_cast<Sx*>(stage), reinterpret_cast<Sx*>(stage + STG));
    // the targets read as one batch, then updated
    double old[16];
    int n = 0;
    tile_for_each<PR>(acc, [&](int i, int j, double) {
      old[n++] = (i < ma && j < mb && rr0 + i >= cc0 + j) ? c.P[(ke + rr0 + i) + (int64_t)(ke + cc0 + j) * c.h] : 0.0;
    });
    n = 0;
    tile_for_each<PR>(acc, [&](int i, int j, double v) {
      if (i < ma && j < mb && rr0 + i >= cc0 + j) c.P[(ke + rr0 + i) + (int64_t)(ke + cc0 + j) * c.h] = old[n] - v;
      ++n;
    });
  } else {
    // complex symmetric: four real FP64 matrix-core products per complex one
    TileAccC acc;
    tile_zero(acc.re);
    tile_zero(acc.im);
    ldl_tile_mfma_c64(c.P, c.h, kb, pw, ke + rr0, ma, ke + cc0, mb, c.d + kb, c.ds + kb, c.two + kb, acc,
                      reinterpret_cast<double*>(stage));
    tile_for_each_c64(acc, [&](int i, int j, cplx v) {
      if (i < ma && j < mb && rr0 + i >= cc0 + j) {
        T* dst = c.P + (ke + rr0 + i) + (int64_t)(ke + cc0 + j) * c.h;
        *dst = *dst - v;
      }
    });
  }
}

// Out-of-line copy for the one-workgroup supernode kernel, so the tile core's
// registers are not kept live across its getf2 and replay phases.
template <int PR, class T>
__device__ __attribute__((noinline)) void panel_schur_tile_call(const SnCtx<T>& c, int kb, int ke, int rr0, int cc0,
                                                                T* stage) {
  panel_schur_tile<PR>(c, kb, ke, rr0, cc0, stage);
}

// Emit maps and inertia for supernode s (multifrontal_ldlt.rs:3196-3256).
template <class T>
__device__ void cdiv_emit(const DevSym& S, const DevFac<T>& Fc, const SnCtx<T>& c,
                          int* cnt /* LDS [4] */, int npert) {
  if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
  __syncthreads();
  for (int p = threadIdx.x; p < c.w; p += NT) {
    const int lp = c.lperm[p];
    if (lp < 0 || lp >= c.w) {  // a pivot map outside the supernode: report, never write through it
      atomicMax(&Fc.status[0], 2);
      continue;
    }
    int g = c.first + lp;
    Fc.e_of_g[g] = c.eo + p;
    Fc.perm_e[c.eo + p] = S.sym_perm[g];
    const bool second = p > 0 && c.two[p - 1];
    if (second) continue;
    if (c.two[p]) {
      T d1 = c.d[p], d2 = c.d[p + 1], e = c.ds[p];
      double det_r = F<T>::real(d1 * d2 - e * e);
      double tr_r = F<T>::real(d1 + d2);
      if (det_r < 0.0) {
        atomicAdd(&cnt[0], 1);
        atomicAdd(&cnt[1], 1);
      } else if (det_r > 0.0) {
        atomicAdd(&cnt[tr_r >= 0.0 ? 0 : 1], 2);
      } else {
        atomicAdd(&cnt[2], 1);
        atomicAdd(&cnt[tr_r >= 0.0 ? 0 : 1], 1);
      }
    } else {
      double r = F<T>::real(c.d[p]);
      atomicAdd(&cnt[r > 0.0 ? 0 : (r < 0.0 ? 1 : 2)], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) Fc.sn_inertia[3 * (int64_t)c.s + threadIdx.x] = cnt[threadIdx.x];
  if (threadIdx.x == 0) {
    atomicAdd(&Fc.counters[0], (unsigned long long)cnt[0]);
    atomicAdd(&Fc.counters[1], (unsigned long long)cnt[1]);
    atomicAdd(&Fc.counters[2], (unsigned long long)cnt[2]);
    if (npert) atomicAdd(&Fc.counters[3], (unsigned long long)npert);
  }
}

// A zero pivot under the Fail policy at column col of supernode s: the
// first failure in (supernode, column) order is kept in counters[4].
template <class T>
__device__ void flag_failure(const DevFac<T>& Fc, int s, int col) {
  if (threadIdx.x == 0) {
    atomicMax(&Fc.status[0], 1);
    atomicMin(&Fc.status[1], s);
    atomicMin(&Fc.counters[4], ((unsigned long long)(unsigned)s << 32) | (unsigned)col);
  }
}

// The packed diagonal block is live only during getf2, the GEMM staging only
// during the Schur update, so they share storage (fits 160 KB for c64, nb 128).
// Deep-row replay (apply_bk_panel_trailing, multifrontal_ldlt.rs:1347-1416) of
// up to RR rows [row0, row0 + nr) below a factored panel: the rows are staged
// in LDS with the panel's composite column permutation applied up front, then
// the panel's steps are replayed in order with the final multipliers L11 (a
// content's multiplier moves with it under the later symmetric interchanges,
// so this is the same sequence of operations), all threads sharing each
// step's update, and finally scaled by the 1x1 / 2x2 blocks of D.
// X holds RR * NBMAX elements, Lc (replay_chunk + 1) * NBMAX; cperm / two_s
// are the panel's composite permutation and 2x2 flags (pw entries).
constexpr int RR = 64;  // deep rows per pass
// L11 columns staged per chunk (LDS: (RR + RCW + 1) * NBMAX elements)
template <class T>
constexpr int replay_chunk() {
  return sizeof(T) == 16 ? 8 : 32;
}

template <class T>
__device__ void replay_rows_block(const SnCtx<T>& c, int kb, int ke, int64_t row0, int nr, T* X, T* Lc,
                                  const int32_t* cperm, const int32_t* two_s) {
  constexpr int RCW = replay_chunk<T>();
  constexpr int SB = 8;  // loads in flight per thread while staging
  const int pw = ke - kb, n = pw * RR;
  const int tid = threadIdx.x, r = tid & (RR - 1), q = tid / RR;
  for (int e0 = tid; e0 < n; e0 += NT * SB) {
    T v[SB];
#pragma unroll
    for (int u = 0; u < SB; ++u) {
      const int e = e0 + u * NT, j = e / RR, rr = e % RR;
      v[u] = (e < n && rr < nr) ? c.P[row0 + rr + (int64_t)(kb + cperm[j]) * c.h] : F<T>::zero();
    }
#pragma unroll
    for (int u = 0; u < SB; ++u)
      if (e0 + u * NT < n) X[e0 + u * NT] = v[u];
  }
  const T* L = c.P + kb + (int64_t)kb * c.h;  // L11(j, k) = L[j + k * h]
  int k = 0;
  for (int k0 = 0; k0 < pw; k0 += RCW) {
    // stage L11 columns [k0, k0 + RCW) (one more when a 2x2 block straddles the chunk end)
    const int kc = min(RCW + 1, pw - k0), nl = kc * NBMAX;
    __syncthreads();
    for (int e0 = tid; e0 < nl; e0 += NT * SB) {
      T v[SB];
#pragma unroll
      for (int u = 0; u < SB; ++u) {
        const int e = e0 + u * NT, j = e % NBMAX, kk = e / NBMAX;
        v[u] = (e < nl && j < pw) ? L[j + (int64_t)(k0 + kk) * c.h] : F<T>::zero();
      }
#pragma unroll
      for (int u = 0; u < SB; ++u)
        if (e0 + u * NT < nl) Lc[e0 + u * NT] = v[u];
    }
    __syncthreads();
    const int kend = min(k0 + RCW, pw);
    while (k < kend) {
      const T* m1 = Lc + (k - k0) * NBMAX;
      constexpr int G = NT / RR;  // column stride of a thread
      if (two_s[k]) {
        const T* m2 = m1 + NBMAX;
        const T wk = X[k * RR + r], wk1 = X[(k + 1) * RR + r];
        int j = k + 2 + q;
        for (; j + 7 * G < pw; j += 8 * G) {
          T a1[8], a2[8], x[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            a1[u] = m1[j + u * G];
            a2[u] = m2[j + u * G];
            x[u] = X[(j + u * G) * RR + r];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) X[(j + u * G) * RR + r] = x[u] - wk * a1[u] - wk1 * a2[u];
        }
        for (; j < pw; j += G) X[j * RR + r] = X[j * RR + r] - wk * m1[j] - wk1 * m2[j];
        k += 2;
      } else {
        const T wk = X[k * RR + r];
        int j = k + 1 + q;
        for (; j + 7 * G < pw; j += 8 * G) {
          T a1[8], x[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            a1[u] = m1[j + u * G];
            x[u] = X[(j + u * G) * RR + r];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u)
            X[(j + u * G) * RR + r] = F<T>::is_zero(a1[u]) ? x[u] : x[u] - wk * a1[u];
        }
        for (; j < pw; j += G) {
          const T m = m1[j];
          if (!F<T>::is_zero(m)) X[j * RR + r] = X[j * RR + r] - wk * m;
        }
        k += 1;
      }
      __syncthreads();
    }
  }
  for (int kk = q; kk < pw; kk += NT / RR) {
    if (kk > 0 && two_s[kk - 1]) continue;  // second column of a 2x2: done with the first
    if (two_s[kk]) {
      const T d11 = c.d[kb + kk], d21 = c.ds[kb + kk], d22 = c.d[kb + kk + 1];
      const T detinv = F<T>::recip(d11 * d22 - d21 * d21);
      const T wk = X[kk * RR + r], wk1 = X[(kk + 1) * RR + r];
      X[kk * RR + r] = (d22 * wk - d21 * wk1) * detinv;
      X[(kk + 1) * RR + r] = (d11 * wk1 - d21 * wk) * detinv;
    } else {
      X[kk * RR + r] = X[kk * RR + r] * F<T>::recip(c.d[kb + kk]);
    }
  }
  __syncthreads();
  for (int e = tid; e < n; e += NT) {
    const int j = e / RR, rr = e % RR;
    if (rr < nr) c.P[row0 + rr + (int64_t)(kb + j) * c.h] = X[e];
  }
  __syncthreads();
}

// Real field: one thread per deep row, its panel columns held in registers,
// RH at a time, so the replay needs neither per-step barriers nor LDS traffic
// for the rows. Pass p takes columns [RH p, RH p + RH): L11's column panel for
// those columns is staged in LDS as the unit-lower factor of the replay (2x2
// coupling entries zero: those sit in D); step k subtracts x_k L11(j, k) from
// every later column j in ascending k, the reference's order (a 2x2 step's
// two subtractions included): inside the pass in registers, and for the
// columns of later passes as running values kept in a scratch of
// pw x NT values (column-major by row, so the threads' accesses coalesce).
// RH = 32 leaves the registers the compiler needs to keep the multiplier
// loads in flight.
constexpr int RH = 32;

// Lp[k * LPW + jj] = L11(RH p + jj, RH p + k) below the diagonal, zero
// elsewhere; dq[kk] = the pass's D coefficients: {1/d} for a 1x1 column,
// {d11, d21, d22, 1/det} for the first column of a 2x2 pair.
constexpr int LPW = NBMAX;
template <int NTH>
__device__ void stage_replay_pass(const SnCtx<double>& c, int kb, int pw, int p0, const int32_t* two_s, double* Lp,
                                  double4* dq) {
  constexpr int SB = 2048 / NTH;  // loads in flight per thread
  const double* L = c.P + kb + (int64_t)kb * c.h;
  const int nj = pw - p0, n = RH * LPW;
  for (int e0 = threadIdx.x; e0 < n; e0 += NTH * SB) {
    double v[SB];
#pragma unroll
    for (int u = 0; u < SB; ++u) {
      const int e = e0 + u * NTH, jj = e % LPW, k = e / LPW, gj = p0 + jj, gk = p0 + k;
      const bool in = e < n && jj < nj && gk < pw && gj > gk && !(gj == gk + 1 && two_s[gk]);
      v[u] = in ? L[gj + (int64_t)gk * c.h] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < SB; ++u)
      if (e0 + u * NTH < n) Lp[e0 + u * NTH] = v[u];
  }
  if (threadIdx.x < RH) {
    const int g = p0 + threadIdx.x;
    double4 q = make_double4(1.0, 0.0, 0.0, 1.0);
    if (g < pw) {
      if (two_s[g]) {
        const double d11 = c.d[kb + g], d21 = c.ds[kb + g], d22 = c.d[kb + g + 1];
        q = make_double4(d11, d21, d22, 1.0 / (d11 * d22 - d21 * d21));
      } else if (!(g > 0 && two_s[g - 1])) {
        q.x = 1.0 / c.d[kb + g];
      }
    }
    dq[threadIdx.x] = q;
  }
}

// D scaling of the pass's columns and their store (unpermuted). A pair whose
// second column opens the next pass is scaled by the caller.
__device__ __forceinline__ void replay_pass_store(const SnCtx<double>& c, int kb, int p0, int n, double (&x)[RH],
                                                  const int32_t* two_s, const double4* dq, bool live, int64_t row,
                                                  bool first_done, bool last_deferred) {
#pragma unroll
  for (int kk = 0; kk < RH; ++kk) {
    const int g = p0 + kk;
    const bool second = (kk == 0) ? first_done : (kk < n && two_s[g - 1] != 0);
    if (kk < n && !second && !(kk == n - 1 && last_deferred)) {
      const double4 q = dq[kk];
      if (kk + 1 < RH && two_s[g]) {
        const double wk = x[kk], wk1 = x[kk + 1];
        x[kk] = (q.z * wk - q.y * wk1) * q.w;
        x[kk + 1] = (q.x * wk1 - q.y * wk) * q.w;
      } else {
        x[kk] = x[kk] * q.x;
      }
    }
  }
  if (live) {
#pragma unroll
    for (int kk = 0; kk < RH; ++kk)
      if (kk < n && !(kk == 0 && first_done) && !(kk == n - 1 && last_deferred))
        c.P[row + (int64_t)(kb + p0 + kk) * c.h] = x[kk];
  }
}

// Rows [row0, row0 + nr), nr <= NTH (the workgroup size). Lp: RH * LPW
// doubles of LDS, dq: RH entries of LDS; ys: (pw - RH) * NTH doubles (scratch
// for the columns past the first pass; ys[(j - RH) * NTH + t] for column j of
// thread t's row).
template <int NTH>
__device__ __forceinline__ void replay_rows_reg(const SnCtx<double>& c, int kb, int ke, int64_t row0, int nr,
                                                double* Lp, double4* dq, const int32_t* cperm, const int32_t* two_s,
                                                double* ys) {
  const int pw = ke - kb;
  const int tid = threadIdx.x;
  const bool live = tid < nr;
  const int64_t row = row0 + tid;
  bool first_done = false;  // column p0 was scaled and stored with the previous pass's last one
  for (int p0 = 0; p0 < pw; p0 += RH) {
    const int n = min(RH, pw - p0);
    stage_replay_pass<NTH>(c, kb, pw, p0, two_s, Lp, dq);
    double x[RH];
#pragma unroll
    for (int j = 0; j < RH; ++j) {
      x[j] = 0.0;
      if (live && j < n) x[j] = p0 == 0 ? c.P[row + (int64_t)(kb + cperm[j]) * c.h] : ys[(int64_t)(p0 + j - RH) * NTH + tid];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RH - 1; ++k) {
      if (k + 1 < n) {
        const double xk = x[k];
#pragma unroll
        for (int j = k + 1; j < RH; ++j) x[j] = x[j] - xk * Lp[k * LPW + j];
      }
    }
    // later columns: every source column of this row is read (first pass)
    // before any of its columns is written
    const bool straddle = p0 + RH < pw && two_s[p0 + RH - 1];
    double s_last = 0.0, s_first = 0.0;
    constexpr int JB = 8;  // later columns per batch: their loads are issued together
    for (int jb = p0 + RH; jb < pw; jb += JB) {
      double y[JB];
#pragma unroll
      for (int u = 0; u < JB; ++u) {
        const int j = jb + u;
        y[u] = 0.0;
        if (live && j < pw) y[u] = p0 == 0 ? c.P[row + (int64_t)(kb + cperm[j]) * c.h] : ys[(int64_t)(j - RH) * NTH + tid];
      }
#pragma unroll
      for (int u = 0; u < JB; ++u) {
        const int j = jb + u;
        if (j < pw) {
#pragma unroll
          for (int k = 0; k < RH; ++k) y[u] = y[u] - x[k] * Lp[k * LPW + (j - p0)];
        }
      }
      if (jb == p0 + RH && straddle) {
        // pair (p0 + RH - 1, p0 + RH): both columns are final now
        const double4 q = dq[RH - 1];
        const double wk = x[RH - 1];
        s_last = (q.z * wk - q.y * y[0]) * q.w;
        s_first = (q.x * y[0] - q.y * wk) * q.w;
      }
      if (live) {
#pragma unroll
        for (int u = 0; u < JB; ++u)
          if (jb + u < pw) ys[(int64_t)(jb + u - RH) * NTH + tid] = y[u];
      }
    }
    replay_pass_store(c, kb, p0, n, x, two_s, dq, live, row, first_done, straddle);
    if (straddle && live) {
      c.P[row + (int64_t)(kb + p0 + RH - 1) * c.h] = s_last;
      c.P[row + (int64_t)(kb + p0 + RH) * c.h] = s_first;
    }
    first_done = straddle;
    __syncthreads();
  }
}

// Deep-row replay of the wide path on matrix cores: one unit is RU = 64 rows
// and four waves. Per pass of RPW = 16 columns (one column tile), wave 0 (one
// row per lane) runs the pass's triangle in ascending k as replay_rows_reg
// does, while waves 1-3 load the next pass's L11 columns and D entries; the
// pass's updates of every later column (the bulk of the flops) are one
// product  Y -= X_pass * L11(later, pass)^T  on v_mfma_f64_16x16x4f64, with Y
// held in registers across passes (wave w owns rows 16w..16w+15 of every
// 16-column tile, in the MFMA result layout); then every wave scales and
// stores 4 of the pass's columns. The L11 staging and D coefficients
// alternate between two LDS buffers by pass. The later-column sums run in
// the matrix core's order, not term by term; every other operation is the
// register replay's.
constexpr int RU = 64;
constexpr int RQLD = NBMAX + 4;  // Lq row stride (doubles)
constexpr int RXLD = RU + 4;     // Xq row stride

// Columns per pass of the matrix-core replay, and per wave in its D scaling.
constexpr int RPW = 16;
constexpr int RCW = RPW / 4;

// D scaling and store of RCW of a pass's columns (k0 .. k0+RCW-1) for
// replay_unit_mfma, one row per lane, from the pass's unscaled values held
// negated in LDS (xq: this row's entry of column 0, column stride RXLD). act:
// columns scaled here (1x1 pivots and 2x2 openers; not a pair's second
// column, nor an opener whose partner is in the next pass); pairs: the 2x2
// openers among them. A pair opened at k0+RCW-1 is finished and stored here, so
// a chunk skips a first column closed from the previous chunk. Same
// arithmetic as replay_pass_store.
__device__ __forceinline__ void replay_chunk_scale_store(const SnCtx<double>& c, int kb, int p0, int k0,
                                                         const double* xq, uint32_t act, uint32_t pairs,
                                                         const double4* dq, bool live, int64_t row) {
  constexpr int CW = RCW;
  static_assert(RPW <= 32 && RPW % 16 == 0, "one mask bit per column of a pass, whole column tiles");
  double v[CW + 1];
#pragma unroll
  for (int i = 0; i <= CW; ++i) v[i] = (k0 + i < RPW) ? -xq[(k0 + i) * RXLD] : 0.0;
#pragma unroll
  for (int i = 0; i < CW; ++i) {
    const int kk = k0 + i;
    const double4 q = dq[kk];
    const bool a = (act >> kk) & 1u, pr = (pairs >> kk) & 1u;
    const double wk = v[i], wk1 = v[i + 1];
    const double p1 = (q.z * wk - q.y * wk1) * q.w;
    const double p2 = (q.x * wk1 - q.y * wk) * q.w;
    v[i] = a ? (pr ? p1 : wk * q.x) : wk;
    v[i + 1] = pr ? p2 : wk1;
  }
  if (live) {
    const uint32_t computed = act | (pairs << 1);
#pragma unroll
    for (int i = 0; i <= CW; ++i) {
      const int kk = k0 + i;
      bool own;
      if (i < CW)
        own = ((computed >> kk) & 1u) && !(i == 0 && kk > 0 && ((pairs >> (kk - 1)) & 1u));
      else
        own = kk < RPW && ((pairs >> (kk - 1)) & 1u);
      if (own) c.P[row + (int64_t)(kb + p0 + kk) * c.h] = v[i];
    }
  }
}

struct ReplayMfmaLds {
  // two buffers, alternating by pass: Lq[b][k * RQLD + j] = L11(j, p0 + k)
  // for j > p0 + k (no 2x2 coupling), else 0
  double Lq[2][RPW * RQLD];
  double Xq[RPW * RXLD];  // the pass's columns: Xq[kk * RXLD + r] (wave 0 writes -x after the triangle)
  double4 dq[2][RPW];  // D coefficients of the pass's columns, alternating like Lq
};

__device__ void replay_unit_mfma(const SnCtx<double>& c, int kb, int ke, int64_t row0, int nr,
                                 ReplayMfmaLds& sh, const int32_t* cperm, const int32_t* two_s) {
  constexpr int NCT = NBMAX / 16;
  constexpr int SB = 8;           // staging loads in flight per thread (first pass)
  constexpr int NPF = NT - 64;    // waves 1-3 prefetch the next pass's staging
  constexpr int PF = (RPW * (NBMAX - RPW) + NPF - 1) / NPF;  // prefetched staging entries per thread
  static_assert(RPW * (NBMAX - RPW) <= PF * NPF, "a later pass's staging fits the prefetch registers");
  const int pw = ke - kb, tid = threadIdx.x, l = tid & 63, w = tid >> 6, cl = l & 15;
  const int rb = 16 * w + (l >> 4);  // unit row of result register r: rb + 4 r
  const double* L = c.P + kb + (int64_t)kb * c.h;
  // Staged columns of pass p0: p0 .. max(pw rounded up to a tile, p0 + RPW) - 1
  // (the triangle reads RPW of them, the products whole 16-column tiles).
  auto stage_cols = [&](int p0) { return min(NBMAX, max((pw + 15) & ~15, p0 + RPW)) - p0; };
  // bit kk: column p0 + kk opens a 2x2 pivot (wave-uniform)
  auto pass_opens = [&](int p0) {
    return (uint32_t)__ballot(l < RPW && p0 + l < pw && two_s[min(p0 + l, NBMAX - 1)] != 0);
  };
  // staging entry e of pass p0: L11(j, gk), zero outside the strict lower part and on 2x2 couplings
  auto lq_value = [&](int p0, int nj, uint32_t opens, int e) {
    const int k = e / nj, j = p0 + e % nj, gk = p0 + k;
    const bool in = j < pw && gk < pw && j > gk && !(j == gk + 1 && ((opens >> (k & 31)) & 1u));
    return in ? L[j + (int64_t)gk * c.h] : 0.0;
  };
  // D coefficients of column g: 1/d (1x1), (d11, d21, d22, 1/det) for a
  // pair's first column. dq_load reads the D entries (w: 1 for a 1x1 pivot,
  // 2 for a pair's first column, 0 otherwise), dq_make forms the coefficients,
  // so a prefetch does not wait for the loads.
  auto dq_load = [&](int g) {
    double4 r = make_double4(1.0, 0.0, 1.0, 0.0);
    if (g < pw) {
      if (two_s[g]) {
        r = make_double4(c.d[kb + g], c.ds[kb + g], c.d[kb + g + 1], 2.0);
      } else if (!(g > 0 && two_s[g - 1])) {
        r.x = c.d[kb + g];
        r.w = 1.0;
      }
    }
    return r;
  };
  auto dq_make = [](const double4& r) {
    double4 q = make_double4(1.0, 0.0, 0.0, 1.0);
    if (r.w == 2.0)
      q = make_double4(r.x, r.y, r.z, 1.0 / (r.x * r.z - r.y * r.y));
    else if (r.w == 1.0)
      q.x = 1.0 / r.x;
    return q;
  };
  // first pass: staging loads (issued ahead of the unit's row loads, so the
  // staging writes wait for them alone), then the rows, then the writes and
  // the D coefficients
  static_assert(RPW * NBMAX <= NT * SB, "the first pass's staging is one batch");
  const uint32_t opens0 = pass_opens(0);
  const int nj0 = stage_cols(0), tot0 = RPW * nj0;
  double v0[SB];
#pragma unroll
  for (int u = 0; u < SB; ++u) {
    const int e = tid + u * NT;
    v0[u] = e < tot0 ? lq_value(0, nj0, opens0, e) : 0.0;
  }
  mf_d4 y[NCT];
  {
    // interchange partners first, so the loads of the unit's rows go out together
    int cp[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) cp[ct] = cperm[min(ct * 16 + cl, pw - 1)];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      const int col = ct * 16 + cl;
      const double* src = c.P + row0 + (int64_t)(kb + cp[ct]) * c.h;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb + 4 * r;
        y[ct][r] = (col < pw && row < nr) ? src[row] : 0.0;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < SB; ++u) {
    const int e = tid + u * NT;
    if (e < tot0) sh.Lq[0][(e / nj0) * RQLD + e % nj0] = v0[u];
  }
  if (tid < RPW) sh.dq[0][tid] = dq_make(dq_load(tid));
  // the unit's rows have arrived before the pass loop (else the wait at the
  // loop head, for them, would also wait out every pass's stores)
  __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
  // waves 1-3 load the next pass's staging and D entries into registers
  // during the triangle
  double pre[PF];
  double4 dqn = make_double4(1.0, 0.0, 1.0, 0.0);
  double xprev = 0.0;  // wave 0: unscaled last column of the previous pass
  double4 qprev = make_double4(1.0, 0.0, 0.0, 1.0);
  bool first_done = false;
  const bool live = l < nr;
  const int64_t row = row0 + l;
  for (int p0 = 0; p0 < pw; p0 += RPW) {
    const int n = min(RPW, pw - p0), pc = p0 / 16;  // first column tile of the pass
    const uint32_t opens = pass_opens(p0);
    const bool straddle = p0 + RPW < pw && two_s[p0 + RPW - 1];
    // columns scaled in this pass: not a pair's second column (closed from
    // this pass or, at column 0, from the previous one), not an opener whose
    // partner is in the next pass
    const uint32_t in_pass = n >= 32 ? 0xffffffffu : ((1u << n) - 1u);
    const uint32_t act = in_pass & ~((opens << 1) | (first_done ? 1u : 0u)) & ~(straddle ? (1u << (n - 1)) : 0u);
    const int cb = (p0 / RPW) & 1;  // this pass's Lq / dq buffer
    // the pass's column tiles, from their owners' registers
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct)
      if (ct >= pc && ct < pc + RPW / 16) {
#pragma unroll
        for (int r = 0; r < 4; ++r) sh.Xq[((ct - pc) * 16 + cl) * RXLD + rb + 4 * r] = y[ct][r];
      }
    __syncthreads();
    if (w == 0) {
      double x[RPW];
#pragma unroll
      for (int kk = 0; kk < RPW; ++kk) x[kk] = sh.Xq[kk * RXLD + l];
      if (first_done && live) {
        // pair (p0 - 1, p0), both final: scaled and stored together
        const double wk = xprev, wk1 = x[0];
        c.P[row + (int64_t)(kb + p0 - 1) * c.h] = (qprev.z * wk - qprev.y * wk1) * qprev.w;
        c.P[row + (int64_t)(kb + p0) * c.h] = (qprev.x * wk1 - qprev.y * wk) * qprev.w;
      }
      // straight-line: Lq is zero in rows and columns at or past n, and so are
      // the columns of x past n, so the terms of those columns add nothing.
      // The coefficients stream through a ring of TW registers, each read TW
      // terms before its use; the scheduling barriers keep the reads from
      // being hoisted further (register pressure).
      // Terms run k ascending, j = k+1 .. RPW-1 within k; (pk, pj) is the term
      // whose coefficient is read next (all indices fold to constants).
      constexpr int TW = 24;
      const double* lq = sh.Lq[cb] + p0;
      double ring[TW];
      int pk = 0, pj = 1, t = 0;
#pragma unroll
      for (int u = 0; u < TW; ++u) {
        ring[u] = lq[pk * RQLD + pj];
        if (++pj == RPW) { ++pk; pj = pk + 1; }
      }
#pragma unroll
      for (int k = 0; k < RPW - 1; ++k) {
#pragma unroll
        for (int j = k + 1; j < RPW; ++j) {
          const double v = ring[t % TW];
          if (pk < RPW - 1) {
            ring[t % TW] = lq[pk * RQLD + pj];
            if (++pj == RPW) { ++pk; pj = pk + 1; }
          }
          x[j] = x[j] - x[k] * v;
          if (t % 8 == 7) __builtin_amdgcn_sched_barrier(0);
          ++t;
        }
      }
      xprev = x[RPW - 1];
      qprev = sh.dq[cb][RPW - 1];
#pragma unroll
      for (int kk = 0; kk < RPW; ++kk) sh.Xq[kk * RXLD + l] = -x[kk];
    } else if (p0 + RPW < pw) {
      const int q0 = p0 + RPW, nj = stage_cols(q0), tot = RPW * nj;
      const uint32_t on = pass_opens(q0);
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int e = (tid - 64) + u * NPF;
        pre[u] = e < tot ? lq_value(q0, nj, on, e) : 0.0;
      }
      if (tid - 64 < RPW) dqn = dq_load(q0 + tid - 64);
    }
    first_done = straddle;
    __syncthreads();
    if (p0 + RPW < pw) {
#pragma unroll
      for (int kk = 0; kk < RPW; kk += 4) {
        const double a = sh.Xq[(kk + (l >> 4)) * RXLD + 16 * w + cl];
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct)
          if (ct >= pc + RPW / 16 && ct * 16 < pw) {
            const double b = sh.Lq[cb][(kk + (l >> 4)) * RQLD + ct * 16 + cl];
            y[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, y[ct], 0, 0, 0);
          }
      }
      if (w > 0) {
        // the next pass's staging and D coefficients into the other buffer
        // (last read in the previous pass) before this pass's stores
        const int q0 = p0 + RPW, nj = stage_cols(q0), tot = RPW * nj;
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int e = (tid - 64) + u * NPF;
          if (e < tot) sh.Lq[cb ^ 1][(e / nj) * RQLD + q0 + e % nj] = pre[u];
        }
        if (tid - 64 < RPW) sh.dq[cb ^ 1][tid - 64] = dq_make(dqn);
      }
    }
    // the pass's D scaling and stores, RCW columns per wave
    replay_chunk_scale_store(c, kb, p0, RCW * w, sh.Xq + l, act, act & opens, sh.dq[cb], live, row);
    __syncthreads();
  }
}

template <class T>
struct CdivShared {
  union alignas(16) {
    T Ab[NBMAX * (NBMAX + 1) / 2];
    T stage[2 * STG];
    T rp[(RR + replay_chunk<T>() + 1) * NBMAX];  // replay: X then Lc
  } u;
  T l1s[NBMAX], l2s[NBMAX], w1s[NBMAX], w2s[NBMAX];
  int32_t deep_s[NBMAX], two_s[NBMAX], cperm_s[NBMAX], lperm_s[NBMAX];
  int cnt[4];
  int npert;
  int any_swap;
};

// One workgroup factors a whole supernode (all its nb-panels).
template <class T, int PR>
__global__ __launch_bounds__(NT) void k_cdiv_small(DevSym S, DevFac<T> Fc,
                                                   const int32_t* __restrict__ list, int nb_base,
                                                   PivotPolicy pol) {
  extern __shared__ unsigned char smem_raw[];
  CdivShared<T>& sh = *reinterpret_cast<CdivShared<T>*>(smem_raw);
  const int s = list[blockIdx.x];
  const int w = S.sn_ncol[s];
  const int nb = (w >= 512 && nb_base < 128) ? 128 : nb_base;
  SnCtx<T> c = make_ctx(S, Fc, s, nb);
  if (threadIdx.x == 0) sh.npert = 0;
  cdiv_init<NT>(c);
  __syncthreads();
  for (int kb = 0; kb < w; kb += nb) {
    const int ke = min(kb + nb, w);
    panel_load<NT>(c, kb, ke, sh.u.Ab);
    int st = panel_getf2<NT>(c, kb, ke, sh.u.Ab, sh.l1s, sh.l2s, sh.w1s, sh.w2s, sh.deep_s, sh.two_s, sh.lperm_s, pol, &sh.npert);
    if (st) {
      flag_failure(Fc, s, kb + st - 1);
      return;
    }
    panel_store<NT, true>(c, kb, ke, sh.u.Ab, sh.deep_s, sh.two_s, sh.lperm_s, sh.cperm_s, &sh.any_swap, sh.w1s);
    bool done = false;
    if constexpr (std::is_same<T, double>::value) {
      // LDS: the pass's L11 panel and D coefficients, then the running values
      // of the columns past the first pass (panels up to 2 RH wide)
      static_assert(sizeof(double) * (RH * LPW + 4 * RH + RH * NT) <= sizeof(sh.u.rp), "replay scratch");
      if (ke - kb <= 2 * RH) {
        double* Lp = reinterpret_cast<double*>(sh.u.rp);
        for (int64_t row0 = ke; row0 < c.h; row0 += NT)
          replay_rows_reg<NT>(c, kb, ke, row0, (int)min<int64_t>(NT, c.h - row0), Lp,
                          reinterpret_cast<double4*>(Lp + RH * LPW), sh.cperm_s, sh.two_s, Lp + RH * LPW + 4 * RH);
        done = true;
      }
    }
    if (!done)
      for (int64_t row0 = ke; row0 < c.h; row0 += RR)
        replay_rows_block(c, kb, ke, row0, (int)min<int64_t>(RR, c.h - row0), sh.u.rp, sh.u.rp + RR * NBMAX,
                          sh.cperm_s, sh.two_s);
    const int64_t mt = c.h - ke;
    const int cw = w - ke;
    if (cw > 0 && mt > 0) {
      const int nrt = (int)((mt + TM - 1) / TM), nct = (cw + TN - 1) / TN;
      for (int cb = 0; cb < nct; ++cb)
        for (int rb = cb; rb < nrt; ++rb) panel_schur_tile_call<PR>(c, kb, ke, rb * TM, cb * TN, sh.u.stage);
      __syncthreads();
    }
  }
  cdiv_emit(S, Fc, c, sh.cnt, sh.npert);
}

// ---- wide supernodes: batched over the level, one launch per phase of each
// panel step (every supernode of the level that has a p-th panel) ----
template <class T>
__global__ __launch_bounds__(GT) void k_wide_getf2(DevSym S, DevFac<T> Fc, const WItem* __restrict__ items,
                                                   PivotPolicy pol) {
  constexpr int G = GT;
  extern __shared__ unsigned char smem_raw[];
  CdivShared<T>& sh = *reinterpret_cast<CdivShared<T>*>(smem_raw);
  const WItem it = items[blockIdx.x];
  const int s = it.s, kb = it.kb;
  SnCtx<T> c = make_ctx(S, Fc, s, S.nb_eff[s]);
  if (Fc.status[0]) return;  // the factorization has already failed
  if (threadIdx.x == 0) sh.npert = 0;
  if (kb == 0) cdiv_init<G>(c);
  __syncthreads();
  const int ke = min(kb + c.nb, c.w);
  panel_load<G>(c, kb, ke, sh.u.Ab);
  int st = panel_getf2<G>(c, kb, ke, sh.u.Ab, sh.l1s, sh.l2s, sh.w1s, sh.w2s, sh.deep_s, sh.two_s, sh.lperm_s, pol, &sh.npert);
  if (st) {
    flag_failure(Fc, s, kb + st - 1);
    return;
  }
  panel_store<G, false>(c, kb, ke, sh.u.Ab, sh.deep_s, sh.two_s, sh.lperm_s, sh.cperm_s, &sh.any_swap, sh.w1s);
  if (threadIdx.x == 0 && sh.npert) atomicAdd(&Fc.counters[3], (unsigned long long)sh.npert);
}

// The panel's row interchanges applied to the already factored columns
// [0, kb) of its rows: blockIdx.y = the step's getf2 unit, blockIdx.x = a
// block of 64 columns. Every thread gathers its permuted sources into
// registers, the workgroup synchronizes, then they are stored.
template <class T>
__global__ __launch_bounds__(NT) void k_wide_swapcols(DevSym S, DevFac<T> Fc, const WItem* __restrict__ items) {
  __shared__ int32_t cperm[NBMAX];
  __shared__ int ident;
  const WItem it = items[blockIdx.y];
  const int s = it.s, kb = it.kb, j0 = 64 * blockIdx.x;
  if (j0 >= kb || Fc.status[0]) return;
  SnCtx<T> c = make_ctx(S, Fc, s, S.nb_eff[s]);
  const int pw = min(kb + c.nb, c.w) - kb;
  if (threadIdx.x == 0) ident = 1;
  __syncthreads();
  for (int q = threadIdx.x; q < pw; q += NT) {
    cperm[q] = c.deep[NBMAX + q];
    if (cperm[q] != q) ident = 0;
  }
  __syncthreads();
  if (ident) return;
  constexpr int PER = NBMAX * 64 / NT;
  const int nc = min(64, kb - j0), n = pw * nc;
  T v[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int e = threadIdx.x + u * NT;
    const int i = e % pw, j = e / pw;
    v[u] = e < n ? c.P[(kb + cperm[i]) + (int64_t)(j0 + j) * c.h] : F<T>::zero();
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int e = threadIdx.x + u * NT;
    const int i = e % pw, j = e / pw;
    if (e < n) c.P[(kb + i) + (int64_t)(j0 + j) * c.h] = v[u];
  }
}

template <class T>
__global__ __launch_bounds__(NT) void k_wide_replay(DevSym S, DevFac<T> Fc, const WItem* __restrict__ items) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ int32_t cperm[NBMAX];
  __shared__ int32_t two_s[NBMAX];
  const WItem it = items[blockIdx.x];
  const int s = it.s, kb = it.kb;
  SnCtx<T> c = make_ctx(S, Fc, s, S.nb_eff[s]);
  if (Fc.status[0]) return;
  const int ke = min(kb + c.nb, c.w), pw = ke - kb;
  const int64_t row0 = ke + (int64_t)it.a;
  const int nr = (int)min<int64_t>(REPLAY_ROWS, c.h - row0);
  for (int j = threadIdx.x; j < pw; j += blockDim.x) {
    cperm[j] = c.deep[NBMAX + j];
    two_s[j] = c.two[kb + j];
  }
  __syncthreads();
  if constexpr (std::is_same<T, double>::value) {
    static_assert(REPLAY_ROWS == RU, "one MFMA unit per work item");
    replay_unit_mfma(c, kb, ke, row0, nr, *reinterpret_cast<ReplayMfmaLds*>(smem_raw), cperm, two_s);
  } else {
    T* X = reinterpret_cast<T*>(smem_raw);  // [pw][RR] column-major
    T* Lc = X + RR * NBMAX;                 // [RCW + 1][NBMAX]: L11 columns of the current chunk
    for (int r0 = 0; r0 < nr; r0 += RR)
      replay_rows_block(c, kb, ke, row0 + r0, min(RR, nr - r0), X, Lc, cperm, two_s);
  }
}

template <class T>
constexpr size_t replay_smem() {
  return std::is_same<T, double>::value ? sizeof(ReplayMfmaLds) : sizeof(T) * (RR + replay_chunk<T>() + 1) * NBMAX;
}

template <class T, int PR>
__global__ __launch_bounds__(NT) void k_wide_schur(DevSym S, DevFac<T> Fc, const WItem* __restrict__ items) {
  __shared__ __align__(16) T stage[2 * STG];
  const WItem it = items[blockIdx.x];
  const int s = it.s, kb = it.kb;
  SnCtx<T> c = make_ctx(S, Fc, s, S.nb_eff[s]);
  if (Fc.status[0]) return;
  const int ke = min(kb + c.nb, c.w);
  panel_schur_tile<PR>(c, kb, ke, it.a * TM, it.b * TN, stage);
}

template <class T>
__global__ __launch_bounds__(NT) void k_wide_emit(DevSym S, DevFac<T> Fc, const int32_t* __restrict__ list) {
  __shared__ int cnt[4];
  const int s = list[blockIdx.x];
  SnCtx<T> c = make_ctx(S, Fc, s, S.nb_eff[s]);
  if (Fc.status[0]) return;
  cdiv_emit(S, Fc, c, cnt, 0);
}

// ---------------------------------------------------------------------------
// Host-callable launchers
// ---------------------------------------------------------------------------
template <class T>
size_t cdiv_smem() {
  return sizeof(CdivShared<T>);
}

template <class T>
void launch_fill(T* p, int64_t n, T v, hipStream_t st) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
}

template <class T>
void launch_rows_e(const DevSym& S, const DevFac<T>& Fc, int64_t nrows, hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_rows_e<T>, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, st, S, Fc, nrows);
}

template <class T>
void launch_assemble(const DevSym& S, const DevFac<T>& Fc, const T* vals, const double* scale,
                     hipStream_t st, LaunchProfiler* pf) {
  if (S.nnz <= 0) return;
  ProfScope ps(pf, LK_ASSEMBLE, 0.0, (double)S.nnz * (2 * sizeof(T) + 24), st);
  hipLaunchKernelGGL(k_assemble<T>, dim3((unsigned)((S.nnz + 255) / 256)), dim3(256), 0, st, S,
                     Fc, vals, scale);
}

// Reduced precisions apply to the real field only; the complex field always
// runs its FP64 (VALU) tile core.
#define RSL_PREC_SWITCH(T, prec, STMT)           \
  do {                                           \
    if (sizeof(T) == 16 || (prec) == TP_F64) {   \
      constexpr int P_ = TP_F64;                 \
      STMT;                                      \
    } else if ((prec) == TP_F32) {               \
      constexpr int P_ = TP_F32;                 \
      STMT;                                      \
    } else {                                     \
      constexpr int P_ = TP_BF16;                \
      STMT;                                      \
    }                                            \
  } while (0)

template <class T>
void launch_cmod(const DevSym& S, const DevFac<T>& Fc, const CItem* items, int nitems,
                 const CTask* tasks, double flops, int prec, hipStream_t st, LaunchProfiler* pf) {
  if (nitems <= 0) return;
  ProfScope ps(pf, LK_CMOD, flops, 0.0, st);
  if constexpr (std::is_same<T, double>::value) {
    RSL_PREC_SWITCH(T, prec, hipLaunchKernelGGL((k_cmod_mfma<P_>), dim3(nitems), dim3(NT), 0, st, S, Fc, items, tasks));
  } else {
    hipLaunchKernelGGL((k_cmod<T, TP_F64>), dim3(nitems), dim3(NT), 0, st, S, Fc, items, tasks);
  }
}

void launch_make_g(const DevSym& S, const DevFac<double>& Fc, const int32_t* list, int count, int max_w,
                   double* gpool, hipStream_t st) {
  if (count <= 0 || max_w <= 0) return;
  hipLaunchKernelGGL(k_make_g, dim3((unsigned)((max_w + 15) / 16), (unsigned)count), dim3(NT), 0, st, S, Fc, list,
                     gpool);
}

void launch_cmod_two_phase(const DevSym& S, const DevFac<double>& Fc, const CItem* items, int nitems,
                           const CTask* tasks, const PTile* ptiles, int64_t nptiles, void* cbuf, int64_t c_elems,
                           const double* gpool, double flops, int prec, hipStream_t st, LaunchProfiler* pf) {
  if (nitems <= 0 || nptiles <= 0) return;
  const size_t esz = prec == 0 ? sizeof(double) : sizeof(float);
  {
    ProfScope ps(pf, LK_CMOD, flops, 0.0, st);
    // FP64 from registers; the reduced-precision products measured faster
    // through the LDS-staged kernel
    if (prec == TP_F64)
      hipLaunchKernelGGL((k_cmod_prod_reg<TP_F64>), dim3((unsigned)nptiles), dim3(NT), 0, st, S, Fc, gpool, ptiles,
                         nptiles, cbuf);
    else if (prec == TP_F32)
      hipLaunchKernelGGL((k_cmod_prod<TP_F32>), dim3((unsigned)nptiles), dim3(NT), 0, st, S, Fc, ptiles, nptiles,
                         cbuf);
    else
      hipLaunchKernelGGL((k_cmod_prod<TP_BF16>), dim3((unsigned)nptiles), dim3(NT), 0, st, S, Fc, ptiles, nptiles,
                         cbuf);
  }
  ProfScope ps(pf, LK_CMOD_SUM, 0.0, (double)c_elems * esz, st);
  if (prec == 0)
    hipLaunchKernelGGL((k_cmod_sum<double>), dim3(nitems), dim3(NT), 0, st, S, Fc, items, tasks,
                       static_cast<const double*>(cbuf));
  else
    hipLaunchKernelGGL((k_cmod_sum<float>), dim3(nitems), dim3(NT), 0, st, S, Fc, items, tasks,
                       static_cast<const float*>(cbuf));
}

template <class T>
void launch_cdiv_small(const DevSym& S, const DevFac<T>& Fc, const int32_t* list, int count,
                       int nb, PivotPolicy pol, double flops, int prec, hipStream_t st, LaunchProfiler* pf) {
  if (count <= 0) return;
  ProfScope ps(pf, LK_CDIV_SMALL, flops, 0.0, st);
  RSL_PREC_SWITCH(T, prec, hipLaunchKernelGGL((k_cdiv_small<T, P_>), dim3(count), dim3(NT), cdiv_smem<T>(), st, S,
                                              Fc, list, nb, pol));
}

template <class T>
void launch_wide_level(const DevSym& S, const DevFac<T>& Fc, const WItem* witems, const WideStep* steps,
                       int nsteps, const int32_t* wide_list, int nwide, PivotPolicy pol, int prec, hipStream_t st, LaunchProfiler* pf) {
  if (nwide <= 0) return;
  const double fm = sizeof(T) == 16 ? 4.0 : 1.0;
  const size_t replay_lds = replay_smem<T>();
  for (int p = 0; p < nsteps; ++p) {
    const WideStep& w = steps[p];
    if (w.g_end > w.g_beg) {
      ProfScope ps(pf, LK_PANEL_GETF2, fm * w.getf2_flops, 0.0, st);
      hipLaunchKernelGGL(k_wide_getf2<T>, dim3((unsigned)(w.g_end - w.g_beg)), dim3(GT), cdiv_smem<T>(), st,
                         S, Fc, witems + w.g_beg, pol);
      if (w.max_kb > 0)
        hipLaunchKernelGGL(k_wide_swapcols<T>, dim3((unsigned)((w.max_kb + 63) / 64), (unsigned)(w.g_end - w.g_beg)),
                           dim3(NT), 0, st, S, Fc, witems + w.g_beg);
    }
    if (w.r_end > w.r_beg) {
      ProfScope ps(pf, LK_PANEL_REPLAY, fm * w.replay_flops, 0.0, st);
      hipLaunchKernelGGL(k_wide_replay<T>, dim3((unsigned)(w.r_end - w.r_beg)), dim3(NT), replay_lds, st, S,
                         Fc, witems + w.r_beg);
    }
    if (w.t_end > w.t_beg) {
      ProfScope ps(pf, LK_PANEL_SCHUR, fm * w.schur_flops, 0.0, st);
      RSL_PREC_SWITCH(T, prec, hipLaunchKernelGGL((k_wide_schur<T, P_>), dim3((unsigned)(w.t_end - w.t_beg)),
                                                  dim3(NT), 0, st, S, Fc, witems + w.t_beg));
    }
  }
  ProfScope ps(pf, LK_EMIT, 0.0, 0.0, st);
  hipLaunchKernelGGL(k_wide_emit<T>, dim3((unsigned)nwide), dim3(NT), 0, st, S, Fc, wide_list);
}

// Drop tolerance at emit (ldlt_emit_and_free, multifrontal_ldlt.rs:2344-2366):
// in every stored column, the entries below the unit diagonal (past the 2x2
// coupling row) whose magnitude is not >= tau * (largest such magnitude) are
// zeroed; the downloaded L then omits them as the reference's emit does, and
// the solve reads zeros there. One workgroup per supernode, one wave per
// column (columns are contiguous, leading dimension h).
template <class T>
__global__ __launch_bounds__(256) void k_drop(DevSym S, DevFac<T> Fc, double tau) {
  const int s = blockIdx.x;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s], eo = S.e_off[s];
  T* P = Fc.pool + S.panel_off[s];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int p = wv; p < w; p += 4) {  // wave-uniform: every lane reaches the reduction
    const int64_t i0 = Fc.two[eo + p] ? p + 2 : p + 1;
    T* col = P + (int64_t)p * h;
    double m = 0.0;
    for (int64_t i = i0 + lane; i < h; i += 64) m = fmax(m, F<T>::mag(col[i]));
    const double thr = tau * wave_max_dpp(m);
    for (int64_t i = i0 + lane; i < h; i += 64)
      if (!(F<T>::mag(col[i]) >= thr)) col[i] = F<T>::zero();
  }
}

template <class T>
void launch_drop(const DevSym& S, const DevFac<T>& Fc, double tau, hipStream_t st) {
  if (S.ns > 0) hipLaunchKernelGGL(k_drop<T>, dim3((unsigned)S.ns), dim3(256), 0, st, S, Fc, tau);
}

template <class T>
hipError_t configure_factor_kernels() {
  hipError_t e = hipSuccess;
  for (int prec = TP_F64; prec <= TP_BF16 && e == hipSuccess; ++prec)
    RSL_PREC_SWITCH(T, prec, e = hipFuncSetAttribute((const void*)k_cdiv_small<T, P_>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)cdiv_smem<T>()));
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)k_wide_replay<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)replay_smem<T>());
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)k_wide_getf2<T>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)cdiv_smem<T>());
}

#define RSL_INST(T)                                                                            \
  template void launch_fill<T>(T*, int64_t, T, hipStream_t);                                   \
  template void launch_rows_e<T>(const DevSym&, const DevFac<T>&, int64_t, hipStream_t);        \
  template void launch_assemble<T>(const DevSym&, const DevFac<T>&, const T*, const double*,   \
                                   hipStream_t, LaunchProfiler*);                              \
  template void launch_cmod<T>(const DevSym&, const DevFac<T>&, const CItem*, int,             \
                               const CTask*, double, int, hipStream_t, LaunchProfiler*);       \
  template void launch_cdiv_small<T>(const DevSym&, const DevFac<T>&, const int32_t*, int, int, \
                                     PivotPolicy, double, int, hipStream_t, LaunchProfiler*);  \
  template void launch_wide_level<T>(const DevSym&, const DevFac<T>&, const WItem*, const WideStep*, \
                                     int, const int32_t*, int, PivotPolicy, int, hipStream_t,    \
                                     LaunchProfiler*);                              \
  template hipError_t configure_factor_kernels<T>();                                           \
  template void launch_drop<T>(const DevSym&, const DevFac<T>&, double, hipStream_t);          \
  template size_t cdiv_smem<T>();

RSL_INST(double)
RSL_INST(cplx)

}  // namespace rsl
