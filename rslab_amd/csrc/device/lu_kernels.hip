// Device kernels of the LU twin: supernodal left-looking LU with threshold
// partial pivoting restricted to each supernode's own rows
// (lu_ll_factor_node, src/numeric/multifrontal_lu.rs:1629-2347), and its
// supernodal forward / backward solve (the arithmetic of solve_lu_block,
// :3209-3283, grouped by supernode).
//
// Storage per supernode s (h rows, w own columns, rows in rs[s] order with the
// own rows first): L panel h x w column-major (unit-lower L below the
// diagonal, U11 on and above it) and U12 w x (h - w) column-major. The own
// rows of the panel are interchanged by the pivoting (rperm records which
// row-structure row sits at each position); the deep rows never are.
//
// Launch sequence of a tree level: cmod (every target tile of the level's
// supernodes sums its updaters' products L_k[rows] U_k[:, cols]), then per
// panel index: getf2 (one workgroup per supernode panel), apply (the panel's
// interchanges on the other columns and the triangular solve of the later
// own columns and of U12, the deep rows' transform), and the trailing GEMM.
#include <hip/hip_runtime.h>

#include <climits>

#include "lu_types.hpp"
#include "scalar.hpp"
#include "tile_mfma.hpp"

namespace rsl {

namespace {

template <class T>
__device__ inline T lu_fma(T a, T b, T c) {
  return c + a * b;
}
template <>
__device__ inline double lu_fma<double>(double a, double b, double c) {
  return fma(a, b, c);
}

// value of lane q (wave-uniform) across the wave, read through the scalar unit
__device__ inline double lu_bcast(double v, int q) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), q);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), q);
  return __hiloint2double(hi, lo);
}
__device__ inline cplx lu_bcast(cplx v, int q) { return {lu_bcast(v.re, q), lu_bcast(v.im, q)}; }

// perturb_pivot (multifrontal_ldlt.rs:786-795)
template <class T>
__device__ inline T lu_perturb(T d, double floor) {
  const double mag = F<T>::mag(d);
  if (mag >= floor) return d;
  if (mag == 0.0) return F<T>::from_real(floor);
  return F<T>::scale(d, floor / mag);
}

// ---- 64x64 tile products on the vector units: thread (ty, tx) owns rows
// 4 ty .. 4 ty + 3 and columns 4 tx .. 4 tx + 3; operands staged through LDS
// 16 columns of K at a time.
constexpr int KC = 16;
constexpr int LDT = 64 + 4;

template <class T>
struct LuAcc {
  T c[4][4];
};

template <class T>
__device__ inline void lu_acc_zero(LuAcc<T>& a) {
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) a.c[r][q] = F<T>::zero();
}

// acc += As(64 x kc) Bs(kc x 64) from the staged chunk
template <class T>
__device__ inline void lu_acc_chunk(const T (*As)[LDT], const T (*Bs)[LDT], int kc, LuAcc<T>& acc) {
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  for (int c = 0; c < kc; ++c) {
    T a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) a[r] = As[c][ty * 4 + r];
#pragma unroll
    for (int q = 0; q < 4; ++q) b[q] = Bs[c][tx * 4 + q];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc.c[r][q] = lu_fma(a[r], b[q], acc.c[r][q]);
  }
}

}  // namespace

// Scatter of the equilibrated, permuted values into the zeroed L and U12
// storage (the assembly of lu_ll_factor_node, multifrontal_lu.rs:1670-1686).
template <class T>
__global__ __launch_bounds__(LU_NT) void k_lu_assemble(LuSymDev S, LuFacDev<T> Fc, const T* __restrict__ vals) {
  const int64_t k = (int64_t)blockIdx.x * LU_NT + threadIdx.x;
  if (k >= S.nnz) return;
  const int64_t d = S.amap[k];
  if (d >= 0) Fc.l[d] = vals[k];
  else Fc.u[-1 - d] = vals[k];
}

// cmod: one workgroup per target tile; every updater's product (ascending)
// is accumulated, the sum is subtracted once.
template <class T>
__global__ __launch_bounds__(LU_NT) void k_lu_cmod(LuSymDev S, LuFacDev<T> Fc, const LuItem* __restrict__ items,
                                                   const CTask* __restrict__ tasks) {
  __shared__ T As[KC][LDT];
  __shared__ T Bs[KC][LDT];
  const int tid = threadIdx.x;
  const LuItem it = items[blockIdx.x];
  LuAcc<T> acc;
  lu_acc_zero(acc);
  for (int t = it.t_beg; t < it.t_end; ++t) {
    const CTask tk = tasks[t];
    const int k = tk.k, nck = S.sn_ncol[k];
    const int64_t hk = S.sn_nrow[k];
    const T* Lk = Fc.l + S.l_off[k];
    const T* Uk = Fc.u + S.u_off[k];
    for (int c0 = 0; c0 < nck; c0 += KC) {
      const int kc = min(KC, nck - c0);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = tid + LU_NT * u, r = e & 63, c = e >> 6;
        const uint64_t below = (1ull << r) - 1ull;
        T va = F<T>::zero(), vb = F<T>::zero();
        if (c < kc) {
          if ((tk.rmask >> r) & 1ull) va = Lk[(nck + tk.a0 + __popcll(tk.rmask & below)) + (int64_t)(c0 + c) * hk];
          if ((tk.cmask >> r) & 1ull) vb = Uk[(c0 + c) + (int64_t)(tk.b0 + __popcll(tk.cmask & below)) * nck];
        }
        As[c][r] = va;
        Bs[c][r] = vb;
      }
      __syncthreads();
      lu_acc_chunk<T>(As, Bs, kc, acc);
      __syncthreads();
    }
  }
  const int s = it.s;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s];
  T* P;
  int64_t ld, mr, mc;
  if (it.region == 0) {
    P = Fc.l + S.l_off[s];
    ld = h;
    mr = h - it.r0;
    mc = w - it.c0;
  } else {
    P = Fc.u + S.u_off[s];
    ld = w;
    mr = w - it.r0;
    mc = h - w - it.c0;
  }
  const int ty = tid >> 4, tx = tid & 15;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = ty * 4 + r, j = tx * 4 + q;
      if (i < mr && j < mc) {
        T* dst = P + (it.r0 + i) + (int64_t)(it.c0 + j) * ld;
        *dst = *dst - acc.c[r][q];
      }
    }
}

// cmod of the real field on the matrix cores (v_mfma_f64_16x16x4f64): the
// same target tiles and task order as k_lu_cmod, operands gathered into LDS
// 32 columns of K at a time (the next chunk's loads issued before the current
// chunk's products), every updater's product accumulated in FP64.
__global__ __launch_bounds__(LU_NT) void k_lu_cmod_mfma(LuSymDev S, LuFacDev<double> Fc,
                                                        const LuItem* __restrict__ items,
                                                        const CTask* __restrict__ tasks) {
  typedef TileCfgW<TP_F64> C;
  constexpr int KCW = C::KC, PER = KCW * 64 / LU_NT;
  __shared__ __align__(16) double As[C::ELEMS];
  __shared__ __align__(16) double Bs[C::ELEMS];
  const int tid = threadIdx.x;
  const LuItem it = items[blockIdx.x];
  double va[PER], vb[PER];
  auto fetch = [&](int tt, int cc) {
    const CTask tk = tasks[tt];
    const int k = tk.k, nck = S.sn_ncol[k];
    const int64_t hk = S.sn_nrow[k];
    const double* Lk = Fc.l + S.l_off[k];
    const double* Uk = Fc.u + S.u_off[k];
    const int kc = min(KCW, nck - cc);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + LU_NT * u, r = e & 63, c = e >> 6;
      const uint64_t below = (1ull << r) - 1ull;
      va[u] = 0.0;
      vb[u] = 0.0;
      if (c < kc) {
        if ((tk.rmask >> r) & 1ull) va[u] = Lk[(nck + tk.a0 + __popcll(tk.rmask & below)) + (int64_t)(cc + c) * hk];
        if ((tk.cmask >> r) & 1ull) vb[u] = Uk[(cc + c) + (int64_t)(tk.b0 + __popcll(tk.cmask & below)) * nck];
      }
    }
  };
  int t = it.t_beg, c0 = 0;
  int nck_t = S.sn_ncol[tasks[t].k];
  fetch(t, 0);
  TileAcc<TP_F64> acc;
  tile_zero(acc);
  for (;;) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + LU_NT * u, r = e & 63, c = e >> 6;
      stage_store_w<TP_F64>(As, r, c, va[u]);
      stage_store_w<TP_F64>(Bs, r, c, vb[u]);
    }
    __syncthreads();
    int tn = t, cn = c0 + KCW, nck_n = nck_t;
    if (cn >= nck_t) {
      tn = t + 1;
      cn = 0;
      if (tn < it.t_end) nck_n = S.sn_ncol[tasks[tn].k];
    }
    if (tn < it.t_end) fetch(tn, cn);
    tile_mma_w<TP_F64>(As, Bs, acc);
    __syncthreads();
    if (tn >= it.t_end) break;
    t = tn;
    c0 = cn;
    nck_t = nck_n;
  }
  const int s = it.s;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s];
  double* P;
  int64_t ld, mr, mc;
  if (it.region == 0) {
    P = Fc.l + S.l_off[s];
    ld = h;
    mr = h - it.r0;
    mc = w - it.c0;
  } else {
    P = Fc.u + S.u_off[s];
    ld = w;
    mr = w - it.r0;
    mc = h - w - it.c0;
  }
  // the targets read as one batch, then updated
  double old[16];
  int n = 0;
  tile_for_each<TP_F64>(acc, [&](int i, int j, double) {
    old[n++] = (i < mr && j < mc) ? P[(it.r0 + i) + (int64_t)(it.c0 + j) * ld] : 0.0;
  });
  n = 0;
  tile_for_each<TP_F64>(acc, [&](int i, int j, double v) {
    if (i < mr && j < mc) P[(it.r0 + i) + (int64_t)(it.c0 + j) * ld] = old[n] - v;
    ++n;
  });
}

// cmod of the complex field on the FP64 matrix cores as four real products
// per complex one (C = (Ar Br - Ai Bi) + i (Ar Bi + Ai Br)), the operands'
// real and imaginary parts gathered into registers once per 32-column chunk
// and staged through LDS pairwise.
__global__ __launch_bounds__(LU_NT) void k_lu_cmod_mfma_c64(LuSymDev S, LuFacDev<cplx> Fc,
                                                            const LuItem* __restrict__ items,
                                                            const CTask* __restrict__ tasks) {
  typedef TileCfgW<TP_F64> C;
  constexpr int KCW = C::KC, PER = KCW * 64 / LU_NT;
  __shared__ __align__(16) double As[C::ELEMS];
  __shared__ __align__(16) double Bs[C::ELEMS];
  const int tid = threadIdx.x;
  const LuItem it = items[blockIdx.x];
  cplx va[PER], vb[PER];
  TileAcc<TP_F64> rr, ii, ri, ir;
  tile_zero(rr);
  tile_zero(ii);
  tile_zero(ri);
  tile_zero(ir);
  auto stage = [&](bool aim, bool bim) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + LU_NT * u, r = e & 63, c = e >> 6;
      stage_store_w<TP_F64>(As, r, c, aim ? va[u].im : va[u].re);
      stage_store_w<TP_F64>(Bs, r, c, bim ? vb[u].im : vb[u].re);
    }
  };
  for (int t = it.t_beg; t < it.t_end; ++t) {
    const CTask tk = tasks[t];
    const int k = tk.k, nck = S.sn_ncol[k];
    const int64_t hk = S.sn_nrow[k];
    const cplx* Lk = Fc.l + S.l_off[k];
    const cplx* Uk = Fc.u + S.u_off[k];
    for (int c0 = 0; c0 < nck; c0 += KCW) {
      const int kc = min(KCW, nck - c0);
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int e = tid + LU_NT * u, r = e & 63, c = e >> 6;
        const uint64_t below = (1ull << r) - 1ull;
        va[u] = {0.0, 0.0};
        vb[u] = {0.0, 0.0};
        if (c < kc) {
          if ((tk.rmask >> r) & 1ull) va[u] = Lk[(nck + tk.a0 + __popcll(tk.rmask & below)) + (int64_t)(c0 + c) * hk];
          if ((tk.cmask >> r) & 1ull) vb[u] = Uk[(c0 + c) + (int64_t)(tk.b0 + __popcll(tk.cmask & below)) * nck];
        }
      }
      stage(false, false);
      __syncthreads();
      tile_mma_w<TP_F64>(As, Bs, rr);
      __syncthreads();
      stage(true, true);
      __syncthreads();
      tile_mma_w<TP_F64>(As, Bs, ii);
      __syncthreads();
      stage(false, true);
      __syncthreads();
      tile_mma_w<TP_F64>(As, Bs, ri);
      __syncthreads();
      stage(true, false);
      __syncthreads();
      tile_mma_w<TP_F64>(As, Bs, ir);
      __syncthreads();
    }
  }
  const int s = it.s;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s];
  cplx* P;
  int64_t ld, mr, mc;
  if (it.region == 0) {
    P = Fc.l + S.l_off[s];
    ld = h;
    mr = h - it.r0;
    mc = w - it.c0;
  } else {
    P = Fc.u + S.u_off[s];
    ld = w;
    mr = w - it.r0;
    mc = h - w - it.c0;
  }
  const int l = tid & 63, wv = tid >> 6, wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = wr + tacc_row<TP_F64>(bi, l, q), j = wc + bj * 16 + (l & 15);
        if (i < mr && j < mc) {
          cplx* dst = P + (it.r0 + i) + (int64_t)(it.c0 + j) * ld;
          const double re = rr.v[bi][bj][q] - ii.v[bi][bj][q], im = ri.v[bi][bj][q] + ir.v[bi][bj][q];
          *dst = {dst->re - re, dst->im - im};
        }
      }
}

// getf2 on the own rows of one panel (multifrontal_lu.rs:2064-2111): per
// column, the pivot is the first row of largest magnitude in [k, w); the rows
// are interchanged only when the diagonal is below u times it; the
// interchange is applied to the panel's columns here and recorded (ipiv,
// rperm) for the other columns; then the column is scaled and the panel's
// later columns get the rank-1 update.
template <class T>
__global__ __launch_bounds__(LU_GT) void k_lu_getf2(LuSymDev S, LuFacDev<T> Fc, const LuUnit* __restrict__ units) {
  __shared__ double red_m[LU_GT / 64];
  __shared__ int red_i[LU_GT / 64];
  __shared__ T ukj[128];
  __shared__ int sh_p;
  __shared__ T sh_pinv;
  const LuUnit un = units[blockIdx.x];
  const int s = un.s, kb = un.kb;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s], first = S.sn_first[s], ke = min(kb + S.nb[s], w);
  T* L = Fc.l + S.l_off[s];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int k = kb; k < ke; ++k) {
    T* ck = L + (int64_t)k * h;
    int p = k;
    if (Fc.thresh_sq > 0.0) {
      double m = -1.0;
      int mi = INT_MAX;
      for (int i = k + tid; i < w; i += LU_GT) {
        const double v = F<T>::magsq(ck[i]);
        if (v > m) {
          m = v;
          mi = i;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double om = __shfl_xor(m, off);
        const int oi = __shfl_xor(mi, off);
        if (om > m || (om == m && oi < mi)) {
          m = om;
          mi = oi;
        }
      }
      if (lane == 0) {
        red_m[wv] = m;
        red_i[wv] = mi;
      }
      __syncthreads();
      if (tid == 0) {
        double bm = red_m[0];
        int bi = red_i[0];
        for (int q = 1; q < LU_GT / 64; ++q)
          if (red_m[q] > bm || (red_m[q] == bm && red_i[q] < bi)) {
            bm = red_m[q];
            bi = red_i[q];
          }
        const double dsq = F<T>::magsq(ck[k]);
        sh_p = (bi != k && dsq < Fc.thresh_sq * bm) ? bi : k;
      }
      __syncthreads();
      p = sh_p;
    }
    if (p != k) {
      for (int j = kb + tid; j < ke; j += LU_GT) {
        T* cj = L + (int64_t)j * h;
        const T t = cj[k];
        cj[k] = cj[p];
        cj[p] = t;
      }
      if (tid == 0) {
        int32_t* rp = Fc.rperm + first;
        const int32_t t = rp[k];
        rp[k] = rp[p];
        rp[p] = t;
      }
    }
    __syncthreads();
    if (tid == 0) {
      Fc.ipiv[first + k] = p;
      T piv = ck[k];
      if (Fc.pol.have_floor) {
        if (F<T>::mag(piv) < Fc.pol.floor) {
          piv = lu_perturb(piv, Fc.pol.floor);
          atomicAdd(&Fc.counters[0], 1ull);
        }
      } else if (F<T>::is_zero(piv)) {
        Fc.status[0] = 1;
        atomicMin(&Fc.counters[1], (unsigned long long)(first + k));
      }
      ck[k] = piv;
      sh_pinv = F<T>::recip(piv);
    }
    for (int j = k + 1 + tid; j < ke; j += LU_GT) ukj[j - kb] = L[(int64_t)j * h + k];
    __syncthreads();
    const T pinv = sh_pinv;
    for (int i = k + 1 + tid; i < w; i += LU_GT) {
      const T lik = ck[i] * pinv;
      ck[i] = lik;
      for (int j = k + 1; j < ke; ++j) {
        const T m = ukj[j - kb];
        if (!F<T>::is_zero(m)) {
          T* cj = L + (int64_t)j * h;
          cj[i] = cj[i] - lik * m;
        }
      }
    }
    __syncthreads();
  }
}

// After a panel's getf2: column units apply the panel's interchanges to the
// columns outside the panel (the finished ones left of it, the later own
// columns and U12) and, on the later ones, the unit-lower triangular solve
// with L11 (the TRSM of :2157-2199); deep-row units apply the panel transform
// to the rows below the own rows (apply_panel_trailing, :1357-1387). The
// panel's diagonal block and pivot reciprocals sit in LDS, each thread's
// column or row segment in registers (the same operations in the same order
// as on the stored values).
template <class T>
__global__ __launch_bounds__(LU_NT) void k_lu_apply(LuSymDev S, LuFacDev<T> Fc, const LuUnit* __restrict__ units) {
  __shared__ T Ls[LU_PANEL][LU_PANEL + 1];  // the panel's diagonal block: Ls[r][c] = entry (kb + r, kb + c)
  __shared__ T pinv[LU_PANEL];
  __shared__ int pv[LU_PANEL];
  const LuUnit un = units[blockIdx.x];
  const int s = un.s, kb = un.kb;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s], first = S.sn_first[s], ke = min(kb + S.nb[s], w), pw = ke - kb;
  T* L = Fc.l + S.l_off[s];
  for (int e = threadIdx.x; e < LU_PANEL * LU_PANEL; e += LU_NT) {
    const int r = e % LU_PANEL, c = e / LU_PANEL;
    Ls[r][c] = (r < pw && c < pw) ? L[(int64_t)(kb + c) * h + kb + r] : F<T>::zero();
  }
  if (threadIdx.x < LU_PANEL) pv[threadIdx.x] = threadIdx.x < pw ? Fc.ipiv[first + kb + threadIdx.x] : 0;
  __syncthreads();
  if (threadIdx.x < LU_PANEL) pinv[threadIdx.x] = threadIdx.x < pw ? F<T>::recip(Ls[threadIdx.x][threadIdx.x]) : F<T>::zero();
  __syncthreads();
  if (un.b == 0) {
    const int64_t c = (int64_t)un.a + threadIdx.x;
    const int64_t nlater = w - ke, ncols = kb + nlater + (h - w);
    if (c >= ncols) return;
    T* col;
    bool later;
    if (c < kb) {
      col = L + c * h;
      later = false;
    } else if (c < kb + nlater) {
      col = L + (ke + c - kb) * h;
      later = true;
    } else {
      col = Fc.u + S.u_off[s] + (c - kb - nlater) * w;
      later = true;
    }
    for (int q = 0; q < pw; ++q) {
      const int p = pv[q];
      if (p != kb + q) {
        const T t = col[kb + q];
        col[kb + q] = col[p];
        col[p] = t;
      }
    }
    if (later) {
      T v[LU_PANEL];
#pragma unroll
      for (int r = 0; r < LU_PANEL; ++r) v[r] = r < pw ? col[kb + r] : F<T>::zero();
      for (int r = 1; r < pw; ++r) {  // a runtime row: the block's entries are read per row, not hoisted
        T acc = F<T>::zero();
#pragma unroll
        for (int i = 0; i < LU_PANEL; ++i)
          if (i == r) acc = v[i];
#pragma unroll
        for (int i = 0; i < LU_PANEL; ++i)
          if (i < r) acc = acc - Ls[r][i] * v[i];
#pragma unroll
        for (int i = 0; i < LU_PANEL; ++i)
          if (i == r) v[i] = acc;
      }
#pragma unroll
      for (int r = 1; r < LU_PANEL; ++r)
        if (r < pw) col[kb + r] = v[r];
    }
  } else {
    const int64_t i = (int64_t)w + un.a + threadIdx.x;
    if (i >= h) return;
    T x[LU_PANEL];
#pragma unroll
    for (int q = 0; q < LU_PANEL; ++q) x[q] = q < pw ? L[(int64_t)(kb + q) * h + i] : F<T>::zero();
    for (int q = 0; q < pw; ++q) {
      T xq = F<T>::zero();
#pragma unroll
      for (int jj = 0; jj < LU_PANEL; ++jj)
        if (jj == q) xq = x[jj];
      xq = xq * pinv[q];
#pragma unroll
      for (int jj = 0; jj < LU_PANEL; ++jj) {
        if (jj == q) x[jj] = xq;
        if (jj > q) {
          const T m = Ls[q][jj];
          if (!F<T>::is_zero(m)) x[jj] = x[jj] - xq * m;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < LU_PANEL; ++q)
      if (q < pw) L[(int64_t)(kb + q) * h + i] = x[q];
  }
}

// Trailing GEMMs of a panel (:2206-2280): L[ke.., ke..w) -= L[ke.., kb..ke)
// U[kb..ke, ke..w) and U12[ke..w, :] -= L[ke..w, kb..ke) U12[kb..ke, :], one
// 64x64 tile per workgroup.
template <class T>
__global__ __launch_bounds__(LU_NT) void k_lu_gemm(LuSymDev S, LuFacDev<T> Fc, const LuUnit* __restrict__ units) {
  __shared__ T As[KC][LDT];
  __shared__ T Bs[KC][LDT];
  const LuUnit un = units[blockIdx.x];
  const int s = un.s, kb = un.kb;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s], ke = min(kb + S.nb[s], w), pw = ke - kb;
  const T* L = Fc.l + S.l_off[s];
  const bool upart = un.a < 0;
  const int64_t r0 = (int64_t)ke + (int64_t)(upart ? -1 - un.a : un.a) * 64;
  const int64_t c0 = (int64_t)un.b * 64;
  T* P;
  int64_t ld, mr, mc;
  const T* B;
  if (!upart) {
    P = Fc.l + S.l_off[s] + (int64_t)ke * h;  // column ke + c0 + j at P + (c0 + j) h
    ld = h;
    mr = h - r0;
    mc = w - ke - c0;
    B = L + (int64_t)ke * h + kb;
  } else {
    P = Fc.u + S.u_off[s];
    ld = w;
    mr = w - r0;
    mc = h - w - c0;
    B = Fc.u + S.u_off[s] + kb;
  }
  const int tid = threadIdx.x;
  LuAcc<T> acc;
  lu_acc_zero(acc);
  for (int q0 = 0; q0 < pw; q0 += KC) {
    const int kc = min(KC, pw - q0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + LU_NT * u, r = e & 63, c = e >> 6;
      T va = F<T>::zero(), vb = F<T>::zero();
      if (c < kc) {
        if (r < mr) va = L[(int64_t)(kb + q0 + c) * h + r0 + r];
        if (r < mc) vb = B[(q0 + c) + (c0 + r) * ld];
      }
      As[c][r] = va;
      Bs[c][r] = vb;
    }
    __syncthreads();
    lu_acc_chunk<T>(As, Bs, kc, acc);
    __syncthreads();
  }
  const int ty = tid >> 4, tx = tid & 15;
  // the targets read as one batch, then updated
  T old[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = ty * 4 + r, j = tx * 4 + q;
      old[r][q] = (i < mr && j < mc) ? P[(r0 + i) + (c0 + j) * ld] : F<T>::zero();
    }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = ty * 4 + r, j = tx * 4 + q;
      if (i < mr && j < mc) P[(r0 + i) + (c0 + j) * ld] = old[r][q] - acc.c[r][q];
    }
}

// ---- solve (row-major n x nrhs vectors)

// bh[g] = d_row[o] b[o] for the original row o = perm[g]
template <class T>
__global__ void k_lu_sprep(int32_t n, int32_t nrhs, const int32_t* __restrict__ perm, const double* __restrict__ d_row,
                           const T* __restrict__ b, T* __restrict__ bh) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * nrhs) return;
  const int64_t g = e / nrhs, c = e - g * nrhs, o = perm[g];
  bh[e] = F<T>::scale(b[o * nrhs + c], d_row[o]);
}

// x[o] = d_col[o] y[e] for the original column o = perm[e]
template <class T>
__global__ void k_lu_sfinish(int32_t n, int32_t nrhs, const int32_t* __restrict__ perm,
                             const double* __restrict__ d_col, const T* __restrict__ y, T* __restrict__ x) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * nrhs) return;
  const int64_t g = e / nrhs, c = e - g * nrhs, o = perm[g];
  x[o * nrhs + c] = F<T>::scale(y[e], d_col[o]);
}

// Forward gather, 64 own rows of one supernode per unit (s, -, first row):
// y_s[p] = b[pivot row of p] minus the descendants' off-diagonal products
// that land on that row.
template <class T>
__global__ __launch_bounds__(LU_NT) void k_lu_fgather(LuSymDev S, LuFacDev<T> Fc, const LuUnit* __restrict__ units,
                                                      int32_t nrhs, const T* __restrict__ bh, const T* __restrict__ scr,
                                                      T* __restrict__ y) {
  const LuUnit un = units[blockIdx.x];
  const int s = un.s, w = S.sn_ncol[s], first = S.sn_first[s];
  const int64_t m = min(64, w - un.a);
  for (int64_t e = threadIdx.x; e < m * nrhs; e += LU_NT) {
    const int64_t p = un.a + e / nrhs, c = e % nrhs;
    const int64_t g = first + Fc.rperm[first + p];
    T v = bh[g * nrhs + c];
#pragma unroll 4
    for (int64_t q = S.gat_ptr[g]; q < S.gat_ptr[g + 1]; ++q) v = v - scr[S.gat_idx[q] * nrhs + c];
    y[(first + p) * nrhs + c] = v;
  }
}

// Backward U12 terms, 16 own rows per unit (s, -, first row), the trailing
// columns split over 16 slices summed in LDS, 8 right-hand sides at a time:
// y_s[p] -= U12[p, :] x[rs[w..]] (the ancestors' x, final, in x).
template <class T>
__global__ __launch_bounds__(LU_NT) void k_lu_boff(LuSymDev S, LuFacDev<T> Fc, const LuUnit* __restrict__ units,
                                                   int32_t nrhs, T* __restrict__ y, const T* __restrict__ x) {
  constexpr int CW = 8;
  __shared__ T part[16][16][CW + 1];
  const LuUnit un = units[blockIdx.x];
  const int s = un.s;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s], first = S.sn_first[s];
  const T* U = Fc.u + S.u_off[s];
  const int32_t* rs = S.rs_idx + S.rs_off[s] + w;
  const int64_t nt = h - w;
  const int i = threadIdx.x & 15, sl = threadIdx.x >> 4;
  const int64_t p = un.a + i;
  for (int c0 = 0; c0 < nrhs; c0 += CW) {
    const int cw = min(CW, nrhs - c0);
    T acc[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c) acc[c] = F<T>::zero();
    if (p < w) {
      // QB columns of the slice per pass, their loads issued together; the
      // next pass's row indices are loaded with them, so a pass waits on one
      // round of loads (ascending columns: the sums keep their order)
      constexpr int QB = 8;
      int64_t t = sl;
      int32_t rn[QB];
#pragma unroll
      for (int q = 0; q < QB; ++q) rn[q] = t + 16 * q < nt ? rs[t + 16 * q] : 0;
      for (; t + 16 * (QB - 1) < nt; t += 16 * QB) {
        T u[QB];
        const T* xr[QB];
#pragma unroll
        for (int q = 0; q < QB; ++q) {
          u[q] = U[p + (t + 16 * q) * w];
          xr[q] = x + (int64_t)rn[q] * nrhs + c0;
        }
        T xv[QB][CW];
#pragma unroll
        for (int q = 0; q < QB; ++q)
#pragma unroll
          for (int c = 0; c < CW; ++c) xv[q][c] = c < cw ? xr[q][c] : F<T>::zero();
#pragma unroll
        for (int q = 0; q < QB; ++q) {
          const int64_t tn = t + 16 * QB + 16 * q;
          rn[q] = tn < nt ? rs[tn] : 0;
        }
#pragma unroll
        for (int q = 0; q < QB; ++q)
#pragma unroll
          for (int c = 0; c < CW; ++c)
            if (c < cw) acc[c] = lu_fma(u[q], xv[q][c], acc[c]);
      }
      for (; t < nt; t += 16) {
        const T u1 = U[p + t * w];
        const T* x1 = x + (int64_t)rs[t] * nrhs + c0;
#pragma unroll
        for (int c = 0; c < CW; ++c)
          if (c < cw) acc[c] = lu_fma(u1, x1[c], acc[c]);
      }
    }
#pragma unroll
    for (int c = 0; c < CW; ++c) part[sl][i][c] = acc[c];
    __syncthreads();
    if (threadIdx.x < 16 * CW) {
      const int ii = threadIdx.x & 15, c = threadIdx.x >> 4;
      const int64_t pp = un.a + ii;
      if (pp < w && c < cw) {
        T sum = part[0][ii][c];
        for (int q = 1; q < 16; ++q) sum = sum + part[q][ii][c];
        T* d = y + (first + pp) * nrhs + c0 + c;
        *d = *d - sum;
      }
    }
    __syncthreads();
  }
}

// One forward block step of a supernode's own block, units (s, j0, first
// row, writer): every unit solves the diagonal block [j0, j0 + m) of L11 from
// the updated, not yet solved values in y (unit-lower substitution, one wave
// per right-hand side, lane per row, the block's L in LDS), then updates its
// 64 rows below the block with it: own rows in y, off-diagonal rows into the
// supernode's scratch rows (written by the first block, added to after). The
// writer unit stores the solved block into yo; no unit of the launch writes
// rows another one reads.
template <class T, int BS>
__global__ __launch_bounds__(LU_NT) void k_lu_fstep(LuSymDev S, LuFacDev<T> Fc, const LuUnit* __restrict__ units,
                                                    int32_t nrhs, T* __restrict__ y, T* __restrict__ yo,
                                                    T* __restrict__ scr) {
  constexpr int CW = 16;
  __shared__ T Lb[BS][BS + 1];
  __shared__ T yb[BS][CW + 1];
  const LuUnit un = units[blockIdx.x];
  const int s = un.s, j0 = un.kb;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s], first = S.sn_first[s], m = min(BS, w - j0);
  const T* L = Fc.l + S.l_off[s];
  {
    // the block's loads issued together, then their LDS stores
    constexpr int PER = BS * BS / LU_NT;
    T v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + u * LU_NT, r = e % BS, q = e / BS;
      v[u] = (r < m && q < r) ? L[(int64_t)(j0 + q) * h + j0 + r] : F<T>::zero();
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + u * LU_NT;
      Lb[e % BS][e / BS] = v[u];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t r0 = un.a;
  for (int c0 = 0; c0 < nrhs; c0 += CW) {
    const int cw = min(CW, nrhs - c0);
    // a wave's right-hand sides (columns wv, wv + 4, ...) as interleaved chains
    constexpr int CPW = CW / (LU_NT / 64);
    T v[CPW];
    const T* yc = y + (int64_t)(first + j0 + lane) * nrhs + c0;
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
      const int c = wv + k * (LU_NT / 64);
      v[k] = (c < cw && lane < m) ? yc[c] : F<T>::zero();
    }
    for (int q = 0; q < m; ++q) {
      const T lq = Lb[lane][q];
#pragma unroll
      for (int k = 0; k < CPW; ++k) {
        const T yq = lu_bcast(v[k], q);
        if (lane > q && lane < m) v[k] = lu_fma(F<T>::zero() - lq, yq, v[k]);
      }
    }
    if (lane < m)
#pragma unroll
      for (int k = 0; k < CPW; ++k) {
        const int c = wv + k * (LU_NT / 64);
        if (c < cw) {
          yb[lane][c] = v[k];
          if (un.b) yo[(int64_t)(first + j0 + lane) * nrhs + c0 + c] = v[k];
        }
      }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * cw; e += LU_NT) {
      const int64_t r = r0 + (e & 63);
      const int c = e >> 6;
      if (r >= h) continue;
      T acc = F<T>::zero();
#pragma unroll 8
      for (int q = 0; q < m; ++q) acc = lu_fma(L[(int64_t)(j0 + q) * h + r], yb[q][c], acc);
      if (r < w) {
        T* d = y + (first + r) * nrhs + c0 + c;
        *d = *d - acc;
      } else {
        T* d = scr + (S.s_off[s] + r - w) * nrhs + c0 + c;
        *d = j0 == 0 ? acc : *d + acc;
      }
    }
    __syncthreads();
  }
}

// One backward block step, units (s, j0, first row, writer): every unit
// solves the diagonal block of U11 (back substitution with the pivots) from
// the values in y, then updates its 64 own rows above the block; the writer
// stores the solved block into xo.
template <class T, int BS>
__global__ __launch_bounds__(LU_NT) void k_lu_bstep(LuSymDev S, LuFacDev<T> Fc, const LuUnit* __restrict__ units,
                                                    int32_t nrhs, T* __restrict__ y, T* __restrict__ xo) {
  constexpr int CW = 16;
  __shared__ T Ub[BS][BS + 1];
  __shared__ T yb[BS][CW + 1];
  const LuUnit un = units[blockIdx.x];
  const int s = un.s, j0 = un.kb;
  const int64_t h = S.sn_nrow[s];
  const int w = S.sn_ncol[s], first = S.sn_first[s], m = min(BS, w - j0);
  const T* L = Fc.l + S.l_off[s];
  {
    // the block's loads issued together, then their LDS stores
    constexpr int PER = BS * BS / LU_NT;
    T v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + u * LU_NT, r = e % BS, q = e / BS;
      v[u] = (q < m && r <= q) ? L[(int64_t)(j0 + q) * h + j0 + r] : F<T>::zero();
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + u * LU_NT;
      Ub[e % BS][e / BS] = v[u];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t r0 = un.a;
  for (int c0 = 0; c0 < nrhs; c0 += CW) {
    const int cw = min(CW, nrhs - c0);
    constexpr int CPW = CW / (LU_NT / 64);
    T v[CPW];
    const T* yc = y + (int64_t)(first + j0 + lane) * nrhs + c0;
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
      const int c = wv + k * (LU_NT / 64);
      v[k] = (c < cw && lane < m) ? yc[c] : F<T>::zero();
    }
    for (int q = m - 1; q >= 0; --q) {
      const T uq = Ub[lane][q];
      const T dinv = F<T>::recip(Ub[q][q]);
#pragma unroll
      for (int k = 0; k < CPW; ++k) {
        if (lane == q) v[k] = v[k] * dinv;
        const T xq = lu_bcast(v[k], q);
        if (lane < q) v[k] = lu_fma(F<T>::zero() - uq, xq, v[k]);
      }
    }
    if (lane < m)
#pragma unroll
      for (int k = 0; k < CPW; ++k) {
        const int c = wv + k * (LU_NT / 64);
        if (c < cw) {
          yb[lane][c] = v[k];
          if (un.b) xo[(int64_t)(first + j0 + lane) * nrhs + c0 + c] = v[k];
        }
      }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * cw; e += LU_NT) {
      const int64_t r = r0 + (e & 63);
      const int c = e >> 6;
      if (r >= j0) continue;
      T acc = F<T>::zero();
#pragma unroll 8
      for (int q = 0; q < m; ++q) acc = lu_fma(L[(int64_t)(j0 + q) * h + r], yb[q][c], acc);
      T* d = y + (first + r) * nrhs + c0 + c;
      *d = *d - acc;
    }
    __syncthreads();
  }
}

// ---- launchers
template <class T>
void launch_lu_assemble(const LuSymDev& S, const LuFacDev<T>& Fc, const T* vals, hipStream_t st) {
  if (S.nnz <= 0) return;
  hipLaunchKernelGGL(k_lu_assemble<T>, dim3((unsigned)((S.nnz + LU_NT - 1) / LU_NT)), dim3(LU_NT), 0, st, S, Fc, vals);
}
template <class T>
void launch_lu_cmod(const LuSymDev& S, const LuFacDev<T>& Fc, const LuItem* items, int nitems, const CTask* tasks,
                    hipStream_t st) {
  if (nitems <= 0) return;
  if constexpr (sizeof(T) == 8)
    hipLaunchKernelGGL(k_lu_cmod_mfma, dim3((unsigned)nitems), dim3(LU_NT), 0, st, S, Fc, items, tasks);
  else
    hipLaunchKernelGGL(k_lu_cmod_mfma_c64, dim3((unsigned)nitems), dim3(LU_NT), 0, st, S, Fc, items, tasks);
}
template <class T>
void launch_lu_getf2(const LuSymDev& S, const LuFacDev<T>& Fc, const LuUnit* units, int n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lu_getf2<T>, dim3((unsigned)n), dim3(LU_GT), 0, st, S, Fc, units);
}
template <class T>
void launch_lu_apply(const LuSymDev& S, const LuFacDev<T>& Fc, const LuUnit* units, int n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lu_apply<T>, dim3((unsigned)n), dim3(LU_NT), 0, st, S, Fc, units);
}
template <class T>
void launch_lu_gemm(const LuSymDev& S, const LuFacDev<T>& Fc, const LuUnit* units, int n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lu_gemm<T>, dim3((unsigned)n), dim3(LU_NT), 0, st, S, Fc, units);
}
template <class T>
void launch_lu_sprep(int32_t n, int32_t nrhs, const int32_t* perm, const double* d_row, const T* b, T* bh,
                     hipStream_t st) {
  const int64_t m = (int64_t)n * nrhs;
  if (m <= 0) return;
  hipLaunchKernelGGL(k_lu_sprep<T>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, n, nrhs, perm, d_row, b, bh);
}
template <class T>
void launch_lu_sfinish(int32_t n, int32_t nrhs, const int32_t* perm, const double* d_col, const T* y, T* x,
                       hipStream_t st) {
  const int64_t m = (int64_t)n * nrhs;
  if (m <= 0) return;
  hipLaunchKernelGGL(k_lu_sfinish<T>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, n, nrhs, perm, d_col, y, x);
}
template <class T>
int lu_solve_bs() {
  return sizeof(T) == 16 ? 32 : 64;
}
// kind: 0 gather (into y), 1 forward block step (y -> yo), 3 backward U12
// terms (on y, x of the ancestors in yo), 4 backward block step (y -> yo)
template <class T>
void launch_lu_sstep(int kind, const LuSymDev& S, const LuFacDev<T>& Fc, const LuUnit* units, int n, int32_t nrhs,
                     const T* bh, T* scr, T* y, T* yo, hipStream_t st) {
  if (n <= 0) return;
  constexpr int BS = sizeof(T) == 16 ? 32 : 64;
  const dim3 g((unsigned)n), b(LU_NT);
  switch (kind) {
    case 0: hipLaunchKernelGGL(k_lu_fgather<T>, g, b, 0, st, S, Fc, units, nrhs, bh, scr, y); break;
    case 1: hipLaunchKernelGGL((k_lu_fstep<T, BS>), g, b, 0, st, S, Fc, units, nrhs, y, yo, scr); break;
    case 3: hipLaunchKernelGGL(k_lu_boff<T>, g, b, 0, st, S, Fc, units, nrhs, y, yo); break;
    default: hipLaunchKernelGGL((k_lu_bstep<T, BS>), g, b, 0, st, S, Fc, units, nrhs, y, yo); break;
  }
}

#define RSL_LU_INST(T)                                                                                          \
  template void launch_lu_assemble<T>(const LuSymDev&, const LuFacDev<T>&, const T*, hipStream_t);              \
  template void launch_lu_cmod<T>(const LuSymDev&, const LuFacDev<T>&, const LuItem*, int, const CTask*,         \
                                  hipStream_t);                                                                 \
  template void launch_lu_getf2<T>(const LuSymDev&, const LuFacDev<T>&, const LuUnit*, int, hipStream_t);       \
  template void launch_lu_apply<T>(const LuSymDev&, const LuFacDev<T>&, const LuUnit*, int, hipStream_t);       \
  template void launch_lu_gemm<T>(const LuSymDev&, const LuFacDev<T>&, const LuUnit*, int, hipStream_t);        \
  template void launch_lu_sprep<T>(int32_t, int32_t, const int32_t*, const double*, const T*, T*, hipStream_t);  \
  template void launch_lu_sfinish<T>(int32_t, int32_t, const int32_t*, const double*, const T*, T*, hipStream_t); \
  template int lu_solve_bs<T>();                                                                              \
  template void launch_lu_sstep<T>(int, const LuSymDev&, const LuFacDev<T>&, const LuUnit*, int, int32_t, const T*, \
                                   T*, T*, T*, hipStream_t);
RSL_LU_INST(double)
RSL_LU_INST(cplx)

}  // namespace rsl
