// Host runtime of the LU twin (see lu_runtime.hpp).
#include "lu_runtime.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace rsl {

template <class T> void launch_fill(T*, int64_t, T, hipStream_t);
template <class T> void launch_lu_assemble(const LuSymDev&, const LuFacDev<T>&, const T*, hipStream_t);
template <class T> void launch_lu_cmod(const LuSymDev&, const LuFacDev<T>&, const LuItem*, int, const CTask*, hipStream_t);
template <class T> void launch_lu_getf2(const LuSymDev&, const LuFacDev<T>&, const LuUnit*, int, hipStream_t);
template <class T> void launch_lu_apply(const LuSymDev&, const LuFacDev<T>&, const LuUnit*, int, hipStream_t);
template <class T> void launch_lu_gemm(const LuSymDev&, const LuFacDev<T>&, const LuUnit*, int, hipStream_t);
template <class T> void launch_lu_sprep(int32_t, int32_t, const int32_t*, const double*, const T*, T*, hipStream_t);
template <class T> void launch_lu_sfinish(int32_t, int32_t, const int32_t*, const double*, const T*, T*, hipStream_t);
template <class T> int lu_solve_bs();
template <class T> void launch_lu_sstep(int, const LuSymDev&, const LuFacDev<T>&, const LuUnit*, int, int32_t, const T*,
                                        T*, T*, T*, hipStream_t);
template <class T> void launch_expand(int64_t, const int64_t*, const T*, T*, hipStream_t);
template <class T> void launch_residual(int32_t, int32_t, const int64_t*, const int32_t*, const T*, const T*, const T*, T*,
                                        hipStream_t, LaunchProfiler*);
template <class T> void launch_colmax(int32_t, int32_t, const T*, unsigned long long*, hipStream_t);
template <class T> void launch_update_cols(int32_t, int32_t, const uint8_t*, T*, const T*, hipStream_t);
template <class T> void launch_copy_cols(int32_t, int32_t, const uint8_t*, T*, const T*, hipStream_t);

LuDeviceSymbolic::~LuDeviceSymbolic() {
  if (stream) (void)hipStreamDestroy(stream);
}

void LuDeviceSymbolic::upload() {
  if (uploaded) return;
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  const LuPlan& P = plan;
  const hipStream_t st = stream;
  sn_first.upload(P.sn_first, st);
  sn_ncol.upload(P.sn_ncol, st);
  sn_nrow.upload(P.sn_nrow, st);
  nb.upload(P.nb, st);
  rs_off.upload(P.rs_off, st);
  rs_idx.upload(P.rs_idx, st);
  l_off.upload(P.l_off, st);
  u_off.upload(P.u_off, st);
  s_off.upload(P.s_off, st);
  gat_ptr.upload(P.gat_ptr, st);
  gat_idx.upload(P.gat_idx, st);
  std::vector<int32_t> pm(P.sym.perm.begin(), P.sym.perm.end());
  perm.upload(pm, st);
  amap.upload(P.amap, st);
  level_lists.upload(P.level_lists, st);
  items.upload(P.items, st);
  tasks.upload(P.tasks, st);
  units.upload(P.units, st);
  rperm_init.upload(P.rperm_init, st);
  csr_ptr.upload(P.csr_ptr, st);
  csr_col.upload(P.csr_col, st);
  csr_src.upload(P.csr_src, st);
  hip_check(hipStreamSynchronize(st), "upload");
  view.n = P.n;
  view.ns = P.ns;
  view.sn_first = sn_first.as<int32_t>();
  view.sn_ncol = sn_ncol.as<int32_t>();
  view.sn_nrow = sn_nrow.as<int32_t>();
  view.nb = nb.as<int32_t>();
  view.rs_off = rs_off.as<int64_t>();
  view.rs_idx = rs_idx.as<int32_t>();
  view.l_off = l_off.as<int64_t>();
  view.u_off = u_off.as<int64_t>();
  view.s_off = s_off.as<int64_t>();
  view.gat_ptr = gat_ptr.as<int64_t>();
  view.gat_idx = gat_idx.as<int64_t>();
  view.perm = perm.as<int32_t>();
  view.amap = amap.as<int64_t>();
  view.nnz = P.nnz;
  uploaded = true;
}

namespace {

template <class T>
struct LuField;
template <>
struct LuField<double> {
  static constexpr int id = 0;
  static double mag(const double* v, int64_t k) { return std::fabs(v[k]); }
};
template <>
struct LuField<cplx> {
  static constexpr int id = 1;
  static double mag(const double* v, int64_t k) { return std::hypot(v[2 * k], v[2 * k + 1]); }
};

// inv_sqrt_scale_guarded (src/scaling/mod.rs:85-96)
double inv_sqrt_guarded(double m) {
  if (m > 0.0) {
    const double c = 1.0 / std::sqrt(m);
    if (std::isfinite(c) && c > 0.0) return c;
  }
  return 1.0;
}

template <class T>
class LuDeviceFactor final : public LuFactorBase {
 public:
  DBuf l, u, rperm, ipiv, status, counters, vals, drow, dcol, bh, yv, y2, scr, bx;
  DBuf r_a, r_acsr, r_b, r_x, r_best, r_r, r_dx, r_max, r_mask;
  // solve work units (their block width depends on the field): per level the
  // gather, the forward block steps (diagonal, update), the U12 units and the
  // backward block steps
  struct SolveLevel {
    int64_t g0 = 0, g1 = 0, o0 = 0, o1 = 0;
    std::vector<int64_t> fs, bs;  // per block step: units begin, end
  };
  std::vector<SolveLevel> slev;
  DBuf su;
  std::vector<T> h_vals;
  LuFacDev<T> fv{};
  hipEvent_t ev_caller = nullptr, ev_done = nullptr;
  bool emit_ready = false;
  LuEmitted em;

  LuDeviceFactor(std::shared_ptr<LuDeviceSymbolic> s, const LuFactorSettings& fs) {
    sym = std::move(s);
    settings = fs;
    sym->upload();
    const LuPlan& P = sym->plan;
    l.alloc(sizeof(T) * std::max<int64_t>(P.l_total, 1));
    u.alloc(sizeof(T) * std::max<int64_t>(P.u_total, 1));
    rperm.alloc(sizeof(int32_t) * std::max<int32_t>(P.n, 1));
    ipiv.alloc(sizeof(int32_t) * std::max<int32_t>(P.n, 1));
    status.alloc(sizeof(int32_t) * 2);
    counters.alloc(sizeof(unsigned long long) * 2);
    vals.alloc(sizeof(T) * std::max<int64_t>(P.nnz, 1));
    drow.alloc(sizeof(double) * std::max<int32_t>(P.n, 1));
    dcol.alloc(sizeof(double) * std::max<int32_t>(P.n, 1));
    fv.l = l.as<T>();
    fv.u = u.as<T>();
    fv.rperm = rperm.as<int32_t>();
    fv.ipiv = ipiv.as<int32_t>();
    fv.status = status.as<int32_t>();
    fv.counters = counters.as<unsigned long long>();
    const double pu = std::min(1.0, std::max(0.0, fs.pivot_u));
    fv.thresh_sq = pu * pu;
    hip_check(hipEventCreateWithFlags(&ev_caller, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&ev_done, hipEventDisableTiming), "hipEventCreate");
    plan_solve();
  }

  void plan_solve() {
    const LuPlan& P = sym->plan;
    const int BS = lu_solve_bs<T>();
    std::vector<LuUnit> u;
    slev.assign(P.levels.size(), SolveLevel());
    for (size_t l = 0; l < P.levels.size(); ++l) {
      const LuLevelPlan& L = P.levels[l];
      SolveLevel& sl = slev[l];
      sl.g0 = (int64_t)u.size();
      for (int32_t s : L.all)
        for (int32_t p0 = 0; p0 < P.sn_ncol[s]; p0 += 64) u.push_back({s, 0, p0, 0});
      sl.g1 = (int64_t)u.size();
      int nblk = 0;
      for (int32_t s : L.all) nblk = std::max(nblk, (P.sn_ncol[s] + BS - 1) / BS);
      for (int i = 0; i < nblk; ++i) {
        sl.fs.push_back((int64_t)u.size());
        for (int32_t s : L.all) {
          const int32_t w = P.sn_ncol[s], h = P.sn_nrow[s], j0 = i * BS;
          if (j0 >= w) continue;
          const int32_t below = j0 + std::min(BS, w - j0);
          if (below >= h) u.push_back({s, j0, h, 1});
          for (int32_t r0 = below; r0 < h; r0 += 64) u.push_back({s, j0, r0, r0 == below ? 1 : 0});
        }
        sl.fs.push_back((int64_t)u.size());
      }
      sl.o0 = (int64_t)u.size();
      for (int32_t s : L.all)
        if (P.sn_nrow[s] > P.sn_ncol[s])
          for (int32_t p0 = 0; p0 < P.sn_ncol[s]; p0 += 16) u.push_back({s, 0, p0, 0});
      sl.o1 = (int64_t)u.size();
      for (int i = 0; i < nblk; ++i) {
        sl.bs.push_back((int64_t)u.size());
        for (int32_t s : L.all) {
          const int32_t nb_s = (P.sn_ncol[s] + BS - 1) / BS;
          if (i >= nb_s) continue;
          const int32_t j0 = (nb_s - 1 - i) * BS;
          if (j0 == 0) u.push_back({s, 0, 0, 1});
          for (int32_t r0 = 0; r0 < j0; r0 += 64) u.push_back({s, j0, r0, r0 == 0 ? 1 : 0});
        }
        sl.bs.push_back((int64_t)u.size());
      }
    }
    if (u.size() > (size_t)INT32_MAX) throw Error(INVALID_INPUT, "lu solve: too many work units");
    su.upload(u, sym->stream);
    hip_check(hipStreamSynchronize(sym->stream), "upload");
  }
  ~LuDeviceFactor() override {
    if (ev_caller) (void)hipEventDestroy(ev_caller);
    if (ev_done) (void)hipEventDestroy(ev_done);
  }

  int field() const override { return LuField<T>::id; }

  // factor_general_lu_numeric, left-looking path (multifrontal_lu.rs:2673-2800)
  void refactor(const double* values) override {
    const LuPlan& P = sym->plan;
    const LuSymDev& S = sym->view;
    const hipStream_t st = sym->stream;
    const auto t0 = std::chrono::steady_clock::now();
    const int32_t n = P.n;
    // ZeroPivotAction floors (:2716-2723): ForceAccept over the unscaled values
    PivotPolicy pol{0, 0.0};
    if (settings.pivot == 1) {
      pol.have_floor = 1;
      pol.floor = std::max(settings.abs_floor, 0.0);
    } else if (settings.pivot == 2) {
      double amax = 0.0;
      for (int64_t k = 0; k < P.nnz; ++k) amax = std::max(amax, LuField<T>::mag(values, k));
      pol.have_floor = 1;
      pol.floor = std::max(amax, 1.0) * 2.220446049250313e-16;
    }
    fv.pol = pol;
    // two-sided equilibration from the original matrix (:2733-2758)
    std::vector<double> rmax(n, 0.0), cmax(n, 0.0);
    for (int32_t j = 0; j < n; ++j)
      for (int64_t k = P.colp[j]; k < P.colp[j + 1]; ++k) {
        const double m = LuField<T>::mag(values, k);
        const int32_t i = P.rowi[k];
        rmax[i] = std::max(rmax[i], m);
        cmax[j] = std::max(cmax[j], m);
      }
    d_row.resize(n);
    d_col.resize(n);
    for (int32_t i = 0; i < n; ++i) {
      d_row[i] = inv_sqrt_guarded(rmax[i]);
      d_col[i] = inv_sqrt_guarded(cmax[i]);
    }
    h_vals.resize(P.nnz);
    const T* v = reinterpret_cast<const T*>(values);
    for (int32_t j = 0; j < n; ++j)
      for (int64_t k = P.colp[j]; k < P.colp[j + 1]; ++k)
        h_vals[k] = F<T>::scale(v[k], d_row[P.rowi[k]] * d_col[j]);
    if (P.nnz) hip_check(hipMemcpyAsync(vals.p, h_vals.data(), sizeof(T) * P.nnz, hipMemcpyHostToDevice, st), "values");
    if (n) {
      hip_check(hipMemcpyAsync(drow.p, d_row.data(), sizeof(double) * n, hipMemcpyHostToDevice, st), "d_row");
      hip_check(hipMemcpyAsync(dcol.p, d_col.data(), sizeof(double) * n, hipMemcpyHostToDevice, st), "d_col");
      hip_check(hipMemcpyAsync(rperm.p, sym->rperm_init.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st), "rperm");
    }
    static const int32_t st0[2] = {0, 0};
    static const unsigned long long c0[2] = {0ull, ~0ull};
    hip_check(hipMemcpyAsync(status.p, st0, sizeof(st0), hipMemcpyHostToDevice, st), "status");
    hip_check(hipMemcpyAsync(counters.p, c0, sizeof(c0), hipMemcpyHostToDevice, st), "counters");
    launch_fill<T>(fv.l, P.l_total, F<T>::zero(), st);
    launch_fill<T>(fv.u, P.u_total, F<T>::zero(), st);
    launch_lu_assemble<T>(S, fv, vals.as<T>(), st);
    const LuItem* items = sym->items.as<LuItem>();
    const CTask* tasks = sym->tasks.as<CTask>();
    const LuUnit* units = sym->units.as<LuUnit>();
    for (const LuLevelPlan& L : P.levels) {
      launch_lu_cmod<T>(S, fv, items + L.item_beg, (int)(L.item_end - L.item_beg), tasks, st);
      for (const LuStepPlan& sp : L.steps) {
        launch_lu_getf2<T>(S, fv, units + sp.g_beg, (int)(sp.g_end - sp.g_beg), st);
        launch_lu_apply<T>(S, fv, units + sp.a_beg, (int)(sp.a_end - sp.a_beg), st);
        launch_lu_gemm<T>(S, fv, units + sp.m_beg, (int)(sp.m_end - sp.m_beg), st);
      }
    }
    hip_check(hipGetLastError(), "lu factor launch");
    int32_t stat[2];
    unsigned long long cnt[2];
    hip_check(hipMemcpyAsync(stat, status.p, sizeof(stat), hipMemcpyDeviceToHost, st), "status");
    hip_check(hipMemcpyAsync(cnt, counters.p, sizeof(cnt), hipMemcpyDeviceToHost, st), "counters");
    hip_check(hipStreamSynchronize(st), "lu factor");
    t_factor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    emit_ready = false;
    if (stat[0] != 0) {
      std::string msg = "numerically rank deficient";
      if (cnt[1] != ~0ull) msg += ": zero pivot at elimination position " + std::to_string(cnt[1]);
      throw Error(NUMERICALLY_RANK_DEFICIENT, msg);
    }
    n_perturbed = (int64_t)cnt[0];
  }

  void ensure_solve(int64_t nrhs) {
    const LuPlan& P = sym->plan;
    const size_t need = sizeof(T) * (size_t)std::max<int64_t>((int64_t)P.n * nrhs, 1);
    if (bh.bytes < need) bh.alloc(need);
    if (yv.bytes < need) yv.alloc(need);
    if (y2.bytes < need) y2.alloc(need);
    const size_t sneed = sizeof(T) * (size_t)std::max<int64_t>(P.s_total * nrhs, 1);
    if (scr.bytes < sneed) scr.alloc(sneed);
  }

  // the supernodal forward / backward sweeps (solve_lu_block's arithmetic,
  // :3209-3283, grouped by supernode) on the library stream
  void enqueue_solve(const T* b, int32_t nrhs, T* x) {
    const LuPlan& P = sym->plan;
    const LuSymDev& S = sym->view;
    const hipStream_t st = sym->stream;
    T* bhp = bh.as<T>();
    T* y = yv.as<T>();
    T* sc = scr.as<T>();
    const LuUnit* u = su.as<LuUnit>();
    launch_lu_sprep<T>(P.n, nrhs, S.perm, drow.as<double>(), b, bhp, st);
    // forward: gather into y, block steps solve from y into y2; backward:
    // the U12 terms and block steps work on y2 and store x into y
    T* y2p = y2.as<T>();
    auto step = [&](int kind, int64_t a, int64_t e, T* yin, T* yout) {
      launch_lu_sstep<T>(kind, S, fv, u + a, (int)(e - a), nrhs, bhp, sc, yin, yout, st);
    };
    for (size_t l = 0; l < slev.size(); ++l) {
      const SolveLevel& sl = slev[l];
      step(0, sl.g0, sl.g1, y, y2p);
      for (size_t i = 0; i < sl.fs.size(); i += 2) step(1, sl.fs[i], sl.fs[i + 1], y, y2p);
    }
    for (size_t l = slev.size(); l-- > 0;) {
      const SolveLevel& sl = slev[l];
      step(3, sl.o0, sl.o1, y2p, y);
      for (size_t i = 0; i < sl.bs.size(); i += 2) step(4, sl.bs[i], sl.bs[i + 1], y2p, y);
    }
    launch_lu_sfinish<T>(P.n, nrhs, S.perm, dcol.as<double>(), y, x, st);
    hip_check(hipGetLastError(), "lu solve launch");
  }

  void solve_host(const double* b, int64_t nrhs, double* x) override {
    const LuPlan& P = sym->plan;
    if (nrhs <= 0 || P.n == 0) return;
    if (nrhs > INT32_MAX / std::max<int32_t>(P.n, 1)) throw Error(INVALID_INPUT, "too many right-hand sides");
    ensure_solve(nrhs);
    const size_t bytes = sizeof(T) * (size_t)P.n * nrhs;
    if (bx.bytes < bytes) bx.alloc(bytes);
    const hipStream_t st = sym->stream;
    hip_check(hipMemcpyAsync(bx.p, b, bytes, hipMemcpyHostToDevice, st), "b");
    enqueue_solve(bx.as<T>(), (int32_t)nrhs, bx.as<T>());
    hip_check(hipMemcpyAsync(x, bx.p, bytes, hipMemcpyDeviceToHost, st), "x");
    hip_check(hipStreamSynchronize(st), "lu solve");
  }

  void solve_device(const void* d_b, int64_t nrhs, void* d_x, hipStream_t caller) override {
    const LuPlan& P = sym->plan;
    if (nrhs <= 0 || P.n == 0) return;
    if (nrhs > INT32_MAX / std::max<int32_t>(P.n, 1)) throw Error(INVALID_INPUT, "too many right-hand sides");
    ensure_solve(nrhs);
    hip_check(hipEventRecord(ev_caller, caller), "hipEventRecord");
    hip_check(hipStreamWaitEvent(sym->stream, ev_caller, 0), "hipStreamWaitEvent");
    enqueue_solve(static_cast<const T*>(d_b), (int32_t)nrhs, static_cast<T*>(d_x));
    hip_check(hipEventRecord(ev_done, sym->stream), "hipEventRecord");
    hip_check(hipStreamWaitEvent(caller, ev_done, 0), "hipStreamWaitEvent");
  }

  void refine(const double* a_values, const double* b, int64_t nrhs, int max_iter, double* x,
              double* info) override {
    const LuPlan& P = sym->plan;
    const int32_t n = P.n;
    if (nrhs <= 0 || n == 0) return;
    if (nrhs > 4096 || nrhs > INT32_MAX / n) throw Error(INVALID_INPUT, "too many right-hand sides");
    const int32_t k = (int32_t)nrhs;
    ensure_solve(nrhs);
    const hipStream_t st = sym->stream;
    const size_t vb = sizeof(T) * (size_t)n * k, ab = sizeof(T) * (size_t)std::max<int64_t>(P.nnz, 1);
    if (r_a.bytes < ab) r_a.alloc(ab);
    if (r_acsr.bytes < ab) r_acsr.alloc(ab);
    for (DBuf* d : {&r_b, &r_x, &r_best, &r_r, &r_dx})
      if (d->bytes < vb) d->alloc(vb);
    if (r_max.bytes < sizeof(unsigned long long) * k) r_max.alloc(sizeof(unsigned long long) * k);
    if (r_mask.bytes < (size_t)2 * k) r_mask.alloc((size_t)2 * k);
    if (P.nnz) hip_check(hipMemcpyAsync(r_a.p, a_values, sizeof(T) * P.nnz, hipMemcpyHostToDevice, st), "A");
    launch_expand<T>(P.nnz, sym->csr_src.as<int64_t>(), r_a.as<T>(), r_acsr.as<T>(), st);
    hip_check(hipMemcpyAsync(r_b.p, b, vb, hipMemcpyHostToDevice, st), "b");
    T* xd = r_x.as<T>();
    enqueue_solve(r_b.as<T>(), k, xd);
    std::vector<double> best(k, INFINITY);
    std::vector<int32_t> iters(k, 0);
    std::vector<uint8_t> keep(k), active(k, 1);
    std::vector<unsigned long long> mx(k);
    uint8_t* keep_d = r_mask.as<uint8_t>();
    uint8_t* act_d = keep_d + k;
    for (int it = 0; it <= max_iter; ++it) {
      launch_residual<T>(n, k, sym->csr_ptr.as<int64_t>(), sym->csr_col.as<int32_t>(), r_acsr.as<T>(), xd,
                         r_b.as<T>(), r_r.as<T>(), st, nullptr);
      hip_check(hipMemsetAsync(r_max.p, 0, sizeof(unsigned long long) * k, st), "memset");
      launch_colmax<T>(n, k, r_r.as<T>(), r_max.as<unsigned long long>(), st);
      hip_check(hipMemcpyAsync(mx.data(), r_max.p, sizeof(unsigned long long) * k, hipMemcpyDeviceToHost, st), "max");
      hip_check(hipStreamSynchronize(st), "refine");
      bool any = false;
      for (int c = 0; c < k; ++c) {
        keep[c] = 0;
        if (!active[c]) continue;
        double res;
        std::memcpy(&res, &mx[c], sizeof(res));
        if (res < best[c]) {
          best[c] = res;
          keep[c] = 1;
        }
        if (res == 0.0 || it == max_iter) active[c] = 0;
        any = any || active[c];
      }
      hip_check(hipMemcpyAsync(keep_d, keep.data(), k, hipMemcpyHostToDevice, st), "mask");
      hip_check(hipMemcpyAsync(act_d, active.data(), k, hipMemcpyHostToDevice, st), "mask");
      launch_copy_cols<T>(n, k, keep_d, r_best.as<T>(), xd, st);
      if (!any) break;
      enqueue_solve(r_r.as<T>(), k, r_dx.as<T>());
      launch_update_cols<T>(n, k, act_d, xd, r_dx.as<T>(), st);
      for (int c = 0; c < k; ++c) iters[c] += active[c];
      hip_check(hipStreamSynchronize(st), "refine");  // host masks are rewritten next step
    }
    hip_check(hipMemcpyAsync(x, r_best.p, vb, hipMemcpyDeviceToHost, st), "x");
    hip_check(hipStreamSynchronize(st), "refine");
    if (info)
      for (int c = 0; c < k; ++c) {
        info[2 * c] = iters[c];
        info[2 * c + 1] = best[c];
      }
  }

  // emit_and_free (multifrontal_lu.rs:1525-1618) over the downloaded panels
  const LuEmitted& emitted() override {
    if (emit_ready) return em;
    const LuPlan& P = sym->plan;
    const hipStream_t st = sym->stream;
    constexpr int ND = sizeof(T) / sizeof(double);
    std::vector<double> hl((size_t)P.l_total * ND), hu((size_t)P.u_total * ND);
    std::vector<int32_t> hr(P.n);
    if (P.l_total) hip_check(hipMemcpyAsync(hl.data(), l.p, sizeof(T) * P.l_total, hipMemcpyDeviceToHost, st), "L");
    if (P.u_total) hip_check(hipMemcpyAsync(hu.data(), u.p, sizeof(T) * P.u_total, hipMemcpyDeviceToHost, st), "U");
    if (P.n) hip_check(hipMemcpyAsync(hr.data(), rperm.p, sizeof(int32_t) * P.n, hipMemcpyDeviceToHost, st), "rperm");
    hip_check(hipStreamSynchronize(st), "lu download");
    const int32_t n = P.n;
    std::vector<int64_t> row_pos(n, 0);
    em.perm.assign(n, 0);
    em.perm_row.assign(n, 0);
    for (int32_t s = 0; s < P.ns; ++s)
      for (int32_t p = 0; p < P.sn_ncol[s]; ++p) {
        const int32_t e = P.sn_first[s] + p, g_row = P.sn_first[s] + hr[e];
        row_pos[g_row] = e;
        em.perm[e] = P.sym.perm[e];
        em.perm_row[e] = P.sym.perm[g_row];
      }
    auto mag = [&](const double* q) { return ND == 2 ? std::hypot(q[0], q[1]) : std::fabs(q[0]); };
    auto nonzero = [&](const double* q) { return ND == 2 ? (q[0] != 0.0 || q[1] != 0.0) : q[0] != 0.0; };
    struct Ent {
      int64_t r;
      const double* v;
    };
    std::vector<Ent> buf;
    const double one[2] = {1.0, 0.0};
    em.l_col_ptr.assign(n + 1, 0);
    em.u_row_ptr.assign(n + 1, 0);
    em.l_row_idx.clear();
    em.u_col_idx.clear();
    em.l_values.clear();
    em.u_values.clear();
    const double tau = settings.drop_tol;
    auto finish = [&](int64_t de, std::vector<int32_t>& idx, std::vector<double>& val) {
      if (tau >= 0.0) {
        double mx = 0.0;
        for (const Ent& x : buf)
          if (x.r != de) mx = std::max(mx, mag(x.v));
        const double thr = tau * mx;
        buf.erase(std::remove_if(buf.begin(), buf.end(), [&](const Ent& x) { return x.r != de && !(mag(x.v) >= thr); }),
                  buf.end());
      }
      std::sort(buf.begin(), buf.end(), [](const Ent& a, const Ent& b) { return a.r < b.r; });
      for (const Ent& x : buf) {
        idx.push_back((int32_t)x.r);
        for (int q = 0; q < ND; ++q) val.push_back(x.v[q]);
      }
    };
    for (int32_t s = 0; s < P.ns; ++s) {
      const int64_t h = P.sn_nrow[s], w = P.sn_ncol[s], first = P.sn_first[s];
      const int32_t* rs = P.rs_idx.data() + P.rs_off[s];
      const double* Ls = hl.data() + (size_t)P.l_off[s] * ND;
      const double* Us = hu.data() + (size_t)P.u_off[s] * ND;
      for (int64_t p = 0; p < w; ++p) {
        const int64_t de = first + p;
        buf.clear();
        buf.push_back({de, one});
        for (int64_t i = p + 1; i < h; ++i) {
          const double* q = Ls + (size_t)(p * h + i) * ND;
          if (!nonzero(q)) continue;
          const int32_t g = i < w ? (int32_t)(first + hr[first + i]) : rs[i];
          buf.push_back({row_pos[g], q});
        }
        finish(de, em.l_row_idx, em.l_values);
        em.l_col_ptr[de + 1] = (int64_t)em.l_row_idx.size();
        buf.clear();
        buf.push_back({de, Ls + (size_t)(p * h + p) * ND});
        for (int64_t j = p + 1; j < w; ++j) {
          const double* q = Ls + (size_t)(j * h + p) * ND;
          if (nonzero(q)) buf.push_back({first + j, q});
        }
        for (int64_t t = 0; t < h - w; ++t) {
          const double* q = Us + (size_t)(p + t * w) * ND;
          if (nonzero(q)) buf.push_back({rs[w + t], q});
        }
        finish(de, em.u_col_idx, em.u_values);
        em.u_row_ptr[de + 1] = (int64_t)em.u_col_idx.size();
      }
    }
    emit_ready = true;
    return em;
  }
};

}  // namespace

std::unique_ptr<LuFactorBase> make_lu_factor(std::shared_ptr<LuDeviceSymbolic> sym, int field,
                                             const double* values, const LuFactorSettings& fs) {
  std::unique_ptr<LuFactorBase> f;
  if (field == 0) f.reset(new LuDeviceFactor<double>(std::move(sym), fs));
  else f.reset(new LuDeviceFactor<cplx>(std::move(sym), fs));
  f->refactor(values);
  return f;
}

}  // namespace rsl
