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
template <class T> void launch_lu_fwd_own(const LuSymDev&, const LuFacDev<T>&, const int32_t*, int, int32_t, const T*,
                                          const T*, T*, hipStream_t);
template <class T> void launch_lu_fwd_off(const LuSymDev&, const LuFacDev<T>&, const int2*, int, int32_t, const T*, T*,
                                          hipStream_t);
template <class T> void launch_lu_bwd_off(const LuSymDev&, const LuFacDev<T>&, const int2*, int, int32_t, T*, hipStream_t);
template <class T> void launch_lu_bwd_own(const LuSymDev&, const LuFacDev<T>&, const int32_t*, int, int32_t, T*,
                                          hipStream_t);

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
  sunits.upload(P.sunits, st);
  rperm_init.upload(P.rperm_init, st);
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
  DBuf l, u, rperm, ipiv, status, counters, vals, drow, dcol, bh, yv, scr, bx;
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
    const int32_t* lists = sym->level_lists.as<int32_t>();
    const int2* su = sym->sunits.as<int2>();
    launch_lu_sprep<T>(P.n, nrhs, S.perm, drow.as<double>(), b, bhp, st);
    for (size_t l = 0; l < P.levels.size(); ++l) {
      const LuLevelPlan& L = P.levels[l];
      launch_lu_fwd_own<T>(S, fv, lists + P.level_off[l], (int)L.all.size(), nrhs, bhp, sc, y, st);
      launch_lu_fwd_off<T>(S, fv, su + L.fo_beg, (int)(L.fo_end - L.fo_beg), nrhs, y, sc, st);
    }
    for (size_t l = P.levels.size(); l-- > 0;) {
      const LuLevelPlan& L = P.levels[l];
      launch_lu_bwd_off<T>(S, fv, su + L.bo_beg, (int)(L.bo_end - L.bo_beg), nrhs, y, st);
      launch_lu_bwd_own<T>(S, fv, lists + P.level_off[l], (int)L.all.size(), nrhs, y, st);
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
