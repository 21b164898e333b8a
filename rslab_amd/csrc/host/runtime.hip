// Host runtime driving the device factor and solve.
//
// LdltSymbolic::factor (src/numeric/sparse_solver.rs:677-711): equilibrate on
// the host (one-pass infinity norm), then the numeric factor runs level by
// level over the supernode tree: cmod tiles of the level, then cdiv of each of
// its supernodes. Solves follow LdltSolver::solve_many (:302-327).
#include "runtime.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace rsl {

// kernel launchers (factor_kernels.hip, solve_kernels.hip, refine_kernels.hip)
template <class T> void launch_fill(T*, int64_t, T, hipStream_t);
template <class T> void launch_assemble(const DevSym&, const DevFac<T>&, const T*, const double*, hipStream_t);
template <class T> void launch_cmod(const DevSym&, const DevFac<T>&, const CItem*, int, const CTask*, hipStream_t);
template <class T> void launch_cdiv_small(const DevSym&, const DevFac<T>&, const int32_t*, int, int, PivotPolicy, hipStream_t);
template <class T> void launch_cdiv_wide(const DevSym&, const DevFac<T>&, int, int, int64_t, int, PivotPolicy, hipStream_t);
template <class T> hipError_t configure_factor_kernels();
template <class T> void launch_gather_rhs(int32_t, int32_t, const int32_t*, const double*, const T*, T*, hipStream_t);
template <class T> void launch_scatter_sol(int32_t, int32_t, const int32_t*, const double*, const T*, T*, hipStream_t);
template <class T> void launch_fwd_tri(const DevSym&, const DevFac<T>&, const int32_t*, int, int32_t, T*, const T*, hipStream_t);
template <class T> void launch_fwd_update(const DevSym&, const DevFac<T>&, const int2*, int, int32_t, const T*, T*, hipStream_t);
template <class T> void launch_dsolve(int32_t, int32_t, const T*, const T*, const uint8_t*, T*, hipStream_t);
template <class T> void launch_bwd_update(const DevSym&, const DevFac<T>&, const int2*, int, int32_t, T*, hipStream_t);
template <class T> void launch_bwd_tri(const DevSym&, const DevFac<T>&, const int32_t*, int, int32_t, T*, hipStream_t);

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory) throw Error(OUT_OF_MEMORY, std::string("device: out of memory in ") + what);
  throw Error(DEVICE_ERROR, std::string("device: ") + what + ": " + hipGetErrorString(e));
}

DBuf::~DBuf() {
  if (p) (void)hipFree(p);
}

void DBuf::alloc(size_t nbytes) {
  if (p) {
    (void)hipFree(p);
    p = nullptr;
  }
  bytes = nbytes;
  if (nbytes) hip_check(hipMalloc(&p, nbytes), "hipMalloc");
}

void onepass_scale(int32_t n, int field, const int64_t* colp, const int32_t* rowi,
                   const double* val, std::vector<double>& scale) {
  scale.assign(n, 0.0);
  for (int32_t j = 0; j < n; ++j)
    for (int64_t k = colp[j]; k < colp[j + 1]; ++k) {
      const int32_t i = rowi[k];
      const double m = field ? std::hypot(val[2 * k], val[2 * k + 1]) : std::fabs(val[k]);
      if (m > scale[i]) scale[i] = m;
      if (i != j && m > scale[j]) scale[j] = m;
    }
  for (int32_t i = 0; i < n; ++i) {
    // inv_sqrt_scale_guarded (src/scaling/mod.rs:85-96)
    const double m = scale[i];
    double out = 1.0;
    if (m > 0.0) {
      const double cand = 1.0 / std::sqrt(m);
      if (std::isfinite(cand) && cand > 0.0) out = cand;
    }
    scale[i] = out;
  }
}

DeviceSymbolic::~DeviceSymbolic() {
  if (stream) (void)hipStreamDestroy(stream);
}

void DeviceSymbolic::upload() {
  if (uploaded) return;
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  const Plan& P = plan;
  sn_first.upload(P.sn_first, stream);
  sn_ncol.upload(P.sn_ncol, stream);
  sn_nrow.upload(P.sn_nrow, stream);
  e_off.upload(P.e_off, stream);
  panel_off.upload(P.panel_off, stream);
  rs_off.upload(P.rs_off, stream);
  rs_idx.upload(P.rs_idx, stream);
  upd_ptr.upload(P.upd_ptr, stream);
  upd_idx.upload(P.upd_idx, stream);
  sym_perm.upload(P.sym.perm, stream);
  snap_off.upload(P.snap_off, stream);
  snap_ld.upload(P.snap_ld, stream);
  amap.upload(P.amap, stream);
  a_row.upload(P.a_row, stream);
  a_col.upload(P.a_col, stream);
  items.upload(P.items, stream);
  tasks.upload(P.tasks, stream);
  u_off.upload(P.u_off, stream);
  gat_ptr.upload(P.gat_ptr, stream);
  gat_idx.upload(P.gat_idx, stream);
  fwd_items.upload(P.fwd_items, stream);
  bwd_items.upload(P.bwd_items, stream);
  std::vector<int32_t> lists;
  level_all_off.clear();
  level_small_off.clear();
  for (const LevelPlan& L : P.levels) {
    level_all_off.push_back((int64_t)lists.size());
    lists.insert(lists.end(), L.all.begin(), L.all.end());
    level_small_off.push_back((int64_t)lists.size());
    lists.insert(lists.end(), L.small.begin(), L.small.end());
  }
  level_lists.upload(lists, stream);
  u_colp.upload(h_colp, stream);
  u_rowi.upload(h_rowi, stream);
  hip_check(hipStreamSynchronize(stream), "plan upload");
  view.n = P.n;
  view.ns = P.ns;
  view.sn_first = sn_first.as<int32_t>();
  view.sn_ncol = sn_ncol.as<int32_t>();
  view.sn_nrow = sn_nrow.as<int32_t>();
  view.e_off = e_off.as<int32_t>();
  view.panel_off = panel_off.as<int64_t>();
  view.rs_off = rs_off.as<int64_t>();
  view.rs_idx = rs_idx.as<int32_t>();
  view.upd_ptr = upd_ptr.as<int32_t>();
  view.upd_idx = upd_idx.as<int32_t>();
  view.sym_perm = sym_perm.as<int32_t>();
  view.snap_off = snap_off.as<int64_t>();
  view.snap_ld = snap_ld.as<int32_t>();
  view.u_off = u_off.as<int64_t>();
  view.gat_ptr = gat_ptr.as<int64_t>();
  view.gat_idx = gat_idx.as<int64_t>();
  view.amap = amap.as<int64_t>();
  view.a_row = a_row.as<int32_t>();
  view.a_col = a_col.as<int32_t>();
  view.nnz = P.nnz;
  uploaded = true;
}

namespace {

template <class T>
struct FieldOf;
template <>
struct FieldOf<double> {
  static constexpr int id = 0;
  static constexpr int ndbl = 1;
};
template <>
struct FieldOf<cplx> {
  static constexpr int id = 1;
  static constexpr int ndbl = 2;
};

template <class T>
class DeviceFactor final : public DeviceFactorBase {
 public:
  DBuf pool, d, ds, two, lperm, e_of_g, perm_e, snap, deep, counters, status, scale_d, vals_d,
      work_y, work_b;
  DevFac<T> fv{};
  std::vector<double> scale;
  bool configured = false;

  DeviceFactor(std::shared_ptr<DeviceSymbolic> s, const FactorSettings& fs) {
    sym = std::move(s);
    settings = fs;
    if (sym->plan.params.panel_nb > 128)
      throw Error(INVALID_INPUT, "panel_nb above 128 is not supported by the device factor");
    const Plan& P = sym->plan;
    const int64_t n = P.n;
    pool.alloc(sizeof(T) * std::max<int64_t>(P.panel_total, 1));
    d.alloc(sizeof(T) * std::max<int64_t>(n, 1));
    ds.alloc(sizeof(T) * std::max<int64_t>(n, 1));
    two.alloc(std::max<int64_t>(n, 1));
    lperm.alloc(sizeof(int32_t) * std::max<int64_t>(n, 1));
    e_of_g.alloc(sizeof(int32_t) * std::max<int64_t>(n, 1));
    perm_e.alloc(sizeof(int32_t) * std::max<int64_t>(n, 1));
    snap.alloc(sizeof(T) * std::max<int64_t>(P.snap_total, 1));
    deep.alloc(sizeof(int32_t) * 128 * std::max<int64_t>(P.ns, 1));
    counters.alloc(sizeof(unsigned long long) * 4);
    status.alloc(sizeof(int32_t) * 2);
    scale_d.alloc(sizeof(double) * std::max<int64_t>(n, 1));
    vals_d.alloc(sizeof(T) * std::max<int64_t>(P.nnz, 1));
    fv.pool = pool.as<T>();
    fv.d = d.as<T>();
    fv.ds = ds.as<T>();
    fv.two = two.as<uint8_t>();
    fv.lperm = lperm.as<int32_t>();
    fv.e_of_g = e_of_g.as<int32_t>();
    fv.perm_e = perm_e.as<int32_t>();
    fv.snap = snap.as<T>();
    fv.deep = deep.as<int32_t>();
    fv.counters = counters.as<unsigned long long>();
    fv.status = status.as<int32_t>();
  }

  int field() const override { return FieldOf<T>::id; }

  void refactor(const double* values) override {
    const Plan& P = sym->plan;
    const DevSym& S = sym->view;
    hipStream_t st = sym->stream;
    if (!configured) {
      hip_check(configure_factor_kernels<T>(), "configure kernels");
      configured = true;
    }
    auto t0 = std::chrono::steady_clock::now();
    if (settings.scaling == 1)
      onepass_scale(P.n, FieldOf<T>::id, sym->h_colp.data(), sym->h_rowi.data(), values, scale);
    else
      scale.assign(P.n, 1.0);
    PivotPolicy pol{0, 0.0};
    if (settings.pivot == 1) {
      pol.have_floor = 1;
      pol.floor = std::max(settings.abs_floor, 0.0);
    } else if (settings.pivot == 2) {
      // ForceAccept floor: max(|A_hat|max, 1) * eps over the equilibrated values
      double amax = 0.0;
      for (int32_t j = 0; j < P.n; ++j)
        for (int64_t k = sym->h_colp[j]; k < sym->h_colp[j + 1]; ++k) {
          const int32_t i = sym->h_rowi[k];
          const double m = FieldOf<T>::id ? std::hypot(values[2 * k], values[2 * k + 1])
                                          : std::fabs(values[k]);
          amax = std::max(amax, m * scale[i] * scale[j]);
        }
      pol.have_floor = 1;
      pol.floor = std::max(amax, 1.0) * 2.220446049250313e-16;
    }
    hip_check(hipMemcpyAsync(scale_d.p, scale.data(), sizeof(double) * P.n, hipMemcpyHostToDevice, st),
              "upload scale");
    hip_check(hipMemcpyAsync(vals_d.p, values, sizeof(T) * P.nnz, hipMemcpyHostToDevice, st),
              "upload values");
    hip_check(hipMemsetAsync(counters.p, 0, counters.bytes, st), "memset");
    const int32_t st_init[2] = {0, INT32_MAX};
    hip_check(hipMemcpyAsync(status.p, st_init, sizeof(st_init), hipMemcpyHostToDevice, st), "status");
    launch_fill<T>(fv.pool, P.panel_total, F<T>::zero(), st);
    launch_assemble<T>(S, fv, vals_d.as<T>(), scale_d.as<double>(), st);
    const int nb = P.params.panel_nb < 8 ? 8 : P.params.panel_nb;
    const int32_t* lists = sym->level_lists.as<int32_t>();
    for (size_t L = 0; L < P.levels.size(); ++L) {
      const LevelPlan& lv = P.levels[L];
      launch_cmod<T>(S, fv, sym->items.as<CItem>() + lv.item_beg, (int)(lv.item_end - lv.item_beg),
                     sym->tasks.as<CTask>(), st);
      launch_cdiv_small<T>(S, fv, lists + sym->level_small_off[L], (int)lv.small.size(), nb, pol, st);
      for (int32_t s : lv.wide)
        launch_cdiv_wide<T>(S, fv, s, P.sn_ncol[s], P.sn_nrow[s], nb, pol, st);
    }
    hip_check(hipGetLastError(), "factor launch");
    unsigned long long cnt[4];
    int32_t stat[2];
    hip_check(hipMemcpyAsync(cnt, counters.p, sizeof(cnt), hipMemcpyDeviceToHost, st), "counters");
    hip_check(hipMemcpyAsync(stat, status.p, sizeof(stat), hipMemcpyDeviceToHost, st), "status");
    hip_check(hipStreamSynchronize(st), "factor");
    stats.t_factor_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stat[0] != 0) throw Error(NUMERICALLY_RANK_DEFICIENT, "numerically rank deficient");
    stats.inertia[0] = (int64_t)cnt[0];
    stats.inertia[1] = (int64_t)cnt[1];
    stats.inertia[2] = (int64_t)cnt[2];
    stats.n_perturbed = (int64_t)cnt[3];
  }

  // y (device, e-order) := solve in place; b/x device user-order buffers
  void solve_device(const void* d_b, int64_t nrhs, void* d_x, hipStream_t st) override {
    const Plan& P = sym->plan;
    const DevSym& S = sym->view;
    if (!st) st = sym->stream;
    const size_t need = sizeof(T) * std::max<int64_t>(P.n * nrhs, 1);
    if (work_y.bytes < need) work_y.alloc(need);
    const size_t needu = sizeof(T) * std::max<int64_t>(P.u_total * nrhs, 1);
    if (work_b.bytes < needu) work_b.alloc(needu);
    T* y = work_y.as<T>();
    T* u = work_b.as<T>();
    const int2* fi = sym->fwd_items.as<int2>();
    const int2* bi = sym->bwd_items.as<int2>();
    launch_gather_rhs<T>(P.n, (int32_t)nrhs, fv.perm_e, scale_d.as<double>(),
                         static_cast<const T*>(d_b), y, st);
    const int32_t* lists = sym->level_lists.as<int32_t>();
    for (size_t L = 0; L < P.levels.size(); ++L) {
      const LevelPlan& lv = P.levels[L];
      launch_fwd_tri<T>(S, fv, lists + sym->level_all_off[L], (int)lv.all.size(), (int32_t)nrhs, y,
                        u, st);
      launch_fwd_update<T>(S, fv, fi + lv.fwd_beg, (int)(lv.fwd_end - lv.fwd_beg), (int32_t)nrhs,
                           y, u, st);
    }
    launch_dsolve<T>(P.n, (int32_t)nrhs, fv.d, fv.ds, fv.two, y, st);
    for (size_t L = P.levels.size(); L-- > 0;) {
      const LevelPlan& lv = P.levels[L];
      launch_bwd_update<T>(S, fv, bi + lv.bwd_beg, (int)(lv.bwd_end - lv.bwd_beg), (int32_t)nrhs, y,
                           st);
      launch_bwd_tri<T>(S, fv, lists + sym->level_all_off[L], (int)lv.all.size(), (int32_t)nrhs, y,
                        st);
    }
    launch_scatter_sol<T>(P.n, (int32_t)nrhs, fv.perm_e, scale_d.as<double>(), y,
                          static_cast<T*>(d_x), st);
    hip_check(hipGetLastError(), "solve launch");
  }

  void solve_host(const double* b, int64_t nrhs, double* x) override {
    const Plan& P = sym->plan;
    hipStream_t st = sym->stream;
    const size_t bytes = sizeof(T) * P.n * nrhs;
    DBuf db, dx;
    db.alloc(std::max<size_t>(bytes, 1));
    dx.alloc(std::max<size_t>(bytes, 1));
    hip_check(hipMemcpyAsync(db.p, b, bytes, hipMemcpyHostToDevice, st), "upload rhs");
    solve_device(db.p, nrhs, dx.p, st);
    hip_check(hipMemcpyAsync(x, dx.p, bytes, hipMemcpyDeviceToHost, st), "download solution");
    hip_check(hipStreamSynchronize(st), "solve");
  }

  void download(std::vector<int64_t>& colp, std::vector<int32_t>& rowi, std::vector<double>& lval,
                std::vector<double>& dd, std::vector<double>& dsd, std::vector<uint8_t>& tw,
                std::vector<int32_t>& perm) override {
    // ldlt_emit_and_free (multifrontal_ldlt.rs:2328-2381) + CSC concat (:3531-3575)
    const Plan& P = sym->plan;
    hipStream_t st = sym->stream;
    const int64_t n = P.n;
    std::vector<T> hp(P.panel_total), hd(n), hds(n);
    std::vector<int32_t> hl(n), heg(n), hpe(n);
    tw.assign(n, 0);
    hip_check(hipMemcpyAsync(hp.data(), pool.p, sizeof(T) * P.panel_total, hipMemcpyDeviceToHost, st), "dl");
    hip_check(hipMemcpyAsync(hd.data(), d.p, sizeof(T) * n, hipMemcpyDeviceToHost, st), "dl");
    hip_check(hipMemcpyAsync(hds.data(), ds.p, sizeof(T) * n, hipMemcpyDeviceToHost, st), "dl");
    hip_check(hipMemcpyAsync(tw.data(), two.p, n, hipMemcpyDeviceToHost, st), "dl");
    hip_check(hipMemcpyAsync(hl.data(), lperm.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st), "dl");
    hip_check(hipMemcpyAsync(heg.data(), e_of_g.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st), "dl");
    hip_check(hipMemcpyAsync(hpe.data(), perm_e.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st), "dl");
    hip_check(hipStreamSynchronize(st), "download");
    const int nd = FieldOf<T>::ndbl;
    dd.assign(n * nd, 0.0);
    dsd.assign(n * nd, 0.0);
    std::memcpy(dd.data(), hd.data(), sizeof(T) * n);
    std::memcpy(dsd.data(), hds.data(), sizeof(T) * n);
    perm = hpe;
    colp.assign(1, 0);
    rowi.clear();
    lval.clear();
    std::vector<std::pair<int32_t, T>> col;
    for (int32_t s = 0; s < P.ns; ++s) {
      const int64_t h = P.sn_nrow[s], w = P.sn_ncol[s], eo = P.e_off[s];
      const int32_t* rs = P.rs_idx.data() + P.rs_off[s];
      const T* pan = hp.data() + P.panel_off[s];
      auto grow = [&](int64_t i) { return i < w ? rs[hl[eo + i]] : rs[i]; };
      for (int64_t p = 0; p < w; ++p) {
        col.clear();
        col.push_back({heg[grow(p)], F<T>::one()});
        const int64_t i0 = tw[eo + p] ? p + 2 : p + 1;
        for (int64_t i = i0; i < h; ++i) {
          const T v = pan[i + p * h];
          if (!F<T>::is_zero(v)) col.push_back({heg[grow(i)], v});
        }
        std::sort(col.begin(), col.end(),
                  [](const std::pair<int32_t, T>& a, const std::pair<int32_t, T>& b) { return a.first < b.first; });
        for (auto& e : col) {
          rowi.push_back(e.first);
          const double* pv = reinterpret_cast<const double*>(&e.second);
          for (int q = 0; q < nd; ++q) lval.push_back(pv[q]);
        }
        colp.push_back((int64_t)rowi.size());
      }
    }
  }

  void refine(const double*, const double*, int64_t, double, double*, double*) override;
};

template <class T>
void DeviceFactor<T>::refine(const double*, const double*, int64_t, double, double*, double*) {
  throw Error(INVALID_INPUT, "refinement ladder not available in this build");
}

}  // namespace

std::unique_ptr<DeviceFactorBase> make_factor(std::shared_ptr<DeviceSymbolic> sym, int field,
                                              const double* values, const FactorSettings& fs) {
  sym->upload();
  std::unique_ptr<DeviceFactorBase> f;
  if (field == 0) f.reset(new DeviceFactor<double>(sym, fs));
  else if (field == 1) f.reset(new DeviceFactor<cplx>(sym, fs));
  else throw Error(INVALID_INPUT, "unknown field");
  f->refactor(values);
  return f;
}

}  // namespace rsl
