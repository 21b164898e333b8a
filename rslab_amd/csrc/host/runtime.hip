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
#include <mutex>

#include "../../../include/rslab_amd.h"
#include "krylov.hpp"

namespace rsl {

// kernel launchers (factor_kernels.hip, solve_kernels.hip, refine_kernels.hip)
template <class T> void launch_fill(T*, int64_t, T, hipStream_t);
void launch_factor_init(unsigned long long*, int32_t*, hipStream_t);
template <class T> void launch_drop(const DevSym&, const DevFac<T>&, double, hipStream_t);
template <class T> void launch_rows_e(const DevSym&, const DevFac<T>&, int64_t, hipStream_t);
template <class T> void launch_assemble(const DevSym&, const DevFac<T>&, const T*, const double*, hipStream_t, LaunchProfiler*);
template <class T> void launch_cmod(const DevSym&, const DevFac<T>&, const CItem*, int, const CTask*, double, int, hipStream_t, LaunchProfiler*);
void launch_device_scale(int32_t, int64_t, const int32_t*, const int32_t*, const double*, int, unsigned long long*,
                         double*, hipStream_t);
void launch_scaled_amax(int64_t, const int32_t*, const int32_t*, const double*, int, const double*,
                        unsigned long long*, hipStream_t);
void launch_row_abs_max(int32_t, const int64_t*, const int64_t*, const double*, int, unsigned long long*,
                        hipStream_t);
void launch_cmod_two_phase(const DevSym&, const DevFac<double>&, const CItem*, int, const CTask*, const PTile*,
                           int64_t, void*, int64_t, const double*, double, int, hipStream_t, LaunchProfiler*);
void launch_make_g(const DevSym&, const DevFac<double>&, const int32_t*, int, int, double*, hipStream_t);
template <class T> void launch_cdiv_small(const DevSym&, const DevFac<T>&, const int32_t*, int, int, PivotPolicy, double, int, hipStream_t, LaunchProfiler*);
template <class T> void launch_wide_level(const DevSym&, const DevFac<T>&, const WItem*, const WideStep*, int, const int32_t*, int, PivotPolicy, int, hipStream_t, LaunchProfiler*);
template <class T> hipError_t configure_factor_kernels();
template <class T> void launch_gather_rhs(int32_t, int32_t, const int32_t*, const double*, const T*, T*, hipStream_t, LaunchProfiler*);
template <class T> void launch_scatter_sol(int32_t, int32_t, const int32_t*, const double*, const T*, T*, hipStream_t, LaunchProfiler*);
template <class T> void launch_fgather(const DevSym&, const DevFac<T>&, const int2*, int, int32_t, T*, const T*, hipStream_t, LaunchProfiler*);
template <class T> void launch_fwd_tri(const DevSym&, const DevFac<T>&, const int32_t*, int, int32_t, T*, double, hipStream_t, LaunchProfiler*);
template <class T> void launch_ftri_step(const DevSym&, const DevFac<T>&, const SItem*, const SolveStep&, int32_t, T*, T*, hipStream_t, LaunchProfiler*);
template <class T> void launch_btri_step(const DevSym&, const DevFac<T>&, const SItem*, const SolveStep&, int32_t, T*, T*, hipStream_t, LaunchProfiler*);
template <class T> void launch_zcopy(const DevSym&, const SItem*, int, int32_t, const T*, T*, hipStream_t);
template <class T> void launch_fwd_update(const DevSym&, const DevFac<T>&, const int2*, int, int32_t, const T*, T*, double, hipStream_t, LaunchProfiler*);
template <class T> void launch_dsolve(int32_t, int32_t, const T*, const T*, const uint8_t*, T*, hipStream_t, LaunchProfiler*);
template <class T> void launch_bwd_update(const DevSym&, const DevFac<T>&, const int2*, int, int32_t, T*, double, hipStream_t, LaunchProfiler*);
template <class T> void launch_bwd_tri(const DevSym&, const DevFac<T>&, const int32_t*, int, int32_t, T*, double, hipStream_t, LaunchProfiler*);
void launch_sync_init(int32_t*, const int32_t*, int32_t, hipStream_t);
template <class T> void launch_fwd_dep(const DevSym&, const DevFac<T>&, const int2*, int, const SolveSync&, int32_t, T*, T*, T*, int, double, hipStream_t, LaunchProfiler*);
template <class T> void launch_bwd_dep(const DevSym&, const DevFac<T>&, const int2*, int, const SolveSync&, int32_t, T*, T*, double, hipStream_t, LaunchProfiler*);
template <class T> void launch_expand(int64_t, const int64_t*, const T*, T*, hipStream_t);
template <class T> void launch_residual(int32_t, int32_t, const int64_t*, const int32_t*, const T*, const T*, const T*, T*, hipStream_t, LaunchProfiler*);
template <class T> void launch_colmax(int32_t, int32_t, const T*, unsigned long long*, hipStream_t);
template <class T> void launch_update_cols(int32_t, int32_t, const uint8_t*, T*, const T*, hipStream_t);
template <class T> void launch_copy_cols(int32_t, int32_t, const uint8_t*, T*, const T*, hipStream_t);
template <class T> void launch_copy_col(int32_t, int32_t, int32_t, const T*, int32_t, int32_t, T*, hipStream_t);

const char* launch_kind_name(int k) {
  static const char* names[LK_COUNT] = {"assemble",     "cmod",        "cdiv_small", "panel_getf2",
                                        "panel_replay", "panel_schur", "emit",       "solve_gather",
                                        "fwd_tri",      "fwd_update",  "dsolve",     "bwd_update",
                                        "bwd_tri",      "solve_scatter", "residual",   "cmod_sum",   "make_g"};
  return (k >= 0 && k < LK_COUNT) ? names[k] : "?";
}

EventProfiler::~EventProfiler() {
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

hipEvent_t EventProfiler::take() {
  if (used < pool.size()) return pool[used++];
  hipEvent_t e;
  hip_check(hipEventCreate(&e), "hipEventCreate");
  pool.push_back(e);
  ++used;
  return e;
}

void EventProfiler::begin(int kind, double flops, double bytes, hipStream_t st) {
  Rec r;
  r.kind = kind;
  r.flops = flops * flop_mult;
  r.bytes = bytes;
  r.a = take();
  hip_check(hipEventRecord(r.a, st), "hipEventRecord");
  recs.push_back(r);
}

void EventProfiler::end(hipStream_t st) {
  Rec& r = recs.back();
  r.b = take();
  hip_check(hipEventRecord(r.b, st), "hipEventRecord");
}

void EventProfiler::collect() {
  for (const Rec& r : recs) {
    hip_check(hipEventSynchronize(r.b), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
    Acc& a = acc[r.kind];
    a.ms += ms;
    a.flops += r.flops;
    a.bytes += r.bytes;
    a.launches += 1;
  }
  recs.clear();
  used = 0;
}

void EventProfiler::reset() {
  recs.clear();
  used = 0;
  for (Acc& a : acc) a = Acc();
}

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
  g_off.upload(P.g_off, stream);
  rs_off.upload(P.rs_off, stream);
  rs_idx.upload(P.rs_idx, stream);
  upd_ptr.upload(P.upd_ptr, stream);
  upd_idx.upload(P.upd_idx, stream);
  sym_perm.upload(P.sym.perm, stream);
  amap.upload(P.amap, stream);
  a_row.upload(P.a_row, stream);
  a_col.upload(P.a_col, stream);
  items.upload(P.items, stream);
  tasks.upload(P.tasks, stream);
  ptiles.upload(P.ptiles, stream);
  u_off.upload(P.u_off, stream);
  gat_ptr.upload(P.gat_ptr, stream);
  gat_idx.upload(P.gat_idx, stream);
  fwd_items.upload(P.fwd_items, stream);
  bwd_items.upload(P.bwd_items, stream);
  std::vector<int32_t> lists;
  level_all_off.clear();
  level_small_off.clear();
  level_wide_off.clear();
  level_snarrow_off.clear();
  for (const LevelPlan& L : P.levels) {
    level_all_off.push_back((int64_t)lists.size());
    lists.insert(lists.end(), L.all.begin(), L.all.end());
    level_small_off.push_back((int64_t)lists.size());
    lists.insert(lists.end(), L.small.begin(), L.small.end());
    level_wide_off.push_back((int64_t)lists.size());
    lists.insert(lists.end(), L.wide.begin(), L.wide.end());
    level_snarrow_off.push_back((int64_t)lists.size());
    lists.insert(lists.end(), L.snarrow.begin(), L.snarrow.end());
  }
  level_lists.upload(lists, stream);
  witems.upload(P.witems, stream);
  gat_items.upload(P.gat_items, stream);
  sitems.upload(P.sitems, stream);
  bitems.upload(P.bitems, stream);
  zitems.upload(P.zitems, stream);
  nb_eff.upload(P.nb_eff, stream);
  dfwd_items.upload(P.dfwd_items, stream);
  dfwd_chain.upload(P.dfwd_chain, stream);
  dbwd_items.upload(P.dbwd_items, stream);
  fu_init.upload(P.fu_init, stream);
  anc_ptr.upload(P.anc_ptr, stream);
  anc_idx.upload(P.anc_idx, stream);
  hip_check(hipStreamSynchronize(stream), "plan upload");
  view.n = P.n;
  view.ns = P.ns;
  view.sn_first = sn_first.as<int32_t>();
  view.sn_ncol = sn_ncol.as<int32_t>();
  view.g_off = g_off.as<int64_t>();
  view.sn_nrow = sn_nrow.as<int32_t>();
  view.e_off = e_off.as<int32_t>();
  view.panel_off = panel_off.as<int64_t>();
  view.rs_off = rs_off.as<int64_t>();
  view.rs_idx = rs_idx.as<int32_t>();
  view.upd_ptr = upd_ptr.as<int32_t>();
  view.upd_idx = upd_idx.as<int32_t>();
  view.sym_perm = sym_perm.as<int32_t>();
  view.u_off = u_off.as<int64_t>();
  view.gat_ptr = gat_ptr.as<int64_t>();
  view.gat_idx = gat_idx.as<int64_t>();
  view.amap = amap.as<int64_t>();
  view.a_row = a_row.as<int32_t>();
  view.a_col = a_col.as<int32_t>();
  view.nb_eff = nb_eff.as<int32_t>();
  view.anc_ptr = anc_ptr.as<int32_t>();
  view.anc_idx = anc_idx.as<int32_t>();
  view.nnz = P.nnz;
  uploaded = true;
}

// CscMatrix::symv (src/sparse/csc.rs:294-309) applies the stored lower
// triangle twice; the device applies the same entries as one full CSR.
void DeviceSymbolic::ensure_full_csr() {
  if (full_ready) return;
  const int32_t n = plan.n;
  std::vector<int64_t> rp(n + 1, 0);
  for (int32_t j = 0; j < n; ++j)
    for (int64_t q = h_colp[j]; q < h_colp[j + 1]; ++q) {
      const int32_t i = h_rowi[q];
      rp[i + 1]++;
      if (i != j) rp[j + 1]++;
    }
  for (int32_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
  nnz_full = rp[n];
  std::vector<int32_t> ci(nnz_full);
  std::vector<int64_t> src(nnz_full);
  std::vector<int64_t> pos(rp.begin(), rp.end() - 1);
  for (int32_t j = 0; j < n; ++j)
    for (int64_t q = h_colp[j]; q < h_colp[j + 1]; ++q) {
      const int32_t i = h_rowi[q];
      ci[pos[i]] = j;
      src[pos[i]++] = q;
      if (i != j) {
        ci[pos[j]] = i;
        src[pos[j]++] = q;
      }
    }
  fa_rp.upload(rp, stream);
  fa_ci.upload(ci, stream);
  fa_src.upload(src, stream);
  hip_check(hipStreamSynchronize(stream), "full CSR upload");
  full_ready = true;
}

// ---- fixed scratch for the no-allocation solve (rsl_solve_device_noalloc) ----
// One array per GPU, part of the library's code object: the loader places it
// in device memory when the library's kernels are first loaded on a GPU and
// it stays until the process ends. Host code never uses its name as an
// address: the device address comes from hipGetSymbolAddress, once per
// device, with that device current.
constexpr size_t kNoallocScratchBytes = RSL_SOLVE_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device (looked up on the first
// call for that device, then cached).
void* noalloc_scratch_current_device() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw Error(DEVICE_ERROR, "device: device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw Error(DEVICE_ERROR, "device: hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
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
  DBuf pool, d, ds, two, lperm, e_of_g, perm_e, deep, counters, status, scale_d, vals_d, sn_inert, rows_e,
      work_y, work_b, work_z, cbuf, gpool, dev_u64;
  // dependency-driven solve: one 128-byte line of counters per supernode + 4 control words
  static constexpr int64_t kSyncStride = 32;
  DBuf solve_sync;
  SolveSync sy{};
  // refinement workspace, kept across calls (grown on demand; calls on one
  // handle are serialized by the C ABI)
  DBuf rf_afull, rf_dx, rf_dbest, rf_dr, rf_ddx, rf_norms, rf_mask, rf_mask2;
  DBuf hs_db, hs_dx, hr_av, hr_db, hr_dxo;                                   // host-buffer calls
  DBuf gm_bv, gm_xv, gm_rv, gm_nrm;  // GMRES-IR: the escalated columns' b, x, r and norms
  KrylovEngine<T> kry;
  const T* kry_A = nullptr;  // operator values (full symmetric CSR) of the current Krylov call
  bool kry_pre = true;       // precondition with this factor (else the identity)

  // The Krylov engine's operator is A (full symmetric CSR values on the
  // device, the plan's pattern), its preconditioner this factor's solve.
  void bind_krylov(const T* A, bool pre) {
    sym->ensure_full_csr();
    kry_A = A;
    kry_pre = pre;
    kry.n = sym->plan.n;
    kry.st = sym->stream;
    const int64_t* rp = sym->fa_rp.as<int64_t>();
    const int32_t* ci = sym->fa_ci.as<int32_t>();
    kry.apply = [this, rp, ci](const T* x, T* y, int32_t k) {
      launch_residual<T>(sym->plan.n, k, rp, ci, kry_A, x, nullptr, y, sym->stream, prof_on ? &prof : nullptr);
    };
    kry.resid = [this, rp, ci](const T* x, const T* b, T* r, int32_t k) {
      launch_residual<T>(sym->plan.n, k, rp, ci, kry_A, x, b, r, sym->stream, prof_on ? &prof : nullptr);
    };
    kry.precond = [this](const T* r, T* z, int32_t k) {
      if (kry_pre)
        solve_device(r, k, z, sym->stream);
      else
        hip_check(hipMemcpyAsync(z, r, sizeof(T) * (size_t)sym->plan.n * k, hipMemcpyDeviceToDevice, sym->stream),
                  "identity preconditioner");
    };
  }

  // Krylov solvers over host buffers (method 0: FGMRES / block GMRES,
  // 1: COCG, 2: COCR); info[4c..4c+4) = {iterations, final relative residual,
  // converged, stop reason}
  void krylov(int method, const double* a_values, const double* b, int64_t nrhs, double tol, int max_iter,
              int restart, int precond, const double* x0, double* x, double* info) override {
    const Plan& P = sym->plan;
    hipStream_t st = sym->stream;
    const int32_t n = P.n, k = (int32_t)nrhs;
    if (precond && !valid) throw Error(INVALID_INPUT, "the last factorization of this handle failed; refactor first");
    if (k > 256) throw Error(INVALID_INPUT, "Krylov solvers take at most 256 right-hand sides per call");
    if (max_iter < 0 || (method == 0 && restart < 1)) throw Error(INVALID_INPUT, "iteration limits must be positive");
    sym->ensure_full_csr();
    const size_t bytes = sizeof(T) * (size_t)n * k;
    grow(hr_av, sizeof(T) * std::max<int64_t>(P.nnz, 1));
    grow(rf_afull, sizeof(T) * std::max<int64_t>(sym->nnz_full, 1));
    grow(hr_db, std::max<size_t>(bytes, 1));
    grow(hr_dxo, std::max<size_t>(bytes, 1));
    hip_check(hipMemcpyAsync(hr_av.p, a_values, sizeof(T) * P.nnz, hipMemcpyHostToDevice, st), "upload A");
    launch_expand<T>(sym->nnz_full, sym->fa_src.as<int64_t>(), hr_av.as<T>(), rf_afull.as<T>(), st);
    hip_check(hipMemcpyAsync(hr_db.p, b, bytes, hipMemcpyHostToDevice, st), "upload b");
    if (x0) hip_check(hipMemcpyAsync(hr_dxo.p, x0, bytes, hipMemcpyHostToDevice, st), "upload x0");
    bind_krylov(rf_afull.as<T>(), precond != 0);
    std::vector<KrylovCol> res;
    if (method == 0)
      kry.fgmres(hr_db.as<T>(), hr_dxo.as<T>(), k, tol, max_iter, restart, x0 != nullptr, res);
    else if (method == 1)
      kry.cocg(hr_db.as<T>(), hr_dxo.as<T>(), k, tol, max_iter, res);
    else if (method == 2)
      kry.cocr(hr_db.as<T>(), hr_dxo.as<T>(), k, tol, max_iter, res);
    else
      throw Error(INVALID_INPUT, "unknown Krylov method");
    hip_check(hipMemcpyAsync(x, hr_dxo.p, bytes, hipMemcpyDeviceToHost, st), "download x");
    hip_check(hipStreamSynchronize(st), "krylov");
    check_solve_waits();
    for (int c = 0; c < k; ++c) {
      info[4 * c + 0] = res[c].iters;
      info[4 * c + 1] = res[c].final_res;
      info[4 * c + 2] = res[c].converged ? 1.0 : 0.0;
      info[4 * c + 3] = res[c].stop;
    }
  }
  static void grow(DBuf& b, size_t nb) {
    if (b.bytes < nb) b.alloc(nb);
  }
  hipEvent_t ev_caller = nullptr, ev_done = nullptr;
  bool ev_ready = false;
  // a level's one-workgroup cdiv runs on the side stream beside its wide
  // panel steps (independent supernodes), joined before the next level
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int64_t c_cap = 0;  // product-buffer elements: levels needing more use the gathered cmod
  void* noalloc_base = nullptr;  // device address of the fixed scratch on this factor's device
  DevFac<T> fv{};
  std::vector<double> scale;
  bool configured = false;

  ~DeviceFactor() override {
    drop_graph();
    if (ev_ready) {
      (void)hipEventDestroy(ev_caller);
      (void)hipEventDestroy(ev_done);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
  }

  DeviceFactor(std::shared_ptr<DeviceSymbolic> s, const FactorSettings& fs) {
    sym = std::move(s);
    settings = fs;
    prof.flop_mult = FieldOf<T>::id ? 4.0 : 1.0;
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
    deep.alloc(sizeof(int32_t) * 2 * 128 * std::max<int64_t>(P.ns, 1));
    counters.alloc(sizeof(unsigned long long) * 8);
    sn_inert.alloc(sizeof(int32_t) * 3 * std::max<int64_t>(P.ns, 1));
    rows_e.alloc(sizeof(int32_t) * std::max<int64_t>(P.rs_off[P.ns], 1));
    // index maps start at 0 so no later read can leave the vectors
    hip_check(hipMemset(e_of_g.p, 0, e_of_g.bytes), "memset");
    hip_check(hipMemset(perm_e.p, 0, perm_e.bytes), "memset");
    hip_check(hipMemset(rows_e.p, 0, rows_e.bytes), "memset");
    if (FieldOf<T>::id == 0 && P.c_elems_max > 0) {
      // two-phase cmod product buffer, at most a quarter of the free memory
      const size_t esz = fs.precision == Precision::FP64 ? sizeof(double) : sizeof(float);
      size_t fr = 0, tot = 0;
      hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
      c_cap = std::min<int64_t>(P.c_elems_max, (int64_t)(fr / 4 / esz));
      if (fs.cmod_cap >= 0) c_cap = std::min<int64_t>(c_cap, fs.cmod_cap);
      if (c_cap > 0) {
        cbuf.alloc(esz * (size_t)c_cap);
        gpool.alloc(sizeof(double) * (size_t)std::max<int64_t>(P.g_total, 1));  // L21 D of every supernode
      }
    }
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
    fv.deep = deep.as<int32_t>();
    fv.counters = counters.as<unsigned long long>();
    fv.sn_inertia = sn_inert.as<int32_t>();
    fv.rs_e = rows_e.as<int32_t>();
    fv.status = status.as<int32_t>();
    solve_sync.alloc(sizeof(int32_t) * (kSyncStride * (size_t)std::max<int64_t>(P.ns, 1) + 4));
    hip_check(hipMemset(solve_sync.p, 0, solve_sync.bytes), "memset");
    sy.cnt = solve_sync.as<int32_t>();
    sy.ctl = reinterpret_cast<uint32_t*>(sy.cnt + kSyncStride * std::max<int64_t>(P.ns, 1));
    hip_check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming), "hipEventCreate");
    // set-up for the calls that must not allocate later: the ordering events
    // and the fixed scratch's address on this (current) device
    make_events();
    noalloc_base = noalloc_scratch_current_device();
  }

  int field() const override { return FieldOf<T>::id; }

  void refactor(const double* values) override {
    const Plan& P = sym->plan;
    hipStream_t st = sym->stream;
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
    run_factor(vals_d.as<T>(), pol, t0);
  }

  // values already in device memory: the equilibration and the pivot floor
  // are computed on the device, to the host path's bits in the real field
  // (the complex field's modulus is the device's hypot: equal to rounding)
  void refactor_device(const void* d_values, hipStream_t caller) override {
    {
      const Plan& P = sym->plan;
      hipStream_t st = sym->stream;
      auto t0 = std::chrono::steady_clock::now();
      wait_for(caller);
      const double* v = static_cast<const double*>(d_values);
      const int cplx = FieldOf<T>::id;
      if (dev_u64.bytes < sizeof(unsigned long long) * (size_t)(P.n + 1))
        dev_u64.alloc(sizeof(unsigned long long) * (size_t)(P.n + 1));
      auto* u64 = dev_u64.as<unsigned long long>();
      if (settings.scaling == 1)
        launch_device_scale(P.n, P.nnz, sym->a_row.as<int32_t>(), sym->a_col.as<int32_t>(), v, cplx, u64,
                            scale_d.as<double>(), st);
      else
        launch_fill<double>(scale_d.as<double>(), P.n, 1.0, st);
      PivotPolicy pol{0, 0.0};
      if (settings.pivot == 1) {
        pol.have_floor = 1;
        pol.floor = std::max(settings.abs_floor, 0.0);
      } else if (settings.pivot == 2) {
        unsigned long long bits = 0;
        launch_scaled_amax(P.nnz, sym->a_row.as<int32_t>(), sym->a_col.as<int32_t>(), v, cplx,
                           scale_d.as<double>(), u64 + P.n, st);
        hip_check(hipMemcpyAsync(&bits, u64 + P.n, sizeof(bits), hipMemcpyDeviceToHost, st), "amax");
        hip_check(hipStreamSynchronize(st), "amax");
        double amax;
        std::memcpy(&amax, &bits, sizeof(amax));
        pol.have_floor = 1;
        pol.floor = std::max(amax, 1.0) * 2.220446049250313e-16;
      }
      // the replayed graph reads the handle's own values buffer
      const T* src = static_cast<const T*>(d_values);
      if (settings.graphs && !prof_on) {
        hip_check(hipMemcpyAsync(vals_d.p, d_values, sizeof(T) * P.nnz, hipMemcpyDeviceToDevice, st),
                  "copy values");
        src = vals_d.as<T>();
      }
      run_factor(src, pol, t0);
    }
  }

  void make_events() {
    if (ev_ready) return;
    hip_check(hipEventCreateWithFlags(&ev_caller, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&ev_done, hipEventDisableTiming), "hipEventCreate");
    ev_ready = true;
  }

  // the library stream waits for work the caller enqueued before the call
  void wait_for(hipStream_t caller) {
    make_events();
    hip_check(hipEventRecord(ev_caller, caller), "hipEventRecord");
    hip_check(hipStreamWaitEvent(sym->stream, ev_caller, 0), "hipStreamWaitEvent");
  }

  // the caller's stream waits for the work this call enqueued on the library stream
  void signal_to(hipStream_t caller) {
    hip_check(hipEventRecord(ev_done, sym->stream), "hipEventRecord");
    hip_check(hipStreamWaitEvent(caller, ev_done, 0), "hipStreamWaitEvent");
  }

  // Asynchronous solve for a caller's stream: runs on the library stream after
  // the caller's earlier work, and the caller's stream waits for it. Every
  // later call on this handle (refactor, refine, solve) also runs on the
  // library stream, so it cannot overwrite the factor or the workspace while
  // this solve still reads them.
  void solve_async(const void* d_b, int64_t nrhs, void* d_x, hipStream_t caller) override {
    wait_for(caller);
    solve_device(d_b, nrhs, d_x, sym->stream);
    signal_to(caller);
  }

  // numeric factor from device values (user lower CSC) with scale_d in place
  // A same-pattern refactor replays the factor's launch sequence as one HIP
  // graph: captured on the first refactor (the plan, the storage and every
  // grid are fixed per handle), replayed while the values buffer and the
  // pivot policy are the ones it was captured with. The sequence holds no
  // memset or host copy (the status is reset by a kernel). Eager when
  // profiling (each launch is bracketed by events) or when graphs are off.
  hipGraphExec_t fgraph = nullptr;
  const T* fgraph_vals = nullptr;
  PivotPolicy fgraph_pol{0, 0.0};
  bool graph_broken = false;  // capture failed once: stay eager
  int64_t factors_done = 0;

  void drop_graph() {
    if (fgraph) (void)hipGraphExecDestroy(fgraph);
    fgraph = nullptr;
  }

  void launch_factor(const T* vals, const PivotPolicy& pol, hipStream_t st) {
    const bool want = settings.graphs && !prof_on && !graph_broken && factors_done > 0;
    if (!want) {
      enqueue_factor(vals, pol, st, prof_on ? &prof : nullptr);
      return;
    }
    if (!fgraph || fgraph_vals != vals || fgraph_pol.have_floor != pol.have_floor || fgraph_pol.floor != pol.floor) {
      drop_graph();
      hipGraph_t g = nullptr;
      hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      try {
        enqueue_factor(vals, pol, st, nullptr);
      } catch (...) {
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      hipError_t e = hipStreamEndCapture(st, &g);
      if (e == hipSuccess && g) e = hipGraphInstantiate(&fgraph, g, nullptr, nullptr, 0);
      if (g) (void)hipGraphDestroy(g);
      if (e != hipSuccess || !fgraph) {
        (void)hipGetLastError();
        fgraph = nullptr;
        graph_broken = true;
        enqueue_factor(vals, pol, st, nullptr);
        return;
      }
      fgraph_vals = vals;
      fgraph_pol = pol;
    }
    hip_check(hipGraphLaunch(fgraph, st), "hipGraphLaunch");
  }

  void run_factor(const T* vals, const PivotPolicy& pol,
                  std::chrono::steady_clock::time_point t0) {
    const Plan& P = sym->plan;
    hipStream_t st = sym->stream;
    if (!configured) {
      hip_check(configure_factor_kernels<T>(), "configure kernels");
      configured = true;
    }
    valid = false;
    launch_factor(vals, pol, st);
    ++factors_done;
    unsigned long long cnt[5];
    int32_t stat[2];
    hip_check(hipMemcpyAsync(cnt, counters.p, sizeof(cnt), hipMemcpyDeviceToHost, st), "counters");
    hip_check(hipMemcpyAsync(stat, status.p, sizeof(stat), hipMemcpyDeviceToHost, st), "status");
    hip_check(hipStreamSynchronize(st), "factor");
    stats.t_factor_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stat[0] == 2) throw Error(DEVICE_ERROR, "device: inconsistent pivot map in the factor");
    if (stat[0] != 0) {
      // the first zero pivot in (supernode, column) order
      const int32_t s = (int32_t)(cnt[4] >> 32), col = (int32_t)(cnt[4] & 0xffffffffu);
      std::string msg = "numerically rank deficient";
      if (cnt[4] != ~0ull && s >= 0 && s < P.ns)
        msg += ": zero pivot at column " + std::to_string(col) + " of supernode " + std::to_string(s) + " (" +
               std::to_string(P.sn_ncol[s]) + " columns, elimination index " + std::to_string(P.e_off[s] + col) +
               ")";
      throw Error(NUMERICALLY_RANK_DEFICIENT, msg);
    }
    stats.inertia[0] = (int64_t)cnt[0];
    stats.inertia[1] = (int64_t)cnt[1];
    stats.inertia[2] = (int64_t)cnt[2];
    stats.n_perturbed = (int64_t)cnt[3];
    valid = true;
  }

  // the numeric factor's launch sequence on `st` (no host synchronization)
  void enqueue_factor(const T* vals, const PivotPolicy& pol, hipStream_t st, LaunchProfiler* pf) {
    const Plan& P = sym->plan;
    const DevSym& S = sym->view;
    launch_factor_init(counters.as<unsigned long long>(), status.as<int32_t>(), st);
    launch_fill<T>(fv.pool, P.panel_total, F<T>::zero(), st);
    launch_assemble<T>(S, fv, vals, scale_d.as<double>(), st, pf);
    const int nb = P.params.panel_nb < 8 ? 8 : P.params.panel_nb;
    const int prec = static_cast<int>(settings.precision);
    const int32_t* lists = sym->level_lists.as<int32_t>();
    // left-looking: a level's cmod (its targets' updates, pulled) before its
    // cdiv; right-looking: after its cdiv, its supernodes' updates of every
    // ancestor (pushed), the same product blocks grouped by updater level
    const bool rl = P.params.schedule == 1;
    auto cmod_level = [&](const LevelPlan& lv) {
      if (lv.item_end == lv.item_beg) return;
      if constexpr (std::is_same<T, double>::value) {
        if (lv.c_elems <= c_cap) {
          launch_cmod_two_phase(S, fv, sym->items.as<CItem>() + lv.item_beg, (int)(lv.item_end - lv.item_beg),
                                sym->tasks.as<CTask>(), sym->ptiles.as<PTile>() + lv.ptile_beg,
                                lv.ptile_end - lv.ptile_beg, cbuf.p, lv.c_elems, gpool.as<double>(), lv.cmod_flops,
                                prec, st, pf);
          return;
        }
      }
      launch_cmod<T>(S, fv, sym->items.as<CItem>() + lv.item_beg, (int)(lv.item_end - lv.item_beg),
                     sym->tasks.as<CTask>(), lv.cmod_flops, prec, st, pf);
    };
    for (size_t L = 0; L < P.levels.size(); ++L) {
      const LevelPlan& lv = P.levels[L];
      if (!rl) cmod_level(lv);
      // the level's small and wide supernodes are independent: with both
      // present, the one-workgroup cdiv runs on the side stream meanwhile
      const bool fork = !lv.small.empty() && !lv.wide.empty();
      if (fork) {
        hip_check(hipEventRecord(ev_fork, st), "hipEventRecord");
        hip_check(hipStreamWaitEvent(side, ev_fork, 0), "hipStreamWaitEvent");
      }
      launch_cdiv_small<T>(S, fv, lists + sym->level_small_off[L], (int)lv.small.size(), nb, pol,
                           lv.small_flops, prec, fork ? side : st, pf);
      launch_wide_level<T>(S, fv, sym->witems.as<WItem>(), lv.wsteps.data(), (int)lv.wsteps.size(),
                           lists + sym->level_wide_off[L], (int)lv.wide.size(), pol, prec, st, pf);
      if (fork) {
        hip_check(hipEventRecord(ev_join, side), "hipEventRecord");
        hip_check(hipStreamWaitEvent(st, ev_join, 0), "hipStreamWaitEvent");
      }
      if constexpr (std::is_same<T, double>::value) {
        if (c_cap > 0 && L + 1 < P.levels.size()) {
          int maxw = 0;
          for (int32_t sn : lv.all) maxw = std::max(maxw, P.sn_ncol[sn]);
          ProfScope ps(pf, LK_MAKE_G, 0.0, 0.0, st);
          launch_make_g(S, fv, lists + sym->level_all_off[L], (int)lv.all.size(), maxw, gpool.as<double>(), st);
        }
      }
      if (rl) cmod_level(lv);
    }
    launch_rows_e<T>(S, fv, P.rs_off[P.ns], st);
    // drop tolerance at emit (ldlt_emit_and_free, multifrontal_ldlt.rs:2358-2366):
    // after the last consumer of every panel, so it changes the stored L only
    if (settings.drop_tol >= 0.0) launch_drop<T>(S, fv, settings.drop_tol, st);
    hip_check(hipGetLastError(), "factor launch");
  }

  // y (device, e-order) := solve in place; b/x device user-order buffers
  void solve_device(const void* d_b, int64_t nrhs, void* d_x, hipStream_t st) override {
    if (!valid) throw Error(INVALID_INPUT, "the last factorization of this handle failed; refactor first");
    const Plan& P = sym->plan;
    const size_t need = sizeof(T) * std::max<int64_t>(P.n * nrhs, 1);
    if (work_y.bytes < need) work_y.alloc(need);
    const size_t needu = sizeof(T) * std::max<int64_t>(P.u_total * nrhs, 1);
    if (work_b.bytes < needu) work_b.alloc(needu);
    if (work_z.bytes < need) work_z.alloc(need);
    solve_core(d_b, nrhs, d_x, st, work_y.as<T>(), work_b.as<T>(), work_z.as<T>(), prof_on ? &prof : nullptr);
  }

  // Scratch bytes of a solve with nrhs right-hand sides, laid out as
  // y (n x nrhs) | u (off-diagonal panel rows x nrhs) | z (n x nrhs), each
  // part 256-byte aligned.
  void solve_scratch_layout(int64_t nrhs, size_t& ny, size_t& nu) const {
    const Plan& P = sym->plan;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    ny = al(sizeof(T) * (size_t)std::max<int64_t>(P.n * nrhs, 1));
    nu = al(sizeof(T) * (size_t)std::max<int64_t>(P.u_total * nrhs, 1));
  }

  // rsl_solve_device without allocation: the scratch is the library's
  // fixed device array (noalloc_scratch, set up when the factor was made),
  // ordered like solve_async. No hipMalloc, no event creation, no profiling.
  void solve_noalloc(const void* d_b, int64_t nrhs, void* d_x, hipStream_t caller) override {
    if (!valid) throw Error(INVALID_INPUT, "the last factorization of this handle failed; refactor first");
    size_t ny = 0, nu = 0;
    solve_scratch_layout(nrhs, ny, nu);
    const size_t need = 2 * ny + nu;
    if (need > kNoallocScratchBytes)
      throw Error(INVALID_INPUT, "rsl_solve_device_noalloc: this solve needs " + std::to_string(need) +
                                     " bytes of scratch (n = " + std::to_string(sym->plan.n) + ", nrhs = " +
                                     std::to_string(nrhs) + "); the fixed scratch holds " +
                                     std::to_string(kNoallocScratchBytes) + " bytes");
    if (!noalloc_base) throw Error(DEVICE_ERROR, "device: no-allocation scratch is not set up");
    wait_for(caller);
    unsigned char* base = static_cast<unsigned char*>(noalloc_base);
    solve_core(d_b, nrhs, d_x, sym->stream, reinterpret_cast<T*>(base), reinterpret_cast<T*>(base + ny),
               reinterpret_cast<T*>(base + ny + nu), nullptr);
    signal_to(caller);
  }

  // the solve's launch sequence on `st` with caller-provided scratch:
  // y, z hold n x nrhs, u holds u_total x nrhs elements
  void solve_core(const void* d_b, int64_t nrhs, void* d_x, hipStream_t st, T* y, T* u, T* z,
                  LaunchProfiler* pf) {
    const Plan& P = sym->plan;
    const DevSym& S = sym->view;
    const int2* fi = sym->fwd_items.as<int2>();
    const int2* bi = sym->bwd_items.as<int2>();
    const double ts = sizeof(T);
    const SItem* si = sym->sitems.as<SItem>();
    const SItem* bi2 = sym->bitems.as<SItem>();
    const SItem* zi = sym->zitems.as<SItem>();
    const int2* gi = sym->gat_items.as<int2>();
    launch_gather_rhs<T>(P.n, (int32_t)nrhs, fv.perm_e, scale_d.as<double>(),
                         static_cast<const T*>(d_b), y, st, pf);
    const int32_t* lists = sym->level_lists.as<int32_t>();
    auto fwd_level = [&](size_t L) {
      const LevelPlan& lv = P.levels[L];
      launch_fgather<T>(S, fv, gi + lv.gat_beg, (int)(lv.gat_end - lv.gat_beg), (int32_t)nrhs, y, u, st, pf);
      launch_fwd_tri<T>(S, fv, lists + sym->level_snarrow_off[L], (int)lv.snarrow.size(), (int32_t)nrhs, y,
                        ts * (lv.tri_entries + 2.0 * lv.ncols * nrhs), st, pf);
      for (const SolveStep& stp : lv.ssteps) launch_ftri_step<T>(S, fv, si, stp, (int32_t)nrhs, y, z, st, pf);
      launch_zcopy<T>(S, zi + lv.z_beg, (int)(lv.z_end - lv.z_beg), (int32_t)nrhs, z, y, st);
      launch_fwd_update<T>(S, fv, fi + lv.fwd_beg, (int)(lv.fwd_end - lv.fwd_beg), (int32_t)nrhs,
                           y, u, ts * lv.off_entries, st, pf);
    };
    auto bwd_level = [&](size_t L) {
      const LevelPlan& lv = P.levels[L];
      launch_bwd_update<T>(S, fv, bi + lv.bwd_beg, (int)(lv.bwd_end - lv.bwd_beg), (int32_t)nrhs, y,
                           ts * lv.off_entries, st, pf);
      launch_bwd_tri<T>(S, fv, lists + sym->level_snarrow_off[L], (int)lv.snarrow.size(), (int32_t)nrhs, y,
                        ts * (lv.tri_entries + 2.0 * lv.ncols * nrhs), st, pf);
      for (const SolveStep& stp : lv.bsteps) launch_btri_step<T>(S, fv, bi2, stp, (int32_t)nrhs, y, z, st, pf);
      launch_zcopy<T>(S, zi + lv.z_beg, (int)(lv.z_end - lv.z_beg), (int32_t)nrhs, z, y, st);
    };
    // Solves of up to 8 right-hand sides: the levels below the split run one
    // launch per level and block step (many small supernodes, bandwidth
    // work), the levels from the split up run as one dependency-driven launch
    // per sweep (few wide supernodes, chains of dependent blocks). Wider
    // solves run every level as launches, which stream each chunk of
    // right-hand sides through one launch instead of walking every chain once
    // per chunk.
    const size_t split = (P.params.solve_sched == 1 && nrhs <= 8) ? (size_t)P.solve_split : P.levels.size();
    for (size_t L = 0; L < split; ++L) fwd_level(L);
    if (split < P.levels.size()) {
      const double sweep_bytes = ts * (P.dep_tri + P.dep_off + 2.0 * P.n * nrhs);
      launch_sync_init(solve_sync.as<int32_t>(), sym->fu_init.as<int32_t>(), P.ns, st);
      // real field: each supernode's own blocks as one chain unit and its
      // helpers (z is free while the sweeps run)
      constexpr int chain = FieldOf<T>::id == 0 ? 1 : 0;
      launch_fwd_dep<T>(S, fv, chain ? sym->dfwd_chain.as<int2>() : sym->dfwd_items.as<int2>(),
                        (int)P.dfwd_items.size(), sy, (int32_t)nrhs, y, u, z, chain, sweep_bytes, st, pf);
      launch_dsolve<T>(P.n, (int32_t)nrhs, fv.d, fv.ds, fv.two, y, st, pf);
      launch_bwd_dep<T>(S, fv, sym->dbwd_items.as<int2>(), (int)P.dbwd_items.size(), sy, (int32_t)nrhs, y, u,
                        sweep_bytes, st, pf);
    } else {
      launch_dsolve<T>(P.n, (int32_t)nrhs, fv.d, fv.ds, fv.two, y, st, pf);
    }
    for (size_t L = split; L-- > 0;) bwd_level(L);
    launch_scatter_sol<T>(P.n, (int32_t)nrhs, fv.perm_e, scale_d.as<double>(), y,
                          static_cast<T*>(d_x), st, pf);
    hip_check(hipGetLastError(), "solve launch");
  }

  void solve_host(const double* b, int64_t nrhs, double* x) override {
    const Plan& P = sym->plan;
    hipStream_t st = sym->stream;
    const size_t bytes = sizeof(T) * P.n * nrhs;
    grow(hs_db, std::max<size_t>(bytes, 1));
    grow(hs_dx, std::max<size_t>(bytes, 1));
    DBuf &db = hs_db, &dx = hs_dx;
    hip_check(hipMemcpyAsync(db.p, b, bytes, hipMemcpyHostToDevice, st), "upload rhs");
    solve_device(db.p, nrhs, dx.p, st);
    hip_check(hipMemcpyAsync(x, dx.p, bytes, hipMemcpyDeviceToHost, st), "download solution");
    hip_check(hipStreamSynchronize(st), "solve");
    check_solve_waits();
  }

  // A dependency wait of a solve that ran past its bound leaves a count in
  // the sticky control word; the synchronous calls report it (and clear it).
  void check_solve_waits() {
    hipStream_t st = sym->stream;
    uint32_t v = 0;
    hip_check(hipMemcpyAsync(&v, sy.ctl + 3, sizeof(v), hipMemcpyDeviceToHost, st), "solve status");
    hip_check(hipStreamSynchronize(st), "solve status");
    if (v == 0) return;
    hip_check(hipMemsetAsync(sy.ctl + 3, 0, sizeof(v), st), "memset");
    hip_check(hipStreamSynchronize(st), "solve status");
    throw Error(DEVICE_ERROR, "device: a solve's dependency wait timed out; its result is invalid");
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
  void refine_device(const void*, const void*, int64_t, double, void*, double*, hipStream_t) override;
  void refine_core(const T* av, const T* b, int32_t k, double target, double anorm, T* xout, double* info);

  // ---- factor file: everything a solve or a refinement reads ----
  // Layout (little-endian): magic "RSLFACT1", u64 plan digest, i32 field,
  // i32 precision, i32 pivot, i32 scaling, f64 abs_floor, f64 drop_tol,
  // i64 inertia[3], i64 n_perturbed, then the arrays, each as u64 byte count
  // and bytes: scale, panel pool, d, d_sub, 2x2 flags, local pivot order,
  // position of each permuted index, user index of each position, position
  // of each panel row, per-supernode inertia.
  struct Arr {
    DBuf* b;
    size_t bytes;
  };
  std::vector<Arr> file_arrays() {
    const Plan& P = sym->plan;
    const size_t n = (size_t)P.n;
    return {{&scale_d, sizeof(double) * n},
            {&pool, sizeof(T) * (size_t)P.panel_total},
            {&d, sizeof(T) * n},
            {&ds, sizeof(T) * n},
            {&two, n},
            {&lperm, sizeof(int32_t) * n},
            {&e_of_g, sizeof(int32_t) * n},
            {&perm_e, sizeof(int32_t) * n},
            {&rows_e, sizeof(int32_t) * (size_t)P.rs_off[P.ns]},
            {&sn_inert, sizeof(int32_t) * 3 * (size_t)P.ns}};
  }

  void save(std::FILE* fh) override {
    if (!valid) throw Error(INVALID_INPUT, "the last factorization of this handle failed; nothing to save");
    hipStream_t st = sym->stream;
    hip_check(hipStreamSynchronize(st), "save");
    auto put = [&](const void* p, size_t nb) {
      if (nb && std::fwrite(p, 1, nb, fh) != nb) throw Error(INVALID_INPUT, "factor file: write failed");
    };
    put("RSLFACT1", 8);
    const uint64_t dg = sym->digest();
    const int32_t hdr[4] = {FieldOf<T>::id, (int32_t)settings.precision, settings.pivot, settings.scaling};
    put(&dg, 8);
    put(hdr, sizeof(hdr));
    put(&settings.abs_floor, 8);
    put(&settings.drop_tol, 8);
    put(stats.inertia, sizeof(stats.inertia));
    put(&stats.n_perturbed, 8);
    std::vector<unsigned char> host;
    for (const Arr& a : file_arrays()) {
      const uint64_t nb = a.bytes;
      put(&nb, 8);
      host.resize(a.bytes);
      if (a.bytes)
        hip_check(hipMemcpyAsync(host.data(), a.b->p, a.bytes, hipMemcpyDeviceToHost, st), "save download");
      hip_check(hipStreamSynchronize(st), "save download");
      put(host.data(), a.bytes);
    }
  }

  void load(std::FILE* fh) override {
    hipStream_t st = sym->stream;
    auto get = [&](void* p, size_t nb) {
      if (nb && std::fread(p, 1, nb, fh) != nb) throw Error(INVALID_INPUT, "factor file: truncated");
    };
    char magic[8];
    get(magic, 8);
    if (std::memcmp(magic, "RSLFACT1", 8) != 0) throw Error(INVALID_INPUT, "factor file: not a factor file");
    uint64_t dg = 0;
    int32_t hdr[4];
    get(&dg, 8);
    get(hdr, sizeof(hdr));
    if (dg != sym->digest())
      throw Error(INVALID_INPUT, "factor file: written for another symbolic analysis (pattern, ordering or partition)");
    if (hdr[0] != FieldOf<T>::id) throw Error(INVALID_INPUT, "factor file: written for the other field");
    settings.precision = static_cast<Precision>(hdr[1]);
    settings.pivot = hdr[2];
    settings.scaling = hdr[3];
    get(&settings.abs_floor, 8);
    get(&settings.drop_tol, 8);
    get(stats.inertia, sizeof(stats.inertia));
    get(&stats.n_perturbed, 8);
    std::vector<unsigned char> host;
    for (const Arr& a : file_arrays()) {
      uint64_t nb = 0;
      get(&nb, 8);
      if (nb != a.bytes) throw Error(INVALID_INPUT, "factor file: array size does not match the plan");
      host.resize(a.bytes);
      get(host.data(), a.bytes);
      if (a.bytes) hip_check(hipMemcpyAsync(a.b->p, host.data(), a.bytes, hipMemcpyHostToDevice, st), "load upload");
      hip_check(hipStreamSynchronize(st), "load upload");
    }
    stats.t_factor_ms = 0.0;
    valid = true;
  }

  void supernode_inertia(std::vector<int32_t>& out) override {
    const Plan& P = sym->plan;
    out.assign(3 * (size_t)P.ns, 0);
    hip_check(hipMemcpyAsync(out.data(), sn_inert.p, sizeof(int32_t) * 3 * P.ns, hipMemcpyDeviceToHost,
                             sym->stream), "inertia download");
    hip_check(hipStreamSynchronize(sym->stream), "inertia download");
  }
};

// ---------------------------------------------------------------- refinement
// refine_ladder (src/numeric/mixed.rs:140-221) with its constants (:97-105),
// run on all right-hand sides at once: every column follows the reference's
// single-vector ladder (IR with an FP64 residual against the original A and a
// contraction guard, then FGMRES warm-started at the best iterate), the IR
// steps of the columns still iterating share one multi-RHS solve.
namespace {
constexpr int IR_MAX = 8;
constexpr double IR_CONTRACT = 0.5;
constexpr int GMRES_RESTART = 30;
constexpr int GMRES_MAX = 120;

inline double hmag(double v) { return std::fabs(v); }
inline double hmag(cplx v) { return std::hypot(v.re, v.im); }
inline double hconj(double v) { return v; }
inline cplx hconj(cplx v) { return {v.re, -v.im}; }
inline double hreal(double v) { return v; }
inline double hreal(cplx v) { return v.re; }
inline bool hzero(double v) { return v == 0.0; }
inline bool hzero(cplx v) { return v.re == 0.0 && v.im == 0.0; }
template <class T>
T hfrom(double v) {
  return F<T>::from_real(v);
}

inline double bits_to_double(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof(d));
  return d;
}
}  // namespace

template <class T>
void DeviceFactor<T>::refine(const double* a_values, const double* b, int64_t nrhs, double target,
                             double* x, double* info) {
  const Plan& P = sym->plan;
  hipStream_t st = sym->stream;
  const int32_t n = P.n, k = (int32_t)nrhs;
  sym->ensure_full_csr();
  // ||A||_inf of the original matrix (sym_norm_inf, mixed.rs:358-371)
  std::vector<double> rows(n, 0.0);
  for (int32_t j = 0; j < n; ++j)
    for (int64_t q = sym->h_colp[j]; q < sym->h_colp[j + 1]; ++q) {
      const int32_t i = sym->h_rowi[q];
      const double m = FieldOf<T>::id ? std::hypot(a_values[2 * q], a_values[2 * q + 1]) : std::fabs(a_values[q]);
      rows[i] += m;
      if (i != j) rows[j] += m;
    }
  double anorm = 0.0;
  for (double r : rows) anorm = std::max(anorm, r);

  grow(hr_av, sizeof(T) * std::max<int64_t>(P.nnz, 1));
  grow(hr_db, sizeof(T) * std::max<int64_t>((int64_t)n * k, 1));
  grow(hr_dxo, sizeof(T) * std::max<int64_t>((int64_t)n * k, 1));
  DBuf &av = hr_av, &db = hr_db, &dxo = hr_dxo;
  hip_check(hipMemcpyAsync(av.p, a_values, sizeof(T) * P.nnz, hipMemcpyHostToDevice, st), "upload A");
  hip_check(hipMemcpyAsync(db.p, b, sizeof(T) * n * k, hipMemcpyHostToDevice, st), "upload b");
  refine_core(av.as<T>(), db.as<T>(), k, target, anorm, dxo.as<T>(), info);
  hip_check(hipMemcpyAsync(x, dxo.p, sizeof(T) * n * k, hipMemcpyDeviceToHost, st), "download x");
  hip_check(hipStreamSynchronize(st), "refine");
}

// a_values, b and x in device memory: ||A||_inf on the device in the host
// loop's order of accumulation, so the ladder takes the same steps (real
// field; the complex field's modulus is the device's hypot)
template <class T>
void DeviceFactor<T>::refine_device(const void* d_a, const void* d_b, int64_t nrhs, double target, void* d_x,
                                    double* info, hipStream_t caller) {
  {
    const Plan& P = sym->plan;
    hipStream_t st = sym->stream;
    sym->ensure_full_csr();
    wait_for(caller);
    if (dev_u64.bytes < sizeof(unsigned long long) * (size_t)(P.n + 1))
      dev_u64.alloc(sizeof(unsigned long long) * (size_t)(P.n + 1));
    unsigned long long bits = 0;
    launch_row_abs_max(P.n, sym->fa_rp.as<int64_t>(), sym->fa_src.as<int64_t>(), static_cast<const double*>(d_a),
                       FieldOf<T>::id, dev_u64.as<unsigned long long>(), st);
    hip_check(hipMemcpyAsync(&bits, dev_u64.p, sizeof(bits), hipMemcpyDeviceToHost, st), "anorm");
    hip_check(hipStreamSynchronize(st), "anorm");
    double anorm;
    std::memcpy(&anorm, &bits, sizeof(anorm));
    refine_core(static_cast<const T*>(d_a), static_cast<const T*>(d_b), (int32_t)nrhs, target, anorm,
                static_cast<T*>(d_x), info);
    hip_check(hipStreamSynchronize(st), "refine");
  }
}

template <class T>
void DeviceFactor<T>::refine_core(const T* av, const T* dbp, int32_t k, double target, double anorm, T* xout,
                                  double* info) {
  const Plan& P = sym->plan;
  hipStream_t st = sym->stream;
  const int32_t n = P.n;
  sym->ensure_full_csr();
  const size_t blk = sizeof(T) * std::max<int64_t>((int64_t)n * k, 1);
  grow(rf_afull, sizeof(T) * std::max<int64_t>(sym->nnz_full, 1));
  grow(rf_dx, blk);
  grow(rf_dbest, blk);
  grow(rf_dr, blk);
  grow(rf_ddx, blk);
  grow(rf_norms, sizeof(unsigned long long) * 3 * std::max(k, 1));
  grow(rf_mask, std::max(k, 1));
  grow(rf_mask2, std::max(k, 1));
  DBuf &afull = rf_afull, &dx = rf_dx, &dbest = rf_dbest, &dr = rf_dr, &ddx = rf_ddx, &norms = rf_norms,
       &mask_d = rf_mask, &mask2_d = rf_mask2;
  launch_expand<T>(sym->nnz_full, sym->fa_src.as<int64_t>(), av, afull.as<T>(), st);
  const int64_t* rp = sym->fa_rp.as<int64_t>();
  const int32_t* ci = sym->fa_ci.as<int32_t>();
  const T* A = afull.as<T>();
  auto* nrm = norms.as<unsigned long long>();

  // x0 = low-precision solve of b; |b|_inf per column
  solve_device(dbp, k, dx.p, st);
  hip_check(hipMemsetAsync(norms.p, 0, sizeof(unsigned long long) * 3 * k, st), "memset");
  launch_colmax<T>(n, k, dbp, nrm + 2 * k, st);

  std::vector<int> ir(k, 0), phase(k, 0);  // 0 iterating, 1 certified, 2 escalate
  std::vector<double> best_be(k, INFINITY), prev_rn(k, INFINITY), bn(k, 0.0);
  std::vector<unsigned long long> hn(3 * k);
  std::vector<uint8_t> keep(k), step(k);
  bool first = true;
  for (;;) {
    launch_residual<T>(n, k, rp, ci, A, dx.as<T>(), dbp, dr.as<T>(), st, prof_on ? &prof : nullptr);
    hip_check(hipMemsetAsync(norms.p, 0, sizeof(unsigned long long) * 2 * k, st), "memset");
    launch_colmax<T>(n, k, dr.as<T>(), nrm, st);
    launch_colmax<T>(n, k, dx.as<T>(), nrm + k, st);
    hip_check(hipMemcpyAsync(hn.data(), norms.p, sizeof(unsigned long long) * 3 * k, hipMemcpyDeviceToHost, st),
              "norms");
    hip_check(hipStreamSynchronize(st), "refine norms");
    if (first) {
      for (int c = 0; c < k; ++c) bn[c] = bits_to_double(hn[2 * k + c]);
      first = false;
    }
    bool any_step = false;
    for (int c = 0; c < k; ++c) {
      keep[c] = step[c] = 0;
      if (phase[c] != 0) continue;
      const double rn = bits_to_double(hn[c]), xn = bits_to_double(hn[k + c]);
      const double be = rn / std::max(anorm * xn + bn[c], 2.2250738585072014e-308);
      if (be < best_be[c]) {
        best_be[c] = be;
        keep[c] = 1;
      }
      if (be <= target) {
        phase[c] = 1;
        continue;
      }
      if (ir[c] >= IR_MAX || (ir[c] > 0 && rn > prev_rn[c] * IR_CONTRACT)) {
        phase[c] = 2;
        continue;
      }
      prev_rn[c] = rn;
      step[c] = 1;
      any_step = true;
    }
    hip_check(hipMemcpyAsync(mask_d.p, keep.data(), k, hipMemcpyHostToDevice, st), "mask");
    launch_copy_cols<T>(n, k, mask_d.as<uint8_t>(), dbest.as<T>(), dx.as<T>(), st);
    if (!any_step) break;
    hip_check(hipMemcpyAsync(mask2_d.p, step.data(), k, hipMemcpyHostToDevice, st), "mask");
    solve_device(dr.p, k, ddx.p, st);
    launch_update_cols<T>(n, k, mask2_d.as<uint8_t>(), dx.as<T>(), ddx.as<T>(), st);
    for (int c = 0; c < k; ++c) ir[c] += step[c];
    // the host vectors keep/step are rewritten next pass: wait for their copies
    hip_check(hipStreamSynchronize(st), "refine step");
  }

  // GMRES-IR (mixed.rs:195-219): FGMRES with this factor as preconditioner,
  // warm-started at each escalated column's best iterate, the escalated
  // columns advancing together; a column keeps the GMRES result when its
  // backward error is smaller
  std::vector<int> gm_iters(k, 0);
  std::vector<int> esc;
  for (int c = 0; c < k; ++c)
    if (phase[c] == 2) esc.push_back(c);
  if (!esc.empty()) {
    const int32_t ke = (int32_t)esc.size();
    const size_t eb = sizeof(T) * (size_t)n * ke;
    grow(gm_bv, eb);
    grow(gm_xv, eb);
    grow(gm_rv, eb);
    T *be_ = gm_bv.as<T>(), *xe = gm_xv.as<T>(), *re_ = gm_rv.as<T>();
    for (int q = 0; q < ke; ++q) {
      launch_copy_col<T>(n, k, esc[q], dbp, ke, q, be_, st);
      launch_copy_col<T>(n, k, esc[q], dbest.as<T>(), ke, q, xe, st);
    }
    bind_krylov(A, true);
    std::vector<KrylovCol> res;
    kry.fgmres(be_, xe, ke, target, GMRES_MAX, GMRES_RESTART, true, res);
    // backward error of each GMRES result
    launch_residual<T>(n, ke, rp, ci, A, xe, be_, re_, st, prof_on ? &prof : nullptr);
    grow(gm_nrm, sizeof(unsigned long long) * 2 * ke);
    auto* nr = gm_nrm.as<unsigned long long>();
    hip_check(hipMemsetAsync(nr, 0, sizeof(unsigned long long) * 2 * ke, st), "memset");
    launch_colmax<T>(n, ke, re_, nr, st);
    launch_colmax<T>(n, ke, xe, nr + ke, st);
    std::vector<unsigned long long> hb(2 * ke);
    hip_check(hipMemcpyAsync(hb.data(), nr, sizeof(unsigned long long) * 2 * ke, hipMemcpyDeviceToHost, st), "norms");
    hip_check(hipStreamSynchronize(st), "gmres norms");
    for (int q = 0; q < ke; ++q) {
      const int c = esc[q];
      gm_iters[c] = res[q].iters;
      const double be2 = bits_to_double(hb[q]) /
                         std::max(anorm * bits_to_double(hb[ke + q]) + bn[c], 2.2250738585072014e-308);
      if (be2 < best_be[c]) {
        best_be[c] = be2;
        launch_copy_col<T>(n, ke, q, xe, k, c, dbest.as<T>(), st);
      }
    }
  }
  hip_check(hipMemcpyAsync(xout, dbest.p, sizeof(T) * n * k, hipMemcpyDeviceToDevice, st), "copy x");
  check_solve_waits();
  for (int c = 0; c < k; ++c) {
    info[4 * c + 0] = ir[c];
    info[4 * c + 1] = gm_iters[c];
    info[4 * c + 2] = best_be[c];
    info[4 * c + 3] = best_be[c] <= target ? 1.0 : 0.0;
  }
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

uint64_t DeviceSymbolic::digest() const {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void* p, size_t nb) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < nb; ++i) {
      h ^= c[i];
      h *= 1099511628211ull;
    }
  };
  const Plan& P = plan;
  const int64_t dims[4] = {P.n, P.ns, P.panel_total, P.nnz};
  mix(dims, sizeof(dims));
  mix(h_colp.data(), sizeof(int64_t) * h_colp.size());
  mix(h_rowi.data(), sizeof(int32_t) * h_rowi.size());
  mix(P.sym.perm.data(), sizeof(int32_t) * P.sym.perm.size());
  mix(P.sn_first.data(), sizeof(int32_t) * P.sn_first.size());
  mix(P.sn_ncol.data(), sizeof(int32_t) * P.sn_ncol.size());
  mix(P.rs_idx.data(), sizeof(int32_t) * P.rs_idx.size());
  mix(P.panel_off.data(), sizeof(int64_t) * P.panel_off.size());
  return h;
}

std::unique_ptr<DeviceFactorBase> load_factor(std::shared_ptr<DeviceSymbolic> sym, const char* path) {
  std::FILE* fh = std::fopen(path, "rb");
  if (!fh) throw Error(INVALID_INPUT, std::string("factor file: cannot open ") + path);
  struct Closer {
    std::FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{fh};
  char magic[8];
  uint64_t dg = 0;
  int32_t hdr[4];
  if (std::fread(magic, 1, 8, fh) != 8 || std::fread(&dg, 1, 8, fh) != 8 || std::fread(hdr, 1, sizeof(hdr), fh) != sizeof(hdr))
    throw Error(INVALID_INPUT, "factor file: truncated header");
  if (std::memcmp(magic, "RSLFACT1", 8) != 0) throw Error(INVALID_INPUT, "factor file: not a factor file");
  if (hdr[1] < 0 || hdr[1] > 2) throw Error(INVALID_INPUT, "factor file: unknown precision");
  FactorSettings fs;
  fs.precision = static_cast<Precision>(hdr[1]);
  fs.pivot = hdr[2];
  fs.scaling = hdr[3];
  sym->upload();
  std::unique_ptr<DeviceFactorBase> f;
  if (hdr[0] == 0) f.reset(new DeviceFactor<double>(sym, fs));
  else if (hdr[0] == 1) f.reset(new DeviceFactor<cplx>(sym, fs));
  else throw Error(INVALID_INPUT, "factor file: unknown field");
  std::rewind(fh);
  f->load(fh);
  return f;
}

}  // namespace rsl
