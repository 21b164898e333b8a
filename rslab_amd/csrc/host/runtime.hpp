// Host runtime: device plan upload, numeric factor and solve orchestration.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../device/devtypes.hpp"
#include "../device/prof.hpp"
#include "../device/scalar.hpp"
#include "plan.hpp"

namespace rsl {

// Status codes shared with the C ABI (rslab_amd.h).
enum Status : int {
  OK = 0,
  NUMERICALLY_RANK_DEFICIENT = 1,
  INVALID_INPUT = 2,
  DIMENSION_MISMATCH = 3,
  DEVICE_ERROR = 4,
  OUT_OF_MEMORY = 5,
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);

// Owning device buffer.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf();
  void alloc(size_t nbytes);
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
  template <class T>
  void upload(const std::vector<T>& v, hipStream_t st) {
    alloc(v.size() * sizeof(T));
    if (!v.empty()) hip_check(hipMemcpyAsync(p, v.data(), bytes, hipMemcpyHostToDevice, st), "upload");
  }
};

enum class Precision : int { FP64 = 0, FP32_SCHUR = 1, BF16_SCHUR = 2 };

struct FactorSettings {
  int pivot = 0;  // 0 Fail, 1 PerturbToEps, 2 ForceAccept
  double abs_floor = 0.0;
  int scaling = 1;  // 1 OnePassInfNorm, 0 Identity
  Precision precision = Precision::FP64;
  double drop_tol = -1.0;  // < 0: None
  int64_t cmod_cap = -1;   // product-buffer element cap, < 0 automatic
  bool graphs = true;      // replay refactors as a captured HIP graph
};

// Device-resident plan of one analysed pattern.
struct DeviceSymbolic {
  Plan plan;
  hipStream_t stream = nullptr;
  DBuf sn_first, sn_ncol, sn_nrow, e_off, panel_off, rs_off, rs_idx, upd_ptr, upd_idx, sym_perm,
      amap, a_row, a_col, items, tasks, level_lists, u_off, gat_ptr, gat_idx,
      witems, nb_eff, gat_items, sitems,
      fwd_items, bwd_items, ptiles, bitems, zitems, g_off, dfwd_items, dfwd_chain, dbwd_items, anc_ptr, anc_idx, fu_init;
  std::vector<int64_t> level_all_off, level_small_off, level_wide_off, level_snarrow_off;  // into level_lists
  // user lower CSC pattern (host) and, built on first refinement, the full
  // symmetric CSR of A with each entry's source index in the CSC values
  std::vector<int64_t> h_colp;
  std::vector<int32_t> h_rowi;
  std::vector<int64_t> batch_off;        // first unknown of each system of a batch (count+1), or empty
  std::vector<int64_t> batch_snode_off;  // first supernode of each system of a batch (count+1)
  DBuf fa_rp, fa_ci, fa_src;
  int64_t nnz_full = 0;
  bool full_ready = false;
  DevSym view{};
  bool uploaded = false;
  ~DeviceSymbolic();
  void upload();
  void ensure_full_csr();
  // FNV-1a digest of the plan's structure (dimension, pattern, permutation,
  // partition, row structures, panel layout): two plans with equal digests
  // lay out a factor identically
  uint64_t digest() const;
};

// Per-launch event timing, attached to the launchers while enabled.
struct EventProfiler final : LaunchProfiler {
  struct Rec {
    int kind;
    double flops, bytes;
    hipEvent_t a, b;
  };
  struct Acc {
    double ms = 0, flops = 0, bytes = 0;
    int64_t launches = 0;
  };
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<Rec> recs;
  Acc acc[LK_COUNT];
  double flop_mult = 1.0;  // 4 for the complex field (4 real products per complex one)
  ~EventProfiler() override;
  void begin(int kind, double flops, double bytes, hipStream_t st) override;
  void end(hipStream_t st) override;
  void collect();  // waits for the recorded events, folds them into acc
  void reset();
 private:
  hipEvent_t take();
};
const char* launch_kind_name(int k);

struct FactorStats {
  double t_factor_ms = 0;
  int64_t inertia[3] = {0, 0, 0};
  int64_t n_perturbed = 0;
};

class DeviceFactorBase {
 public:
  virtual ~DeviceFactorBase() = default;
  virtual int field() const = 0;  // 0 f64, 1 c64
  virtual void refactor(const double* values) = 0;
  virtual void solve_host(const double* b, int64_t nrhs, double* x) = 0;
  virtual void solve_device(const void* d_b, int64_t nrhs, void* d_x, hipStream_t st) = 0;
  // solve_device on the library stream, ordered after and before `caller`'s work
  virtual void solve_async(const void* d_b, int64_t nrhs, void* d_x, hipStream_t caller) = 0;
  // solve_async with the library's fixed scratch array: allocates nothing
  virtual void solve_noalloc(const void* d_b, int64_t nrhs, void* d_x, hipStream_t caller) = 0;
  virtual void download(std::vector<int64_t>& colp, std::vector<int32_t>& rowi,
                        std::vector<double>& lval, std::vector<double>& d,
                        std::vector<double>& ds, std::vector<uint8_t>& two,
                        std::vector<int32_t>& perm) = 0;
  // Mixed-precision refinement ladder (IR -> GMRES-IR) against the original
  // matrix values; b, x row-major n x nrhs host buffers. info: per RHS
  // {ir_iters, gmres_iters, backward_error, certified}.
  virtual void refine(const double* a_values, const double* b, int64_t nrhs, double target,
                      double* x, double* info) = 0;
  // device-resident variants (real field): inputs in device memory, ordered
  // after the work already enqueued on `caller`; the calls return when done
  virtual void refactor_device(const void* d_values, hipStream_t caller) = 0;
  virtual void refine_device(const void* d_a_values, const void* d_b, int64_t nrhs, double target, void* d_x,
                             double* info, hipStream_t caller) = 0;
  // per-supernode (positive, negative, zero) pivot counts of the last factor
  virtual void supernode_inertia(std::vector<int32_t>& out) = 0;
  // Krylov solvers with A (values in the plan's pattern, host) as operator
  // and this factor (precond != 0) or the identity as preconditioner; b, x0,
  // x host row-major n x nrhs; method 0 FGMRES, 1 COCG, 2 COCR; info per
  // column {iterations, final relative residual, converged, stop reason}
  virtual void krylov(int method, const double* a_values, const double* b, int64_t nrhs, double tol,
                      int max_iter, int restart, int precond, const double* x0, double* x, double* info) = 0;
  // factor file (a valid factor's device state, tied to the plan by a digest)
  virtual void save(std::FILE* fh) = 0;
  virtual void load(std::FILE* fh) = 0;
  FactorStats stats;
  FactorSettings settings;
  EventProfiler prof;
  bool prof_on = false;
  bool valid = false;  // the last numeric factorization succeeded
  std::shared_ptr<DeviceSymbolic> sym;
};

std::unique_ptr<DeviceFactorBase> make_factor(std::shared_ptr<DeviceSymbolic> sym, int field,
                                              const double* values, const FactorSettings& fs);
// A factor read from a file written by DeviceFactorBase::save for the same
// plan (no numeric factorization runs).
std::unique_ptr<DeviceFactorBase> load_factor(std::shared_ptr<DeviceSymbolic> sym, const char* path);

// onepass_scale (src/numeric/sparse_solver.rs:399-418) on the host
void onepass_scale(int32_t n, int field, const int64_t* colp, const int32_t* rowi,
                   const double* val, std::vector<double>& scale);

}  // namespace rsl
