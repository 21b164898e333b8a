// Host runtime of the LU twin: device plan upload, numeric factor and solve
// orchestration, and the host emit of the global L (CSC) / U (CSR).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "lu_plan.hpp"
#include "runtime.hpp"

namespace rsl {

struct LuDeviceSymbolic {
  LuPlan plan;
  hipStream_t stream = nullptr;
  DBuf sn_first, sn_ncol, sn_nrow, nb, rs_off, rs_idx, l_off, u_off, s_off, gat_ptr, gat_idx, perm, amap,
      level_lists, items, tasks, units, rperm_init, csr_ptr, csr_col, csr_src;
  LuSymDev view{};
  bool uploaded = false;
  ~LuDeviceSymbolic();
  void upload();
};

struct LuFactorSettings {
  int pivot = 0;           // 0 Fail, 1 PerturbToEps, 2 ForceAccept
  double abs_floor = 0.0;
  double pivot_u = 0.1;    // threshold, clamped to [0, 1]
  double drop_tol = -1.0;  // < 0: None (applied at emit)
};

// LuFactors (multifrontal_lu.rs:321-352): e-ordered L CSC / U CSR,
// permutations, equilibration.
struct LuEmitted {
  std::vector<int64_t> l_col_ptr, u_row_ptr;
  std::vector<int32_t> l_row_idx, u_col_idx;
  std::vector<double> l_values, u_values;  // interleaved re/im for the complex field
  std::vector<int32_t> perm, perm_row;
};

class LuFactorBase {
 public:
  virtual ~LuFactorBase() = default;
  virtual int field() const = 0;  // 0 f64, 1 c64
  virtual void refactor(const double* values) = 0;
  virtual void solve_host(const double* b, int64_t nrhs, double* x) = 0;
  // b, x device buffers (row-major n x nrhs); ordered after `caller`'s work,
  // and `caller` waits for the solve
  virtual void solve_device(const void* d_b, int64_t nrhs, void* d_x, hipStream_t caller) = 0;
  virtual const LuEmitted& emitted() = 0;
  // solve_lu_refined (multifrontal_lu.rs:3285-3322) per column, on the
  // device: x += (LU)^-1 (b - A x) with the residual against a_values (the
  // analysed pattern's values, host), keeping each column's iterate of least
  // max |r|; info[2c..2c+2) = {corrections applied, that least max |r|}
  virtual void refine(const double* a_values, const double* b, int64_t nrhs, int max_iter, double* x,
                      double* info) = 0;
  std::vector<double> d_row, d_col;
  int64_t n_perturbed = 0;
  double t_factor_ms = 0;
  LuFactorSettings settings;
  std::shared_ptr<LuDeviceSymbolic> sym;
};

std::unique_ptr<LuFactorBase> make_lu_factor(std::shared_ptr<LuDeviceSymbolic> sym, int field,
                                             const double* values, const LuFactorSettings& fs);

}  // namespace rsl
