// Device execution plan of the LU twin, built once per analysed pattern
// (value independent; reused across refactorizations, LuSymbolic::factor,
// src/numeric/multifrontal_lu.rs:880-907).
#pragma once
#include <cstdint>
#include <vector>

#include "../device/lu_types.hpp"
#include "symbolic.hpp"

namespace rsl {

struct LuStepPlan {
  int64_t g_beg = 0, g_end = 0;  // getf2 units
  int64_t a_beg = 0, a_end = 0;  // apply units (columns, deep rows)
  int64_t m_beg = 0, m_end = 0;  // trailing GEMM tiles
  double gemm_flops = 0;
};

struct LuLevelPlan {
  std::vector<int32_t> all;        // supernodes of the level
  int64_t item_beg = 0, item_end = 0;  // cmod target tiles
  double cmod_flops = 0;
  std::vector<LuStepPlan> steps;   // one per panel index
};

struct LuPlan {
  Symbolic sym;  // analysis of the symmetrized pattern A + A^T
  int32_t n = 0, ns = 0;
  int64_t nnz = 0;
  std::vector<int64_t> colp;  // the full (unsymmetric) user pattern
  std::vector<int32_t> rowi;
  std::vector<int32_t> sn_first, sn_ncol, sn_nrow, nb;  // nb: LU_PANEL
  std::vector<int64_t> rs_off;
  std::vector<int32_t> rs_idx;
  std::vector<int64_t> l_off, u_off, s_off;
  int64_t l_total = 0, u_total = 0, s_total = 0;
  std::vector<int64_t> amap;            // user nnz -> L slot, or -1 - U slot
  std::vector<int64_t> csr_ptr, csr_src;  // CSR of A (residuals): row pointers, CSC index of each entry
  std::vector<int32_t> csr_col;
  std::vector<int64_t> gat_ptr, gat_idx;
  std::vector<int32_t> rperm_init;      // identity own-row positions (first + p -> p)
  std::vector<LuLevelPlan> levels;
  std::vector<int32_t> level_lists;     // the levels' supernodes, level after level
  std::vector<int64_t> level_off;
  std::vector<LuItem> items;
  std::vector<CTask> tasks;
  std::vector<LuUnit> units;
  double factor_flops = 0;              // sum_s sum_{k<w} 2 (h-k-1)^2 + (h-k-1) (real field)
  int64_t l_entries = 0, u_entries = 0; // stored L (incl. diagonal) and U (incl. diagonal) of the supernodes
};

// Lower triangle of the pattern A + A^T (symmetrized_lower_pattern,
// multifrontal_lu.rs:364-409).
void symmetrized_lower(int32_t n, const int64_t* colp, const int32_t* rowi, std::vector<int64_t>& scp,
                       std::vector<int32_t>& sri);

// The plan of P.sym (already analysed) for the user pattern colp / rowi.
void build_lu_plan(LuPlan& P, const int64_t* colp, const int32_t* rowi);

}  // namespace rsl
