// Plain-old-data views of the LU twin's device plan and factor storage,
// passed to kernels by value (shared by the HIP translation units and the
// host runtime).
#pragma once
#include <cstdint>

#include "devtypes.hpp"

namespace rsl {

struct LuSymDev {
  int32_t n, ns;
  const int32_t* sn_first;   // first column (permuted numbering)
  const int32_t* sn_ncol;    // own columns w
  const int32_t* sn_nrow;    // rows h = |rs[s]| (own rows first)
  const int32_t* nb;         // device panel width of s
  const int64_t* rs_off;     // row structures
  const int32_t* rs_idx;
  const int64_t* l_off;      // h x w column-major: L below the diagonal, U11 on and above
  const int64_t* u_off;      // w x (h - w) column-major: U12
  const int64_t* s_off;      // solve scratch: the h - w off-diagonal rows of s
  const int64_t* gat_ptr;    // per permuted row: scratch rows that update it (ascending updater)
  const int64_t* gat_idx;
  const int32_t* perm;       // permuted index -> original index
  const int64_t* amap;       // original nnz -> L slot (>= 0) or U slot (-1 - slot)
  int64_t nnz;
};

template <class T>
struct LuFacDev {
  T* l;
  T* u;
  int32_t* rperm;   // own row position p of s (at first + p): its row-structure index
  int32_t* ipiv;    // at first + k: the row interchanged with row k at step k (k if none)
  int32_t* status;  // [0] 1 when a pivot is exactly zero and zero pivots fail
  unsigned long long* counters;  // [0] perturbed pivots [1] first zero pivot (permuted column)
  double thresh_sq;              // pivot threshold u squared
  PivotPolicy pol;
};

// cmod work item: a 64x64 tile of target s's L panel (region 0, rows over h,
// columns over w) or of its U12 block (region 1, rows over w, columns over
// h - w) and its task range (CTask: A rows from the updater's L, B columns
// from its U12).
struct LuItem {
  int32_t s, r0, c0, t_beg, t_end, region;
};

// Panel-step work unit of supernode s for the panel starting at column kb:
// getf2 (a, b unused); apply: b = 0 a column unit (first column a of the
// columns outside the panel: interchanges, then the triangular solve for the
// later ones), b = 1 a deep-row unit (first deep row a); GEMM: tile (a, b) of
// the trailing L block, or (-1 - a, b) of U12.
struct LuUnit {
  int32_t s, kb, a, b;
};

constexpr int LU_NT = 256;     // threads per workgroup of the LU kernels but getf2
constexpr int LU_GT = 1024;    // threads of the getf2 workgroup (one per supernode panel)
constexpr int LU_COLS = 256;   // columns (rows) per apply unit
constexpr int LU_PANEL = 16;   // device panel width of the LU cdiv

}  // namespace rsl
