// Device execution plan built once per analysed pattern (value independent).
#pragma once
#include <cstdint>
#include <vector>

#include "../device/devtypes.hpp"
#include "symbolic.hpp"

namespace rsl {

struct PlanParams {
  int panel_nb = 64;            // SolverSettings::panel_nb default (gemm_tuning.rs:27)
  double wide_flops = 2.5e5;    // supernodes with h*w*w above this get multi-workgroup cdiv
  int solve_wide = 32;          // own blocks wider than this are solved in 64-column steps
  int schedule = 0;             // 0 left-looking (cmod pulled at the target's level),
                                // 1 right-looking (pushed after the updater's level)
  int solve_sched = 1;          // 1 dependency-driven solve above the split level,
                                // 0 level-by-level launches throughout
  int solve_split = -1;         // first level of the dependency-driven launch (< 0: automatic)
};

struct LevelPlan {
  std::vector<int32_t> all;     // supernodes of the level (solve)
  std::vector<int32_t> small;   // one workgroup per supernode (cdiv)
  std::vector<int32_t> wide;    // multi-launch panel path (cdiv)
  int64_t item_beg = 0, item_end = 0;  // cmod items of the level
  int64_t ptile_beg = 0, ptile_end = 0;  // two-phase cmod: product tiles of the level
  int64_t c_elems = 0;                   // product-buffer elements the level needs
  int64_t fwd_beg = 0, fwd_end = 0;    // solve: (supernode, row chunk) L21 y items
  int64_t bwd_beg = 0, bwd_end = 0;    // solve: (supernode, column chunk) L21^T x items
  int64_t gat_beg = 0, gat_end = 0;    // solve: (supernode, own-row chunk) gather items
  std::vector<int32_t> snarrow, swide; // solve: own block in one workgroup / in block steps
  std::vector<SolveStep> ssteps;       // solve: forward block steps of the wide ones
  std::vector<SolveStep> bsteps;       // solve: backward block steps (last block first)
  int64_t z_beg = 0, z_end = 0;        // solve: every own block of the wide ones (copy-back units)
  // algorithmic work of the level (real field; 2 flops per multiply-add)
  double cmod_flops = 0;   // sum over cmod tasks of 2 * nck * (lower-trapezoid pairs)
  double small_flops = 0;  // cdiv of the small supernodes: sum_k (h-k)^2 - (h-w)^2
  double tri_entries = 0;  // own-block entries sum w(w+1)/2 (solve)
  double off_entries = 0;  // off-diagonal entries sum (h-w) w (solve)
  int64_t ncols = 0;       // columns eliminated in the level
  std::vector<WideStep> wsteps;
};

struct Int2 {
  int32_t x, y;
};


struct Plan {
  Symbolic sym;
  PlanParams params;
  int32_t n = 0, ns = 0;
  std::vector<int32_t> sn_first, sn_ncol, sn_nrow, e_off, nb_eff;
  std::vector<int64_t> panel_off, rs_off;
  std::vector<int64_t> g_off;   // offset of supernode s's L21 D block in the G pool ((h - w) x w)
  int64_t g_total = 0;
  std::vector<int32_t> rs_idx;
  std::vector<int32_t> upd_ptr, upd_idx;
  std::vector<int32_t> level;   // level of each supernode (leaves 0)
  std::vector<LevelPlan> levels;
  std::vector<CItem> items;
  std::vector<CTask> tasks;
  std::vector<PTile> ptiles;
  int64_t c_elems_max = 0;      // largest c_elems over the levels
  std::vector<WItem> witems;    // batched wide-path work units
  // solve plan
  std::vector<int64_t> u_off;   // scratch rows of supernode s (h - w of them)
  int64_t u_total = 0;
  std::vector<int64_t> gat_ptr, gat_idx;  // per permuted column, ascending updater
  std::vector<Int2> fwd_items, bwd_items, gat_items;
  std::vector<SItem> sitems;
  std::vector<SItem> bitems;    // backward step units: (s, j0, i0) pushes of block j0 into block i0
  std::vector<SItem> zitems;    // (s, j0) own blocks of the wide supernodes
  // dependency-driven solve: work units in the order they are taken
  std::vector<Int2> dfwd_items;  // (s, c): own 64-row blocks c < nblk of a level's supernodes,
                                 // then their off-diagonal 64-row chunks (c >= nblk); levels ascending
  std::vector<Int2> dfwd_chain;  // real field: (s, -1) chain unit, (s, c) helpers 1 <= c < nblk,
                                 // per supernode; then the off-diagonal chunks; levels ascending
  std::vector<Int2> dbwd_items;  // (s, -1 - (q + ng r)): off-diagonal rows of chunk r
                                 // (bwd_row_chunks) for own columns 16q..16q+15,
                                 // then (s, c): own 64-column blocks, last first; levels descending
  std::vector<int32_t> anc_ptr, anc_idx;  // distinct supernodes owning s's off-diagonal rows
  int32_t solve_split = 0;                // levels below run as launches, from it up as units
  std::vector<int32_t> fu_init;           // off-diagonal chunk counter at solve start (full below the split)
  double dep_tri = 0, dep_off = 0;        // own-block lower / off-diagonal entries from the split up
  int64_t panel_total = 0;      // elements in the panel pool
  // user lower-CSC entries -> panel slot
  int64_t nnz = 0;
  std::vector<int64_t> amap;
  std::vector<int32_t> a_row, a_col;
  // work figures (real-field flop convention; x4 for complex)
  double factor_flops = 0;      // sum_s sum_{k<w} (h-k)^2
  double geom_flops = 0;        // sum_s h^2 w (reference proxy)
  int64_t nnzl_panel = 0;       // sum_s sum_{k<w} (h-k): stored L incl. diagonal
};

// Build the plan; colp/rowi is the user lower CSC the symbolic came from.
void build_plan(Plan& P, const int64_t* colp, const int32_t* rowi, const PlanParams& prm);

}  // namespace rsl
