// Plain-old-data views of the device plan and factor, passed to kernels by
// value. Shared by the HIP translation units and the host runtime.
#pragma once
#include <cstdint>

namespace rsl {

// Backward dependency-driven solve: the off-diagonal terms of a supernode
// with w own columns and `off` off-diagonal rows run as row chunks of
// BWD_CHUNK rows, each chunk's partial sums (w x nrhs) stored in the
// supernode's forward scratch rows (off x nrhs, no longer read), when they
// fit; otherwise one chunk that subtracts its sums from the solution itself.
constexpr int BWD_CHUNK = 1024;
constexpr int bwd_row_chunks(int32_t w, int64_t off) {
  return ((off + BWD_CHUNK - 1) / BWD_CHUNK > 1 && (int64_t)w * ((off + BWD_CHUNK - 1) / BWD_CHUNK) <= off)
             ? (int)((off + BWD_CHUNK - 1) / BWD_CHUNK)
             : 1;
}

// Value-free per-pattern arrays (uploaded once per analysis).
struct DevSym {
  int32_t n, ns;
  const int32_t* sn_first;   // first column (permuted numbering)
  const int32_t* sn_ncol;    // eliminated columns w
  const int32_t* sn_nrow;    // panel height h = |rs[s]|
  const int32_t* e_off;      // first elimination position of s
  const int64_t* panel_off;  // panel offset in the value pool (column-major h x w)
  const int64_t* g_off;      // offset of L21 D in the G pool ((h - w) x w, column-major)
  const int64_t* rs_off;     // offset of rs[s] in rs_idx
  const int32_t* rs_idx;     // row structures (global permuted indices)
  const int32_t* upd_ptr;    // update lists (ns+1)
  const int32_t* upd_idx;
  const int32_t* sym_perm;   // permuted -> user index
  const int64_t* u_off;      // solve scratch row offset of supernode s (h - w rows)
  const int64_t* gat_ptr;    // per permuted column: scratch rows to gather (ascending updater)
  const int64_t* gat_idx;
  const int64_t* amap;       // user nnz -> panel offset
  const int32_t* a_row;      // user nnz -> user row
  const int32_t* a_col;      // user nnz -> user column
  const int32_t* nb_eff;     // panel width of s (64, or 128 on supernodes >= 512 wide)
  const int32_t* anc_ptr;    // distinct supernodes owning s's off-diagonal rows (ns+1)
  const int32_t* anc_idx;
  int64_t nnz;
};

// Dependency-driven solve: per-supernode progress counters, cleared before
// every solve, and control words. A unit waits on counters with bounded
// spins; a wait that exceeds its bound sets abort (every later wait returns
// at once, so the launch always drains) and counts itself in ctl[3].
struct SolveSync {
  int32_t* cnt;    // supernode s at cnt[32 s + kind] (one 128-byte line each): kind 0 forward
                   // own 64-blocks solved (in block order), 1 forward off-diagonal 64-row
                   // chunks written to the scratch, 2 backward own blocks solved (from the
                   // last), 3 backward 16-column off-diagonal units done
  uint32_t* ctl;   // [0] forward ticket [1] backward ticket [2] abort [3] timed-out waits (sticky)
};

// Numeric factor storage for field T.
template <class T>
struct DevFac {
  T* pool;            // all panels
  T* d;               // D diagonal, elimination order
  T* ds;              // D sub-diagonal (first column of a 2x2 block)
  uint8_t* two;       // first column of a 2x2 block
  int32_t* lperm;     // pivoted position -> local column, elimination order
  int32_t* e_of_g;    // permuted index -> elimination position
  int32_t* rs_e;      // per panel row (rs_idx order): its elimination position
  int32_t* perm_e;    // elimination position -> user index
  int32_t* deep;      // per-panel interchange partners (scratch, nb per supernode)
  unsigned long long* counters;  // [0] pos [1] neg [2] zero [3] perturbed [4] first failure (s << 32 | column)
  int32_t* sn_inertia;          // per supernode: pos, neg, zero
  int32_t* status;    // [0] 0 ok / 1 rank deficient, [1] failing supernode
};

// One cmod work item: a 64x64 tile of target panel s and its task range.
struct CItem {
  int32_t s, r0, c0, t_beg, t_end;
};
// One contribution of updater k to a tile: off-diagonal row ranges [a0,a1)
// (tile rows) and [b0,b1) (tile columns), indices into k's off-diagonal rows.
// When the level runs the two-phase cmod, entry (a, b) of the updater's
// product block for this target sits at cbuf[cbase + a * ldc + b].
struct CTask {
  int32_t k, a0, a1, b0, b1, ldc;
  uint64_t rmask;  // tile rows hit by rows a0..a1-1 (in order)
  uint64_t cmask;  // tile columns hit by rows b0..b1-1 (in order)
  int64_t cbase;
};

// One 64x64 tile of a product block  C = L_k[p0.., :] D_k L_k[p0..p0+N, :]^T
// (rows p0..p0+M of k's off-diagonal rows against the N of them that fall in
// the target's columns), stored row-major with leading dimension N from
// element c_off of the level's product buffer.
struct PTile {
  int32_t k, p0, m, n, mt, nt;
  int64_t c_off;
};

// One unit of the batched wide-supernode path: nb-panel [kb, kb+nb) of
// supernode s; a = deep-row offset (replay) or row tile (Schur), b = column tile.
struct WItem {
  int32_t s, kb, a, b;
};

// One panel step of a level's wide supernodes, run as three batched launches
// (getf2 on every supernode that has a p-th panel, then the deep-row replay,
// then the deferred Schur tiles). Ranges index the WItem array.
// Deep rows per replay work unit of the wide path: four waves, one 16-row
// MFMA tile each, with one row per lane of wave 0 in the in-pass triangle
// (the unit's time is bound by its per-row dependence chains, so small units
// spread a step's rows over more CUs).
constexpr int REPLAY_ROWS = 64;

struct WideStep {
  int64_t g_beg = 0, g_end = 0, r_beg = 0, r_end = 0, t_beg = 0, t_end = 0;
  double getf2_flops = 0, replay_flops = 0, schur_flops = 0;
  int32_t max_kb = 0;  // largest panel start of the step (columns whose rows are interchanged)
};

// Solve work unit on a wide supernode's own block, for the 64-column block
// step j0 of supernode s: a forward row unit has r0 = the first own row of a
// 64-row chunk below the block, a backward push unit r0 = the first column of
// an earlier block; a diagonal unit (which stores the solved block) r0 = -1.
struct SItem {
  int32_t s, j0, r0, x;
};

// One 64-column block step of a level's wide supernodes in the solve:
// diagonal units [d_beg, d_end), update units [u_beg, u_end) (of sitems).
struct SolveStep {
  int64_t d_beg = 0, d_end = 0, u_beg = 0, u_end = 0;
};

struct PivotPolicy {
  int32_t have_floor;
  double floor;
};

}  // namespace rsl
