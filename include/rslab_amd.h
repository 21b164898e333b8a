/*
 * rslab_amd -- MI355X (gfx950) numeric factor and solve for rslab's
 * supernodal left-looking LDL^T with Bunch-Kaufman pivoting.
 *
 * C ABI: plain pointers and sizes, no exceptions cross it, every entry point
 * returns an int status (RSL_OK = 0). Input arrays are borrowed for the
 * duration of the call; the library owns all device memory; handles are freed
 * explicitly. A message for the last failure on the calling thread is
 * available from rsl_last_error().
 *
 * The rslab (Rust) entry points each function stands in for are cited as
 * path:line in the rslab source tree.
 */
#ifndef RSLAB_AMD_H
#define RSLAB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, mapped from RslabError (src/error.rs:3-43). */
#define RSL_OK 0
#define RSL_NUMERICALLY_RANK_DEFICIENT 1 /* RslabError::NumericallyRankDeficient */
#define RSL_INVALID_INPUT 2              /* RslabError::InvalidInput */
#define RSL_DIMENSION_MISMATCH 3         /* RslabError::DimensionMismatch */
#define RSL_DEVICE_ERROR 4               /* InvalidInput("device: ...") in rslab terms */
#define RSL_OUT_OF_MEMORY 5

/* Fields. */
#define RSL_F64 0 /* f64, real symmetric */
#define RSL_C64 1 /* Complex<f64>, complex symmetric (PARDISO mtype 6), interleaved re/im */

/* Orderings (OrderingMethod, src/symbolic/mod.rs:28-138). */
#define RSL_ORDER_AMD 0
#define RSL_ORDER_METIS_ND 1
#define RSL_ORDER_NATURAL 2

/* Pivot policy (ZeroPivotAction, src/numeric/multifrontal_ldlt.rs:165-180). */
#define RSL_PIVOT_FAIL 0
#define RSL_PIVOT_PERTURB 1 /* PerturbToEps { abs_floor } */
#define RSL_PIVOT_FORCE 2   /* ForceAccept */

/* Precision of the Schur / contribution updates; pivot search, 2x2 inverse,
 * D and residuals always stay FP64. */
#define RSL_PREC_FP64 0
#define RSL_PREC_FP32_SCHUR 1
#define RSL_PREC_BF16_SCHUR 2

/* Numeric schedule (FactorMethod, src/numeric/multifrontal_ldlt.rs:257-295).
 * Both give the same factor: left-looking pulls each level's updates before
 * its cdiv, right-looking pushes a level's updates into every ancestor right
 * after its cdiv. FactorMethod::Multifrontal is not provided (rejected). */
#define RSL_METHOD_LEFT_LOOKING 0
#define RSL_METHOD_RIGHT_LOOKING 1
#define RSL_METHOD_MULTIFRONTAL 2

/* SolverSettings (src/numeric/multifrontal_ldlt.rs:302-394, defaults :583-614)
 * restricted to the knobs the device path honours. */
typedef struct {
  int32_t ordering;       /* default RSL_ORDER_AMD (OrderingMethod::default) */
  int32_t nemin;          /* default 16 */
  int32_t relax;          /* 1 = Some(RelaxAmalgamation), default 1 */
  int32_t relax_width;    /* default 256 */
  int32_t relax_rows;     /* default 64 */
  int32_t panel_nb;       /* default 64 (128 on supernodes >= 512 wide), max 128 */
  int32_t pivot;          /* default RSL_PIVOT_FAIL */
  double abs_floor;       /* PerturbToEps floor */
  int32_t scaling;        /* 1 = OnePassInfNorm (default), 0 = Identity */
  int32_t precision;      /* default RSL_PREC_FP64 */
  uint64_t nd_seed;       /* rslab-metis seed, default 1 */
  double wide_flops;      /* h*w*w above which a supernode's cdiv uses the
                             multi-workgroup panel path (tuning knob) */
  double drop_tol;        /* SolverSettings::drop_tol (multifrontal_ldlt.rs:640-643):
                             < 0 = None (default); tau >= 0 drops stored L
                             entries below tau * (largest in their column) */
  int64_t cmod_cap;       /* cmod product-buffer cap in elements; < 0 = automatic
                             (a quarter of the free device memory); levels that
                             need more take the gathered-tile cmod, 0 forces it
                             for every level (real field) */
  int32_t graphs;         /* 1 (default): a same-pattern refactor captures the
                             factor's launch sequence once as a HIP graph and
                             replays it; 0: every factor launches eagerly */
  int32_t solve_wide;     /* supernodes wider than this solve their own block in
                             multi-workgroup 64-column steps (tuning knob,
                             default 32; narrower ones use one workgroup) */
  int32_t method;         /* RSL_METHOD_*, default RSL_METHOD_LEFT_LOOKING
                             (SolverSettings::with_method); a plan setting */
  int32_t solve_dep;      /* 1 (default): a solve of up to 8 right-hand sides
                             runs each sweep as one launch whose work units
                             order themselves through device counters; wider
                             solves, and every solve when 0, use one launch
                             per level and block step (tuning knob, a plan
                             setting) */
  int32_t solve_split;    /* with solve_dep: the first tree level (leaves are
                             level 0) in the dependency-driven launch, the
                             levels below run one launch per level; < 0
                             (default): the first level with at most 128
                             supernodes (tuning knob, a plan setting) */
  double pivot_u;         /* LU twin's pivot threshold u (SolverSettings::
                             pivot_u, multifrontal_ldlt.rs:382, default 0.1,
                             clamped to [0, 1] as with_pivot_u :742-745): the
                             diagonal is kept unless |a_kk| < u max |a_ik| */
} rsl_settings;

typedef struct rsl_symbolic rsl_symbolic; /* LdltSymbolic + device plan */
typedef struct rsl_solver rsl_solver;     /* LdltSolver<T>: device factor */

/* Fill *s with the defaults listed above. */
void rsl_settings_default(rsl_settings *s);

/* LdltSymbolic::analyze_with (src/numeric/sparse_solver.rs:505-515):
 * lower-triangle CSC, rows sorted (CscMatrix, src/sparse/csc.rs:15-20). */
int rsl_analyze(int64_t n, const int64_t *col_ptr, const int64_t *row_idx,
                const rsl_settings *s, rsl_symbolic **out);

/* LdltSymbolic from the caller's own symbolic analysis, the input that
 * factor_numeric takes (&MultifrontalSymbolic, src/numeric/
 * multifrontal_ldlt.rs:1851-1855, whose SymbolicFactorization fields are
 * src/symbolic/mod.rs:333-405): the lower-CSC pattern of A, the permutation
 * perm (new -> old), and the supernode partition of the permuted columns in
 * postorder -- supernode s holds columns sn_first[s] .. sn_first[s] +
 * sn_ncol[s] - 1, consecutive from 0 -- with its tree (sn_parent[s] a later
 * supernode, or -1 for a root). Row structures and update lists are derived as
 * the reference derives them (multifrontal_ldlt.rs:2121-2152, 3386-3401).
 * RSL_INVALID_INPUT for a perm that is not a permutation, a partition that
 * does not tile 0..n-1, a parent that is not a later supernode, or a
 * supernode with rows in a supernode outside its ancestors. The ordering and
 * amalgamation fields of s are unused; its numeric fields apply. */
int rsl_plan_create(int64_t n, const int64_t *col_ptr, const int64_t *row_idx,
                    const int32_t *perm, int32_t nsuper, const int32_t *sn_first,
                    const int32_t *sn_ncol, const int32_t *sn_parent,
                    const rsl_settings *s, rsl_symbolic **out);

/* The same analysis (permutation, partition) with a device plan built for
 * other plan settings (panel_nb, wide_flops, solve_wide, method; the analysis fields
 * of s are unused) -- what the autotuner times candidates with, without
 * repeating the ordering. No counterpart in rslab. */
int rsl_symbolic_replan(const rsl_symbolic *sym, const rsl_settings *s, rsl_symbolic **out);

/* Batch of independent systems (no counterpart in rslab, which factors one
 * matrix per call; for callers with many small systems of uneven size):
 * each of the `count` lower-CSC matrices is analysed on its own exactly as by
 * rsl_analyze, and they are laid out as one block-diagonal system
 * diag(A_0, ..., A_count-1). rsl_factor, rsl_refactor, rsl_solve and
 * rsl_refine then act on the whole batch in one launch sequence (values, b and
 * x concatenated in system order); rsl_factor_inertia_batch gives each
 * system's inertia. */
int rsl_analyze_batch(int32_t count, const int64_t *n, const int64_t *const *col_ptr,
                      const int64_t *const *row_idx, const rsl_settings *s, rsl_symbolic **out);
int rsl_factor_inertia_batch(rsl_solver *f, int32_t count, int64_t *inertia);

/* Symbolic sizes: out[0]=n out[1]=nsuper out[2]=sum of column counts
 * out[3]=panel entries out[4]=levels out[5]=children entries
 * out[6]=full permuted pattern nnz out[7]=stored L entries (padded)
 * out[8]=cmod product-buffer elements of the largest level
 * out[9]=solve scratch rows (off-diagonal panel rows) and
 * work figures: flops[0]=factor flops sum (h-k)^2, flops[1]=geom flops
 * sum h^2 w (MultifrontalSymbolic / estimate_memory,
 * src/numeric/sparse_solver.rs:556-672). */
int rsl_symbolic_info(const rsl_symbolic *sym, int64_t out[10], double flops[2]);

/* SymbolicFactorization fields (src/symbolic/mod.rs:333-405): permutation
 * (new->old), inverse, supernode first column / width / panel height,
 * children (CSR), full pattern of P A P^T (CSC). Any pointer may be NULL. */
int rsl_symbolic_export(const rsl_symbolic *sym, int32_t *perm, int32_t *perm_inv,
                        int32_t *sn_first, int32_t *sn_ncol, int32_t *sn_nrow,
                        int32_t *ch_ptr, int32_t *ch_idx, int64_t *pp_colp,
                        int32_t *pp_rowi);

/* LdltSymbolic::factor (src/numeric/sparse_solver.rs:677-711) -> factor_numeric
 * (src/numeric/multifrontal_ldlt.rs:1851-1893), left-looking branch (or the
 * right-looking one, :1950-1977, for a symbolic planned with
 * RSL_METHOD_RIGHT_LOOKING). values are the CSC values in the analysed
 * pattern (interleaved for RSL_C64). s may be NULL (defaults); its method
 * must be the one the symbolic was planned for. The symbolic handle must
 * outlive the factor. */
int rsl_factor(rsl_symbolic *sym, int32_t field, const double *values,
               const rsl_settings *s, rsl_solver **out);

/* Same-pattern numeric refactorization, reusing plan and device storage
 * (LdltSymbolic reuse, src/numeric/sparse_solver.rs:457-488). */
int rsl_refactor(rsl_solver *f, const double *values);

/* Device-resident variant of rsl_refactor: d_values in device memory, same
 * layout (interleaved for RSL_C64). The equilibration and the ForceAccept
 * floor are computed on the device; in the real field they give the host
 * path's bits, so the factor equals rsl_refactor's on the same values (the
 * complex field's modulus is the device's hypot, equal to the host's to
 * rounding). Ordered after the work enqueued on `stream` (a hipStream_t; NULL
 * is the null stream); returns when done. */
int rsl_refactor_device(rsl_solver *f, const void *d_values, void *stream);

/* LdltSolver::solve_many (src/numeric/sparse_solver.rs:302-327) with
 * solve_ldlt_many (src/dense/ldlt_generic.rs:564-614): b, x host buffers,
 * row-major n x nrhs. */
int rsl_solve(rsl_solver *f, const double *b, int64_t nrhs, double *x);

/* Device-resident variant of rsl_solve: d_b, d_x device pointers (row-major
 * n x nrhs), asynchronous with respect to the host. The solve starts after
 * the work enqueued on `stream` (a hipStream_t; NULL is the null stream)
 * before the call, and work enqueued on `stream` after the call starts after
 * the solve. It runs on the handle's own stream, so every later call on the
 * same handle (refactor, refine, solve, on any stream) is ordered after it and
 * cannot overwrite the factor or the shared workspace it reads. */
int rsl_solve_device(rsl_solver *f, const void *d_b, int64_t nrhs, void *d_x,
                     void *stream);

/* rsl_solve_device without any allocation. No counterpart in rslab: for
 * callers that cannot allocate device memory during the call (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, result and stream ordering as rsl_solve_device, but the solve's
 * scratch (y and z: n x nrhs each, u: one row per off-diagonal panel row x
 * nrhs, each part 256-byte aligned) lives in one fixed device array of
 * RSL_SOLVE_NOALLOC_SCRATCH_BYTES per GPU, held from the library's first use
 * on that GPU until the process ends. A solve needing more returns
 * RSL_INVALID_INPUT with the sizes. Every call on one GPU shares that array:
 * calls that overlap in time on one GPU (two handles, two streams) race;
 * calls on one handle are ordered on its stream and do not. */
#define RSL_SOLVE_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int rsl_solve_device_noalloc(rsl_solver *f, const void *d_b, int64_t nrhs, void *d_x,
                             void *stream);

/* inertia[3] = (positive, negative, zero) (LdltSolver::inertia,
 * src/numeric/sparse_solver.rs:89-91); n_perturbed (:76-78); t_ms = last
 * numeric factor wall time. Any pointer may be NULL. */
int rsl_factor_info(const rsl_solver *f, int64_t inertia[3], int64_t *n_perturbed,
                    double *t_ms);

/* LdltFactors (src/dense/ldlt_generic.rs:31-66) in CSC: first call
 * rsl_factor_nnz to size l_row_idx / l_values, then rsl_factor_download. */
int rsl_factor_nnz(rsl_solver *f, int64_t *nnzl);
int rsl_factor_download(rsl_solver *f, int64_t *l_col_ptr, int32_t *l_row_idx,
                        double *l_values, double *d_diag, double *d_subdiag,
                        uint8_t *two_by_two, int32_t *perm);

/* Factor once, solve elsewhere (rslab keeps LdltFactors in memory only and
 * has no file format; this one is the library's own). rsl_factor_save writes
 * the handle's factor -- the scale, panels, D, pivot maps and inertia, which
 * is all a solve or a refinement reads -- with a digest of the symbolic plan.
 * rsl_factor_load reads it into a new handle for a symbolic of the same plan
 * (the same pattern and analysis, e.g. rsl_analyze with the same settings in
 * another process) without a numeric factorization; a file written for
 * another plan or field is RSL_INVALID_INPUT. The loaded handle solves,
 * refines and refactors like the original. */
int rsl_factor_save(rsl_solver *f, const char *path);
int rsl_factor_load(rsl_symbolic *sym, const char *path, rsl_solver **out);

/* MixedLdltSolver::solve_to (src/numeric/mixed.rs:267-285) around
 * refine_ladder (:140-221): IR with FP64 residuals against the original
 * matrix values a_values, escalating to FGMRES-IR. b, x row-major n x nrhs;
 * info[4*c..4*c+4) = {ir_iters, gmres_iters, backward_error, certified}. */
int rsl_refine(rsl_solver *f, const double *a_values, const double *b, int64_t nrhs,
               double target, double *x, double *info);

/* Device-resident variant of rsl_refine: d_a_values, d_b, d_x in device
 * memory (interleaved for RSL_C64), info on the host. Takes the same steps as
 * rsl_refine on the same inputs (real field; in the complex field ||A||inf
 * is formed with the device's hypot). Ordered after the work enqueued on
 * `stream`; returns when done. */
int rsl_refine_device(rsl_solver *f, const void *d_a_values, const void *d_b, int64_t nrhs,
                      double target, void *d_x, double *info, void *stream);

/* Krylov solvers with this factor as preconditioner (Preconditioner for
 * LdltSolver, src/numeric/iterative.rs:178-191: apply_block is the
 * multi-RHS device solve) and A as operator. a_values are A's values in the
 * analysed pattern (the operator may differ in value from the factored
 * matrix: a factor of a nearby matrix preconditions A); b, x (and x0) host
 * buffers, row-major n x nrhs (interleaved for complex); precond = 1 applies
 * the factor, 0 the identity (NoPreconditioner, :167-176). The right-hand
 * sides advance in lockstep, each following the single-vector recurrence on
 * its own, at most 256 per call. info[4*c..4*c+4) = {iterations, final
 * relative residual |b - A x|_2 / |b|_2, converged, stop reason (0
 * converged, 1 iteration budget, 2 breakdown; StopReason :335-348)}.
 *   rsl_gmres_block  gmres_block / gmres (:1881, :822-1021): FGMRES(restart),
 *                    MGS with DGKS re-orthogonalization per column, x0 may
 *                    be NULL (zero);
 *   rsl_cocg         cocg (:402-489), x0 = 0;
 *   rsl_cocr         cocr (:495-583), x0 = 0. */
int rsl_gmres_block(rsl_solver *f, const double *a_values, const double *b, int64_t nrhs, double tol,
                    int32_t max_iter, int32_t restart, int32_t precond, const double *x0, double *x,
                    double *info);
int rsl_cocg(rsl_solver *f, const double *a_values, const double *b, int64_t nrhs, double tol,
             int32_t max_iter, int32_t precond, double *x, double *info);
int rsl_cocr(rsl_solver *f, const double *a_values, const double *b, int64_t nrhs, double tol,
             int32_t max_iter, int32_t precond, double *x, double *info);

/* Per-kernel timing (no counterpart in rslab; a measurement aid). While
 * enabled, every launch of this factor's numeric factor, solve and refinement
 * is bracketed by HIP events on the stream it runs on. rsl_profile_enable
 * clears the counters. rsl_profile_read returns, for one launch kind,
 * out = {total ms, algorithmic flops, algorithmic bytes, launches}; flops are
 * real-arithmetic (a complex multiply-add counts 8), 2 per multiply-add.
 * Kinds run from 0 until rsl_profile_kind_name returns NULL. */
int rsl_profile_enable(rsl_solver *f, int32_t on);
int rsl_profile_read(rsl_solver *f, int32_t kind, double out[4]);
const char *rsl_profile_kind_name(int32_t kind);

void rsl_free_factor(rsl_solver *f);
void rsl_free_symbolic(rsl_symbolic *s);
const char *rsl_last_error(void);

/* Ordering only: AMD / nested dissection permutation (new->old) of a full
 * symmetric pattern (rslab_ordering_core::CscPattern contract,
 * crates/rslab-ordering-core/src/lib.rs:1-75). stats may be NULL;
 * AMD stats[0..6] = ncmpa, ndense, lnz, ndiv, nms_lu, nms_ldl, dmax. */
int rsl_order(int32_t method, int64_t n, const int64_t *col_ptr, const int64_t *row_idx,
              uint64_t seed, int32_t *perm, double *stats);

/* ---- LU twin (src/numeric/multifrontal_lu.rs): unsymmetric matrices given
 * as the full CSC (GeneralCsc, src/sparse/general.rs; rows strictly
 * increasing within each column). The analysis runs on the pattern A + A^T
 * (symmetrized_lower_pattern :364-409) with the settings' ordering and
 * amalgamation; the numeric factor is the supernodal left-looking LU with
 * threshold partial pivoting restricted to each supernode's own rows
 * (factor_lu_left_looking :2428-2671, lu_ll_factor_node :1629-2347) after the
 * two-sided equilibration D_r A D_c (:2733-2758). RSL_METHOD_LEFT_LOOKING
 * only; the multifrontal LU is not provided. */
typedef struct rsl_lu_symbolic rsl_lu_symbolic; /* LuSymbolic + device plan */
typedef struct rsl_lu_solver rsl_lu_solver;     /* LuSolver<T>: device factor */

/* LuSymbolic::analyze_with (:850-869). */
int rsl_lu_analyze(int64_t n, const int64_t *col_ptr, const int64_t *row_idx, const rsl_settings *s,
                   rsl_lu_symbolic **out);
/* out[8] = {n, nnz, supernodes, tree levels, stored L entries, stored U
 * entries (both with the diagonal, before dropping), cmod tasks, panel-step
 * units}; flops[0] = sum over supernodes and own columns k of
 * 2 (h-k-1)^2 + (h-k-1) (real arithmetic). Either pointer may be NULL. */
int rsl_lu_symbolic_info(const rsl_lu_symbolic *sym, int64_t *out, double *flops);
/* The analysis of A + A^T, as rsl_symbolic_export. */
int rsl_lu_symbolic_export(const rsl_lu_symbolic *sym, int32_t *perm, int32_t *perm_inv, int32_t *sn_first,
                           int32_t *sn_ncol, int32_t *sn_nrow, int32_t *ch_ptr, int32_t *ch_idx,
                           int64_t *pp_colp, int32_t *pp_rowi);
/* LuSymbolic::factor (:880-907) -> factor_general_lu_numeric (:2673-2800):
 * values in the analysed pattern (interleaved for RSL_C64); s may be NULL;
 * uses pivot (ZeroPivotAction), abs_floor, pivot_u, drop_tol (applied when
 * the factors are emitted, :1562-1602) and method. The symbolic handle must
 * outlive the factor. */
int rsl_lu_factor(rsl_lu_symbolic *sym, int32_t field, const double *values, const rsl_settings *s,
                  rsl_lu_solver **out);
/* Same-pattern refactor in place (storage and plan reused). */
int rsl_lu_refactor(rsl_lu_solver *f, const double *values);
/* LuSolver::solve_many (:1212-1217) -> solve_lu_many (:3158-3207): b, x host
 * buffers, row-major n x nrhs (interleaved for RSL_C64). */
int rsl_lu_solve(rsl_lu_solver *f, const double *b, int64_t nrhs, double *x);
/* solve_lu_many with b, x in device memory, ordered after the work enqueued
 * on `stream` (hipStream_t; NULL the default stream), which then waits for
 * the solve; returns without waiting. */
int rsl_lu_solve_device(rsl_lu_solver *f, const void *d_b, int64_t nrhs, void *d_x, void *stream);
/* LuSolver::solve_refined (:1219-1227) -> solve_lu_refined (:3285-3322), per
 * column, on the device: x = (LU)^-1 b, then up to max_iter corrections
 * x += (LU)^-1 (b - A x) with the residual against a_values (A's values in
 * the analysed pattern), every correction evaluated; each column returns its
 * iterate of least max |b - A x| and stops early at a zero residual. b, x
 * host buffers, row-major n x nrhs (interleaved for RSL_C64), nrhs <= 4096;
 * info (may be NULL) = 2 per column: {corrections applied, least max |r|}. */
int rsl_lu_refine(rsl_lu_solver *f, const double *a_values, const double *b, int64_t nrhs, int32_t max_iter,
                  double *x, double *info);
/* out[6] = {n, n_perturbed (LuSolver::n_perturbed :1234), stored L entries,
 * stored U entries (LuFactors, after drop_tol), field, 0}; t[0] = the last
 * factor's wall time in ms (host to host). Either pointer may be NULL. */
int rsl_lu_info(rsl_lu_solver *f, int64_t *out, double *t);
/* LuFactors (:321-352): L CSC (n+1, nnz_l; unit diagonal first in each
 * column), U CSR (n+1, nnz_u; the pivot first in each row), perm and
 * perm_row (position -> original column / row), d_row, d_col; sizes from
 * rsl_lu_info. Any pointer may be NULL. */
int rsl_lu_factor_download(rsl_lu_solver *f, int64_t *l_col_ptr, int32_t *l_row_idx, double *l_values,
                           int64_t *u_row_ptr, int32_t *u_col_idx, double *u_values, int32_t *perm,
                           int32_t *perm_row, double *d_row, double *d_col);
void rsl_lu_free_factor(rsl_lu_solver *f);
void rsl_lu_free_symbolic(rsl_lu_symbolic *s);

/* Synthetic inputs: `count` draws of rslab's xorshift64* Rng::range(lo, hi)
 * from Rng::new(seed) (src/matgen/mod.rs:63-96). */
void rsl_rng_uniform(uint64_t seed, double lo, double hi, int64_t count, double *out);

#ifdef __cplusplus
}
#endif
#endif
