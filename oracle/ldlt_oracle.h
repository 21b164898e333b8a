/*
 * CPU oracle for the supernodal left-looking LDL^T path -- TEST INFRASTRUCTURE.
 *
 * This is a plain-C restatement of rslab's numeric factor and solve, used only
 * by tests/, by __graft_entry__.smoke() as the checker, and by bench.py as the
 * reported CPU baseline. Nothing in the product links or calls it.
 *
 * Reference functions restated (paths relative to the rslab source tree):
 *   factor_left_looking   src/numeric/multifrontal_ldlt.rs:3346-3576
 *   ll_factor_node        src/numeric/multifrontal_ldlt.rs:2400-2750
 *   ll_cdiv_emit          src/numeric/multifrontal_ldlt.rs:2759-3258
 *   apply_bk_panel_trailing src/numeric/multifrontal_ldlt.rs:1347-1416
 *   ldlt_emit_and_free    src/numeric/multifrontal_ldlt.rs:2328-2381
 *   row structures        src/numeric/multifrontal_ldlt.rs:2121-2152
 *   perturb_pivot         src/numeric/multifrontal_ldlt.rs:786-795
 *   swap_sym_lower_bounded src/dense/ldlt_generic.rs:88-110
 *   solve_ldlt_block      src/dense/ldlt_generic.rs:618-714
 *   onepass_scale         src/numeric/sparse_solver.rs:399-418
 *   inv_sqrt_scale_guarded src/scaling/mod.rs:85-96
 *   CscMatrix::symv       src/sparse/csc.rs:294-309
 *   sym_norm_inf          src/numeric/mixed.rs:358-371
 *
 * How each function is pinned (independently of this file) is recorded in the
 * tests: closed-form inertia, dense LAPACK Bunch-Kaufman (scipy.linalg.ldl) on
 * single-front matrices, the reference's own known-answer tests, residuals.
 */
#ifndef RSLAB_LDLT_ORACLE_H
#define RSLAB_LDLT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t n, nsuper;
  const int32_t *perm, *perm_inv;       /* symbolic permutation new->old, old->new */
  const int32_t *sn_first, *sn_ncol;    /* per supernode, postorder */
  const int32_t *ch_ptr, *ch_idx;       /* children lists (CSR, nsuper+1) */
  const int64_t *pp_colp;               /* full pattern of P A P^T */
  const int32_t *pp_rowi;
} ora_sym;

typedef struct {
  int32_t n, cplx, status; /* status: 0 ok, 1 numerically rank deficient */
  int64_t nnzl, n_perturbed;
  int64_t inertia[3];      /* positive, negative, zero */
  int64_t *l_col_ptr;      /* n+1 */
  int32_t *l_row_idx;      /* nnzl */
  double *l_values;        /* nnzl (x2 interleaved re/im when cplx) */
  double *d_diag, *d_subdiag; /* n (x2 when cplx) */
  uint8_t *two_by_two;     /* n */
  int32_t *perm;           /* n: elimination position -> user index */
} ora_factors;

/* pivot: 0 = Fail, 1 = PerturbToEps{abs_floor}, 2 = ForceAccept.
 * a_val holds the (already equilibrated) values of the user-order lower CSC.
 * drop_tol < 0 means None. prec (real field): 0 = FP64; 1 = FP32, 2 = BF16
 * Schur products -- the cmod and deferred-Schur products take operands
 * rounded to that format and sum in FP32, the sum is subtracted in FP64
 * (emulation of the device's reduced-precision paths; not in rslab). */
ora_factors *ora_factor(const ora_sym *s, int cplx, const int64_t *a_colp,
                        const int32_t *a_rowi, const double *a_val, int panel_nb,
                        int pivot, double abs_floor, double drop_tol, int prec);
/* v rounded as the reduced-precision products round their operands. */
float ora_round_lowp(double v, int prec);
/* Threads for ora_factor (OpenMP; 0 = all available). The supernodes of one
 * tree level run concurrently, or, on narrow levels, the column loops inside
 * one supernode; the result does not depend on the count. */
void ora_set_threads(int n);
int ora_get_threads(void);
void ora_free(ora_factors *f);
/* solve_ldlt_many: b, x row-major n x nrhs (complex interleaved). 0 = ok. */
int ora_solve_many(const ora_factors *f, const double *b, int32_t nrhs, double *x);
void ora_onepass_scale(int32_t n, int cplx, const int64_t *colp, const int32_t *rowi,
                       const double *val, double *scale);
void ora_symv(int32_t n, int cplx, const int64_t *colp, const int32_t *rowi,
              const double *val, const double *x, double *y);
double ora_sym_norm_inf(int32_t n, int cplx, const int64_t *colp, const int32_t *rowi,
                        const double *val);

#ifdef __cplusplus
}
#endif
#endif
