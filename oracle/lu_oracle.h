/*
 * CPU oracle for the supernodal left-looking LU twin -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of rslab's unsymmetric LU (the general-matrix twin
 * of the LDL^T path), used only by tests/ and bench.py's CPU baseline leg.
 * Nothing in the product links or calls it.
 *
 * Reference functions restated (paths relative to the rslab source tree):
 *   factor_general_lu_numeric  src/numeric/multifrontal_lu.rs:2673-2800
 *       (ZeroPivotAction floors :2716-2723, two-sided equilibration
 *        :2733-2758, A_perm and its transpose :2760-2778)
 *   factor_lu_left_looking     src/numeric/multifrontal_lu.rs:2428-2671
 *   lu_ll_factor_node          src/numeric/multifrontal_lu.rs:1629-2347
 *       (assembly :1670-1686, sequential cmod :1881-2002, blocked cdiv
 *        with restricted threshold partial pivoting :2005-2283, index
 *        maps :2328-2344)
 *   apply_panel_trailing       src/numeric/multifrontal_lu.rs:1357-1387
 *   emit_and_free              src/numeric/multifrontal_lu.rs:1525-1618
 *   solve_lu_block             src/numeric/multifrontal_lu.rs:3209-3283
 *   symmetrized_lower_pattern  src/numeric/multifrontal_lu.rs:364-409
 *   perturb_pivot              src/numeric/multifrontal_ldlt.rs:786-795
 *   inv_sqrt_scale_guarded     src/scaling/mod.rs:85-96
 *
 * The reference runs each update as one of three kernels (scalar loop,
 * serial GEMM, column-tiled parallel GEMM) picked by size; all three compute
 * each updater's product and subtract it, in ascending updater order, and
 * differ only in the summation order inside a product. This file uses the
 * scalar loop's order for every update.
 */
#ifndef RSLAB_LU_ORACLE_H
#define RSLAB_LU_ORACLE_H
#include <stdint.h>

#include "ldlt_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t n, cplx, status;  /* status: 0 ok, 1 numerically rank deficient */
  int64_t n_perturbed;
  int64_t nnz_l, nnz_u;
  int64_t *l_col_ptr;       /* n+1: L in CSC, unit lower, diagonal first */
  int32_t *l_row_idx;
  double *l_values;         /* x2 interleaved re/im when cplx */
  int64_t *u_row_ptr;       /* n+1: U in CSR, pivot first in each row */
  int32_t *u_col_idx;
  double *u_values;
  int32_t *perm;            /* factorization position -> original column */
  int32_t *perm_row;        /* factorization position -> original row */
  double *d_row, *d_col;    /* the factor is of diag(d_row) A diag(d_col) */
} oru_factors;

/* s: the analysis of the symmetrized pattern A + A^T (its pp pattern is that
 * pattern, permuted). a_*: the full unsymmetric matrix, CSC, sorted rows,
 * unscaled values. pivot: 0 Fail, 1 PerturbToEps{abs_floor}, 2 ForceAccept.
 * pivot_u: threshold in [0, 1]. drop_tol < 0 means None. */
oru_factors *oru_factor(const ora_sym *s, int cplx, const int64_t *a_colp, const int32_t *a_rowi,
                        const double *a_val, int pivot, double abs_floor, double pivot_u,
                        double drop_tol);
void oru_free(oru_factors *f);
/* solve_lu_block: b, x row-major n x nrhs (complex interleaved). 0 = ok. */
int oru_solve_many(const oru_factors *f, const double *b, int32_t nrhs, double *x);

#ifdef __cplusplus
}
#endif
#endif
