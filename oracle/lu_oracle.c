/*
 * CPU oracle for the supernodal left-looking LU twin -- TEST INFRASTRUCTURE
 * ONLY. See lu_oracle.h for the reference functions restated here. The
 * field-generic body is instantiated for f64 and c64 (|z| the modulus,
 * 1/z = conj(z)/|z|^2 as num-complex's Complex::inv).
 */
#include "lu_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- real field ---- */
#define T double
#define SFX f64
#define MAG(x) fabs(x)
#define MAGSQ(x) ((x) * (x))
#define RECIP(x) (1.0 / (x))
#define FROM_REAL(x) ((double)(x))
#define IS_ZERO(x) ((x) == 0.0)
#define ZERO 0.0
#define ONE 1.0
#define FMADD(a, b, c) fma((a), (b), (c))
#include "lu_oracle_body.h"
#undef T
#undef SFX
#undef MAG
#undef MAGSQ
#undef RECIP
#undef FROM_REAL
#undef IS_ZERO
#undef ZERO
#undef ONE
#undef FMADD

/* ---- complex field ---- */
typedef double _Complex lc64;
static inline lc64 lc64_recip(lc64 z) {
  double ns = creal(z) * creal(z) + cimag(z) * cimag(z);
  return CMPLX(creal(z) / ns, -cimag(z) / ns);
}
#define T lc64
#define SFX c64
#define MAG(x) hypot(creal(x), cimag(x))
#define MAGSQ(x) (creal(x) * creal(x) + cimag(x) * cimag(x))
#define RECIP(x) lc64_recip(x)
#define FROM_REAL(x) CMPLX((double)(x), 0.0)
#define IS_ZERO(x) (creal(x) == 0.0 && cimag(x) == 0.0)
#define ZERO CMPLX(0.0, 0.0)
#define ONE CMPLX(1.0, 0.0)
#define FMADD(a, b, c) ((a) * (b) + (c))
#include "lu_oracle_body.h"

oru_factors *oru_factor(const ora_sym *s, int cplx, const int64_t *a_colp, const int32_t *a_rowi,
                        const double *a_val, int pivot, double abs_floor, double pivot_u,
                        double drop_tol) {
  oru_factors *f = calloc(1, sizeof(oru_factors));
  f->n = s->n;
  f->cplx = cplx;
  if (cplx)
    lu_factor_c64(s, a_colp, a_rowi, (const lc64 *)a_val, pivot, abs_floor, pivot_u, drop_tol, f);
  else
    lu_factor_f64(s, a_colp, a_rowi, a_val, pivot, abs_floor, pivot_u, drop_tol, f);
  return f;
}

void oru_free(oru_factors *f) {
  if (!f) return;
  free(f->l_col_ptr);
  free(f->l_row_idx);
  free(f->l_values);
  free(f->u_row_ptr);
  free(f->u_col_idx);
  free(f->u_values);
  free(f->perm);
  free(f->perm_row);
  free(f->d_row);
  free(f->d_col);
  free(f);
}

int oru_solve_many(const oru_factors *f, const double *b, int32_t nrhs, double *x) {
  if (f->status) return 1;
  if (f->cplx) lu_solve_c64(f, (const lc64 *)b, nrhs, (lc64 *)x);
  else lu_solve_f64(f, b, nrhs, x);
  return 0;
}
