/*
 * CPU oracle for the supernodal LDL^T path -- TEST INFRASTRUCTURE ONLY.
 * See ldlt_oracle.h for the reference functions restated here.
 * The field-generic body is instantiated for f64 and complex-symmetric c64
 * (no conjugation anywhere; |z| is the modulus, 1/z = conj(z)/|z|^2 as
 * num-complex's Complex::inv).
 */
#include "ldlt_oracle.h"

#include <complex.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

/* Reduced-precision Schur products (real field only): 0 = FP64, 1 = FP32,
 * 2 = BF16 inputs. The cmod products L_k G^T and the deferred Schur products
 * L21 (L D)^T then take both operands rounded to FP32 (to BF16 by way of FP32,
 * round to nearest even) and sum in FP32; the sum is subtracted from the
 * FP64 panel. Pivoting, D, the 2x2 inverses and the replay stay FP64. */
static int ora_prec = 0;

static inline float ora_lowp(double v) {
  float f = (float)v;
  if (ora_prec == 2) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
  }
  return f;
}

float ora_round_lowp(double v, int prec) {
  int keep = ora_prec;
  ora_prec = prec;
  float f = ora_lowp(v);
  ora_prec = keep;
  return f;
}

/* Threads of the factor (OpenMP): all available unless ora_set_threads
 * names a count. Every entry is computed in the same order whatever the
 * count, so the factor is the same bit for bit. */
static int ora_nthreads = 0;

static inline int ora_threads(void) {
#ifdef _OPENMP
  return ora_nthreads > 0 ? ora_nthreads : omp_get_max_threads();
#else
  return 1;
#endif
}

static inline int ora_thread(void) {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

void ora_set_threads(int n) { ora_nthreads = n > 0 ? n : 0; }
int ora_get_threads(void) { return ora_threads(); }

/* ---- real field ---- */
#define LOWP 1
#define T double
#define SFX f64
#define MAG(x) fabs(x)
#define MAGSQ(x) ((x) * (x))
#define REAL(x) (x)
#define RECIP(x) (1.0 / (x))
#define FROM_REAL(x) ((double)(x))
#define IS_ZERO(x) ((x) == 0.0)
#define ZERO 0.0
#define ONE 1.0
#define FMADD(a, b, c) fma((a), (b), (c))
#include "ldlt_oracle_body.h"
#undef T
#undef SFX
#undef MAG
#undef MAGSQ
#undef REAL
#undef RECIP
#undef FROM_REAL
#undef IS_ZERO
#undef ZERO
#undef ONE
#undef FMADD
#undef LOWP

/* ---- complex-symmetric field ---- */
#define LOWP 0
typedef double _Complex c64;
static inline c64 c64_recip(c64 z) {
  double ns = creal(z) * creal(z) + cimag(z) * cimag(z);
  return CMPLX(creal(z) / ns, -cimag(z) / ns);
}
#define T c64
#define SFX c64
#define MAG(x) hypot(creal(x), cimag(x))
#define MAGSQ(x) (creal(x) * creal(x) + cimag(x) * cimag(x))
#define REAL(x) creal(x)
#define RECIP(x) c64_recip(x)
#define FROM_REAL(x) CMPLX((double)(x), 0.0)
#define IS_ZERO(x) (creal(x) == 0.0 && cimag(x) == 0.0)
#define ZERO CMPLX(0.0, 0.0)
#define ONE CMPLX(1.0, 0.0)
#define FMADD(a, b, c) ((a) * (b) + (c))
#include "ldlt_oracle_body.h"

ora_factors *ora_factor(const ora_sym *s, int cplx, const int64_t *a_colp,
                        const int32_t *a_rowi, const double *a_val, int panel_nb,
                        int pivot, double abs_floor, double drop_tol, int prec) {
  ora_factors *f = calloc(1, sizeof(ora_factors));
  f->n = s->n;
  f->cplx = cplx;
  ora_prec = cplx ? 0 : prec;
  if (cplx)
    factor_c64(s, a_colp, a_rowi, (const c64 *)a_val, panel_nb, pivot, abs_floor, drop_tol, f);
  else
    factor_f64(s, a_colp, a_rowi, a_val, panel_nb, pivot, abs_floor, drop_tol, f);
  return f;
}

void ora_free(ora_factors *f) {
  if (!f) return;
  free(f->l_col_ptr);
  free(f->l_row_idx);
  free(f->l_values);
  free(f->d_diag);
  free(f->d_subdiag);
  free(f->two_by_two);
  free(f->perm);
  free(f);
}

int ora_solve_many(const ora_factors *f, const double *b, int32_t nrhs, double *x) {
  if (f->status) return 1;
  if (f->cplx) return solve_many_c64(f, (const c64 *)b, nrhs, (c64 *)x);
  return solve_many_f64(f, b, nrhs, x);
}

void ora_onepass_scale(int32_t n, int cplx, const int64_t *colp, const int32_t *rowi,
                       const double *val, double *scale) {
  if (cplx) onepass_scale_c64(n, colp, rowi, (const c64 *)val, scale);
  else onepass_scale_f64(n, colp, rowi, val, scale);
}

void ora_symv(int32_t n, int cplx, const int64_t *colp, const int32_t *rowi,
              const double *val, const double *x, double *y) {
  if (cplx) symv_c64(n, colp, rowi, (const c64 *)val, (const c64 *)x, (c64 *)y);
  else symv_f64(n, colp, rowi, val, x, y);
}

double ora_sym_norm_inf(int32_t n, int cplx, const int64_t *colp, const int32_t *rowi,
                        const double *val) {
  if (cplx) return norm_inf_c64(n, colp, rowi, (const c64 *)val);
  return norm_inf_f64(n, colp, rowi, val);
}
