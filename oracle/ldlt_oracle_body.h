/*
 * Field-generic body of the oracle, included once per field by ldlt_oracle.c
 * with T, SFX and the scalar helpers defined. TEST INFRASTRUCTURE ONLY.
 */

#define FN2(a, b) a##_##b
#define FN1(a, b) FN2(a, b)
#define FN(name) FN1(name, SFX)

/* swap_sym_lower_bounded (ldlt_generic.rs:88-110) */
static void FN(swap_sym)(T *a, int64_t n, int64_t p, int64_t q, int64_t row_limit) {
  T t;
  for (int64_t i = q + 1; i < row_limit; ++i) {
    t = a[p * n + i]; a[p * n + i] = a[q * n + i]; a[q * n + i] = t;
  }
  for (int64_t i = p + 1; i < q; ++i) {
    t = a[p * n + i]; a[p * n + i] = a[i * n + q]; a[i * n + q] = t;
  }
  t = a[p * n + p]; a[p * n + p] = a[q * n + q]; a[q * n + q] = t;
  for (int64_t j = 0; j < p; ++j) {
    t = a[j * n + p]; a[j * n + p] = a[j * n + q]; a[j * n + q] = t;
  }
}

/* perturb_pivot (multifrontal_ldlt.rs:786-795) */
static T FN(perturb)(T d, double floor) {
  double mag = MAG(d);
  if (mag >= floor) return d;
  if (mag == 0.0) return FROM_REAL(floor);
  return d * FROM_REAL(floor / mag);
}

/* apply_bk_panel_trailing (multifrontal_ldlt.rs:1347-1416) over rows [r0,r1) */
static void FN(apply_trailing)(T *base, int64_t nrow, int64_t kb, int64_t ke, const T *d,
                               const T *dsub, const uint8_t *two, const int64_t *deep_swaps,
                               const T *snap, int64_t nb, int64_t r0, int64_t r1) {
  int64_t k = kb;
  while (k < ke) {
    int64_t kp = deep_swaps[k - kb];
    if (kp >= 0) {
      int64_t src = two[k] ? k + 1 : k;
      T *ca = base + src * nrow, *cb = base + kp * nrow;
      for (int64_t i = r0; i < r1; ++i) { T t = ca[i]; ca[i] = cb[i]; cb[i] = t; }
    }
    if (two[k]) {
      T d11 = d[k], d21 = dsub[k], d22 = d[k + 1];
      T detinv = RECIP(d11 * d22 - d21 * d21);
      T *ck = base + k * nrow, *ck1 = base + (k + 1) * nrow;
      for (int64_t j = k + 2; j < ke; ++j) {
        T l1 = snap[(k - kb) * nb + (j - kb)], l2 = snap[(k + 1 - kb) * nb + (j - kb)];
        T *cj = base + j * nrow;
        for (int64_t i = r0; i < r1; ++i) cj[i] = cj[i] - ck[i] * l1 - ck1[i] * l2;
      }
      for (int64_t i = r0; i < r1; ++i) {
        T a = ck[i], b = ck1[i];
        ck[i] = (d22 * a - d21 * b) * detinv;
        ck1[i] = (d11 * b - d21 * a) * detinv;
      }
      k += 2;
    } else {
      T dinv = RECIP(d[k]);
      T *ck = base + k * nrow;
      for (int64_t j = k + 1; j < ke; ++j) {
        T m = snap[(k - kb) * nb + (j - kb)];
        if (!IS_ZERO(m)) {
          T *cj = base + j * nrow;
          for (int64_t i = r0; i < r1; ++i) cj[i] = cj[i] - ck[i] * m;
        }
      }
      for (int64_t i = r0; i < r1; ++i) ck[i] = ck[i] * dinv;
      k += 1;
    }
  }
}

/* G[npk x nck] = L_k[rows] * D_k (block diagonal 1x1 / 2x2), column-major. */
static void FN(apply_d)(const T *pk, int64_t nrk, int64_t row0, int64_t npk, int64_t nck,
                        const T *dk, const T *dsk, const uint8_t *twok, T *g) {
  int64_t ck = 0;
  while (ck < nck) {
    if (twok[ck]) {
      T d11 = dk[ck], d21 = dsk[ck], d22 = dk[ck + 1];
      for (int64_t i = 0; i < npk; ++i) {
        T a = pk[(row0 + i) + ck * nrk], b = pk[(row0 + i) + (ck + 1) * nrk];
        g[i + ck * npk] = d11 * a + d21 * b;
        g[i + (ck + 1) * npk] = d21 * a + d22 * b;
      }
      ck += 2;
    } else {
      for (int64_t i = 0; i < npk; ++i) g[i + ck * npk] = pk[(row0 + i) + ck * nrk] * dk[ck];
      ck += 1;
    }
  }
}

typedef struct {
  T *panel, *d, *dsub;
  uint8_t *two;
  int64_t *lperm;
} FN(node);

/* Shared state of one factorization: the symbolic, the permuted matrix, the
 * row structures and update lists, and the output maps (each supernode writes
 * only its own entries of them). */
typedef struct {
  const ora_sym *s;
  int64_t n, ns, panel_nb;
  int have_floor;
  double floor;
  const int64_t *pcnt, *prow;
  const T *pval;
  int64_t **rs;
  const int64_t *rsn, *ucnt, *uidx, *e_off;
  int64_t *e_of_g;
  T *d_out, *ds_out;
  uint8_t *two_out;
  int32_t *perm_out;
  FN(node) *nodes;
  int inner; /* run the loops inside a supernode on all threads */
} FN(ctx);

/* Per-thread scratch. gloc maps a global row to its local panel row while a
 * supernode is being factored and is all -1 between supernodes. */
typedef struct {
  int64_t *gloc;
  T *gbuf, *ubuf;
  float *fbuf;
  int64_t gcap, ucap;
  int64_t cnt[4]; /* positive, negative, zero, perturbed */
} FN(work);

/* cmod (ll_factor_node, multifrontal_ldlt.rs:2452-2728) for one target
 * column: U[r, c] = sum_ck L_k[p0 + r, ck] * G[c, ck] for r >= c, subtracted
 * from the panel. ub / fb are mrows of scratch. */
static void FN(cmod_col)(T *P, int64_t nrow, int64_t first, const FN(node) *K, int64_t nck, int64_t nrk,
                         const int64_t *ok, int64_t p0, int64_t mrows, int64_t npk, const T *gbuf,
                         const int64_t *gloc, int64_t c, T *ub, float *fb) {
  int64_t tcol = ok[p0 + c] - first;
  T *dst = P + tcol * nrow;
  if (LOWP && ora_prec) {
    /* reduced-precision product: rounded operands, FP32 sums, FP64 subtract */
    for (int64_t r = c; r < mrows; ++r) fb[r] = 0.0f;
    for (int64_t ck = 0; ck < nck; ++ck) {
      float g = ora_lowp(REAL(gbuf[c + ck * npk]));
      const T *lc = K->panel + nck + p0 + ck * nrk;
      for (int64_t r = c; r < mrows; ++r) fb[r] = fb[r] + ora_lowp(REAL(lc[r])) * g;
    }
    for (int64_t r = c; r < mrows; ++r) {
      int64_t li = gloc[ok[p0 + r]];
      dst[li] = dst[li] - (double)fb[r];
    }
    return;
  }
  for (int64_t r = c; r < mrows; ++r) ub[r] = ZERO;
  for (int64_t ck = 0; ck < nck; ++ck) {
    T g = gbuf[c + ck * npk];
    const T *lc = K->panel + nck + p0 + ck * nrk;
    for (int64_t r = c; r < mrows; ++r) ub[r] = ub[r] + lc[r] * g;
  }
  for (int64_t r = c; r < mrows; ++r) {
    int64_t li = gloc[ok[p0 + r]];
    dst[li] = dst[li] - ub[r];
  }
}

/* deferred Schur (multifrontal_ldlt.rs:3085-3167) for target column cc of
 * the trailing block: P[ke + rr, ke + cc] -= sum_q L21[rr, q] G[cc, q], rr >= cc,
 * with g = L21 D (mt x pw, column-major). */
static void FN(schur_col)(T *P, int64_t nrow, int64_t kb, int64_t ke, int64_t pw, int64_t mt, const T *g,
                          int64_t cc) {
  T *dst = P + (ke + cc) * nrow + ke;
  if (LOWP && ora_prec) {
    /* reduced-precision product L21 (L D)^T: rounded operands, FP32 sums
     * over the panel's columns, FP64 subtract */
    for (int64_t rr = cc; rr < mt; ++rr) {
      float acc = 0.0f;
      for (int64_t q = 0; q < pw; ++q)
        acc = acc + ora_lowp(REAL(P[(ke + rr) + (kb + q) * nrow])) * ora_lowp(REAL(g[q * mt + cc]));
      dst[rr] = dst[rr] - (double)acc;
    }
    return;
  }
  for (int64_t q = 0; q < pw; ++q) {
    T r = P[(ke + cc) + (kb + q) * nrow];
    const T *gq = g + q * mt;
    for (int64_t rr = cc; rr < mt; ++rr) dst[rr] = dst[rr] - gq[rr] * r;
  }
}

/* ll_factor_node + ll_cdiv_emit for supernode sn (its updaters are finished).
 * Returns 0 or 1 (rank deficient). */
static int FN(factor_node)(const FN(ctx) *X, FN(work) *W, int64_t sn) {
  const ora_sym *s = X->s;
  const int64_t panel_nb = X->panel_nb;
  const int have_floor = X->have_floor;
  const double floor = X->floor;
  const int64_t *pcnt = X->pcnt, *prow = X->prow;
  const T *pval = X->pval;
  int64_t **rs = X->rs;
  const int64_t *rsn = X->rsn, *ucnt = X->ucnt, *uidx = X->uidx, *e_off = X->e_off;
  int64_t *e_of_g = X->e_of_g;
  T *d_out = X->d_out, *ds_out = X->ds_out;
  uint8_t *two_out = X->two_out;
  int32_t *perm_out = X->perm_out;
  FN(node) *nodes = X->nodes;
  int64_t *gloc = W->gloc;
  T *gbuf = W->gbuf, *ubuf = W->ubuf;
  float *fbuf = W->fbuf;
  int64_t gcap = W->gcap, ucap = W->ucap;
  const double alpha = (1.0 + sqrt(17.0)) / 8.0;
  int status = 0;
  int64_t ipos = 0, ineg = 0, izero = 0, nperturbed = 0;
  const int64_t first = s->sn_first[sn], ncol = s->sn_ncol[sn], nrow = rsn[sn];
  T *P = calloc((size_t)(nrow * ncol > 0 ? nrow * ncol : 1), sizeof(T));
  for (int64_t li = 0; li < nrow; ++li) gloc[rs[sn][li]] = li;
  /* assembly */
  for (int64_t p = 0; p < ncol; ++p) {
    int64_t c = first + p;
    for (int64_t k = pcnt[c]; k < pcnt[c + 1]; ++k) {
      int64_t li = gloc[prow[k]];
      P[li + p * nrow] = P[li + p * nrow] + pval[k];
    }
  }
  /* cmod: every descendant updater, ascending */
  for (int64_t u = ucnt[sn]; u < ucnt[sn + 1]; ++u) {
    int64_t kk = uidx[u], nck = s->sn_ncol[kk], nrk = rsn[kk];
    const int64_t *ok = rs[kk] + nck;
    int64_t nok = nrk - nck, p0 = 0, p1;
    while (p0 < nok && ok[p0] < first) ++p0;
    p1 = p0;
    while (p1 < nok && ok[p1] < first + ncol) ++p1;
    int64_t npk = p1 - p0;
    if (npk == 0) continue;
    const FN(node) *K = &nodes[kk];
    int64_t mrows = nok - p0;
    if (gcap < npk * nck) { free(gbuf); gcap = npk * nck; gbuf = malloc(sizeof(T) * gcap); }
    if (ucap < mrows) {
      free(ubuf); free(fbuf);
      ucap = mrows; ubuf = malloc(sizeof(T) * ucap); fbuf = malloc(sizeof(float) * ucap);
    }
    FN(apply_d)(K->panel, nrk, nck + p0, npk, nck, K->d, K->dsub, K->two, gbuf);
    /* each target column c is written by its own product column only */
    if (X->inner && npk > 1) {
#pragma omp parallel num_threads(ora_threads())
      {
        T *ub = malloc(sizeof(T) * mrows);
        float *fb = malloc(sizeof(float) * mrows);
#pragma omp for schedule(dynamic, 4)
        for (int64_t c = 0; c < npk; ++c)
          FN(cmod_col)(P, nrow, first, K, nck, nrk, ok, p0, mrows, npk, gbuf, gloc, c, ub, fb);
        free(ub);
        free(fb);
      }
    } else {
      for (int64_t c = 0; c < npk; ++c)
        FN(cmod_col)(P, nrow, first, K, nck, nrk, ok, p0, mrows, npk, gbuf, gloc, c, ubuf, fbuf);
    }
  }
  /* cdiv */
  int64_t nb = panel_nb;
  if (ncol >= 512 && nb < 128) nb = 128;
  T *d = calloc(ncol ? ncol : 1, sizeof(T));
  T *dsub = calloc(ncol ? ncol : 1, sizeof(T));
  uint8_t *two = calloc(ncol ? ncol : 1, 1);
  int64_t *lperm = malloc(sizeof(int64_t) * (nrow ? nrow : 1));
  for (int64_t i = 0; i < nrow; ++i) lperm[i] = i;
  T *l1 = calloc(nrow ? nrow : 1, sizeof(T)), *l2 = calloc(nrow ? nrow : 1, sizeof(T));
  int64_t *deep = malloc(sizeof(int64_t) * nb);
  T *snap = calloc(nb * nb, sizeof(T));
  for (int64_t kb = 0; kb < ncol && status == 0;) {
    int64_t ke = kb + nb < ncol ? kb + nb : ncol;
    for (int64_t q = 0; q < nb; ++q) deep[q] = -1;
    int64_t k = kb;
    while (k < ke) {
      double absakk = MAG(P[k + k * nrow]);
      double colmax_sq = 0.0;
      int64_t imax = k;
      for (int64_t i = k + 1; i < ke; ++i) {
        double m = MAGSQ(P[k * nrow + i]);
        if (m > colmax_sq) { colmax_sq = m; imax = i; }
      }
      double colmax = sqrt(colmax_sq);
      int kstep;
      int64_t kp;
      if ((absakk > colmax ? absakk : colmax) == 0.0) {
        if (!have_floor) { status = 1; break; }
        kstep = 1; kp = k;
      } else if (absakk >= alpha * colmax) {
        kstep = 1; kp = k;
      } else {
        double rowmax_sq = 0.0;
        for (int64_t j = k; j < imax; ++j) {
          double m = MAGSQ(P[j * nrow + imax]);
          if (m > rowmax_sq) rowmax_sq = m;
        }
        for (int64_t i = imax + 1; i < ke; ++i) {
          double m = MAGSQ(P[imax * nrow + i]);
          if (m > rowmax_sq) rowmax_sq = m;
        }
        double rowmax = sqrt(rowmax_sq);
        if (absakk >= alpha * colmax * (colmax / rowmax)) { kstep = 1; kp = k; }
        else if (MAG(P[imax * nrow + imax]) >= alpha * rowmax) { kstep = 1; kp = imax; }
        else { kstep = 2; kp = imax; }
      }
      if (kstep == 1) {
        if (kp != k) {
          FN(swap_sym)(P, nrow, k, kp, ke);
          int64_t t = lperm[k]; lperm[k] = lperm[kp]; lperm[kp] = t;
          deep[k - kb] = kp;
        }
        T dk = P[k + k * nrow];
        if (have_floor) {
          if (MAG(dk) < floor) { dk = FN(perturb)(dk, floor); P[k + k * nrow] = dk; nperturbed++; }
        } else if (IS_ZERO(dk)) { status = 1; break; }
        d[k] = dk;
        T dinv = RECIP(dk);
        for (int64_t j = k + 1; j < ke; ++j) {
          T m = P[k * nrow + j] * dinv;
          snap[(k - kb) * nb + (j - kb)] = m;
          if (!IS_ZERO(m))
            for (int64_t i = j; i < ke; ++i) P[j * nrow + i] = P[j * nrow + i] - P[k * nrow + i] * m;
        }
        for (int64_t i = k + 1; i < ke; ++i) P[k * nrow + i] = P[k * nrow + i] * dinv;
        k += 1;
      } else {
        if (kp != k + 1) {
          FN(swap_sym)(P, nrow, k + 1, kp, ke);
          int64_t t = lperm[k + 1]; lperm[k + 1] = lperm[kp]; lperm[kp] = t;
          deep[k - kb] = kp;
        }
        T d11 = P[k + k * nrow], d21 = P[k * nrow + k + 1], d22 = P[(k + 1) + (k + 1) * nrow];
        T det = d11 * d22 - d21 * d21;
        double scale = MAG(d11);
        if (MAG(d22) > scale) scale = MAG(d22);
        if (MAG(d21) > scale) scale = MAG(d21);
        double growth_floor = 1e-14 * scale * scale;
        if (have_floor) {
          double fl = floor * floor > growth_floor ? floor * floor : growth_floor;
          if (MAG(det) < fl) {
            double lift = floor > scale * sqrt(1e-14) ? floor : scale * sqrt(1e-14);
            d11 = d11 + FROM_REAL(lift);
            d22 = d22 + FROM_REAL(lift);
            det = d11 * d22 - d21 * d21;
            if (MAG(det) < fl) det = det + FROM_REAL(fl);
            nperturbed++;
          }
        } else if (MAG(det) <= growth_floor) { status = 1; break; }
        T detinv = RECIP(det);
        d[k] = d11; dsub[k] = d21; d[k + 1] = d22; two[k] = 1;
        for (int64_t i = k + 2; i < ke; ++i) {
          T wik = P[k * nrow + i], wik1 = P[(k + 1) * nrow + i];
          l1[i] = (d22 * wik - d21 * wik1) * detinv;
          l2[i] = (d11 * wik1 - d21 * wik) * detinv;
          snap[(k - kb) * nb + (i - kb)] = l1[i];
          snap[(k + 1 - kb) * nb + (i - kb)] = l2[i];
        }
        for (int64_t j = k + 2; j < ke; ++j)
          for (int64_t i = j; i < ke; ++i)
            P[j * nrow + i] = P[j * nrow + i] - P[k * nrow + i] * l1[j] - P[(k + 1) * nrow + i] * l2[j];
        for (int64_t i = k + 2; i < ke; ++i) { P[k * nrow + i] = l1[i]; P[(k + 1) * nrow + i] = l2[i]; }
        k += 2;
      }
    }
    if (status) break;
    /* deep-row replay: every row on its own, so row blocks run concurrently */
    if (nrow > ke) {
      if (X->inner && nrow - ke > 256) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(ora_threads())
        for (int64_t r0 = ke; r0 < nrow; r0 += 128)
          FN(apply_trailing)(P, nrow, kb, ke, d, dsub, two, deep, snap, nb, r0, r0 + 128 < nrow ? r0 + 128 : nrow);
      } else {
        FN(apply_trailing)(P, nrow, kb, ke, d, dsub, two, deep, snap, nb, ke, nrow);
      }
    }
    /* deferred Schur: P[ke.., ke..ncol] -= L21 D R^T (lower part); a column
     * cc reads only the panel's columns and g, and writes only column ke + cc */
    int64_t pw = ke - kb, cw = ncol - ke, mt = nrow - ke;
    if (pw > 0 && cw > 0 && mt > 0) {
      T *g = malloc(sizeof(T) * mt * pw);
      /* G = L21 * D with L21 = P[ke.., kb..ke) (the panel is contiguous) */
      FN(apply_d)(P + kb * nrow, nrow, ke, mt, pw, d + kb, dsub + kb, two + kb, g);
      if (X->inner && cw > 1) {
#pragma omp parallel for schedule(dynamic, 2) num_threads(ora_threads())
        for (int64_t cc = 0; cc < cw; ++cc) FN(schur_col)(P, nrow, kb, ke, pw, mt, g, cc);
      } else {
        for (int64_t cc = 0; cc < cw; ++cc) FN(schur_col)(P, nrow, kb, ke, pw, mt, g, cc);
      }
      free(g);
    }
    kb = ke;
  }
  for (int64_t li = 0; li < nrow; ++li) gloc[rs[sn][li]] = -1;
  free(l1); free(l2); free(deep); free(snap);
  if (status) { free(P); free(d); free(dsub); free(two); free(lperm); goto out; }
  /* emit maps + inertia */
  int64_t eoff = e_off[sn];
  for (int64_t pp = 0; pp < ncol;) {
    int64_t g = rs[sn][lperm[pp]], e = eoff + pp;
    e_of_g[g] = e;
    perm_out[e] = s->perm[g];
    d_out[e] = d[pp];
    if (two[pp]) {
      int64_t g2 = rs[sn][lperm[pp + 1]];
      e_of_g[g2] = e + 1;
      perm_out[e + 1] = s->perm[g2];
      d_out[e + 1] = d[pp + 1];
      ds_out[e] = dsub[pp];
      two_out[e] = 1;
      double det_r = REAL(d[pp] * d[pp + 1] - dsub[pp] * dsub[pp]);
      double tr_r = REAL(d[pp] + d[pp + 1]);
      if (det_r < 0.0) { ipos++; ineg++; }
      else if (det_r > 0.0) { if (tr_r >= 0.0) ipos += 2; else ineg += 2; }
      else { izero++; if (tr_r >= 0.0) ipos++; else ineg++; }
      pp += 2;
    } else {
      double r = REAL(d[pp]);
      if (r > 0.0) ipos++; else if (r < 0.0) ineg++; else izero++;
      pp += 1;
    }
  }
  nodes[sn].panel = P; nodes[sn].d = d; nodes[sn].dsub = dsub;
  nodes[sn].two = two; nodes[sn].lperm = lperm;
out:
  W->gbuf = gbuf; W->ubuf = ubuf; W->fbuf = fbuf; W->gcap = gcap; W->ucap = ucap;
  W->cnt[0] += ipos; W->cnt[1] += ineg; W->cnt[2] += izero; W->cnt[3] += nperturbed;
  return status;
}

/* Left-looking factor; returns 0 or 1 (rank deficient). */
static int FN(factor)(const ora_sym *s, const int64_t *a_colp, const int32_t *a_rowi,
                      const T *a_val, int panel_nb, int pivot, double abs_floor,
                      double drop_tol, ora_factors *out) {
  const int64_t n = s->n, ns = s->nsuper;
  /* perturb floor */
  int have_floor = 0;
  double floor = 0.0;
  if (pivot == 1) { have_floor = 1; floor = abs_floor > 0 ? abs_floor : 0.0; }
  if (pivot == 2) {
    double amax = 0.0;
    for (int64_t k = 0; k < a_colp[n]; ++k) if (MAG(a_val[k]) > amax) amax = MAG(a_val[k]);
    have_floor = 1;
    floor = (amax > 1.0 ? amax : 1.0) * 2.220446049250313e-16;
  }
  if (panel_nb < 8) panel_nb = 8;

  /* A_perm = P^T A P lower (from_triplets: bucket by column, sort, sum dups) */
  int64_t nnz = a_colp[n];
  int64_t *pcnt = calloc(n + 1, sizeof(int64_t));
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = a_colp[j]; k < a_colp[j + 1]; ++k) {
      int64_t gi = s->perm_inv[a_rowi[k]], gj = s->perm_inv[j];
      pcnt[(gi >= gj ? gj : gi) + 1]++;
    }
  for (int64_t j = 0; j < n; ++j) pcnt[j + 1] += pcnt[j];
  int64_t *prow = malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  T *pval = malloc(sizeof(T) * (nnz ? nnz : 1));
  int64_t *pos = malloc(sizeof(int64_t) * (n ? n : 1));
  for (int64_t j = 0; j < n; ++j) pos[j] = pcnt[j];
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = a_colp[j]; k < a_colp[j + 1]; ++k) {
      int64_t gi = s->perm_inv[a_rowi[k]], gj = s->perm_inv[j];
      int64_t r = gi >= gj ? gi : gj, c = gi >= gj ? gj : gi;
      prow[pos[c]] = r;
      pval[pos[c]] = a_val[k];
      pos[c]++;
    }
  for (int64_t j = 0; j < n; ++j) { /* insertion sort: columns are short */
    for (int64_t k = pcnt[j] + 1; k < pcnt[j + 1]; ++k) {
      int64_t r = prow[k];
      T v = pval[k];
      int64_t q = k - 1;
      while (q >= pcnt[j] && prow[q] > r) { prow[q + 1] = prow[q]; pval[q + 1] = pval[q]; --q; }
      prow[q + 1] = r;
      pval[q + 1] = v;
    }
  }
  /* duplicates (cannot occur for a validated input, summed for fidelity) */

  /* row structures */
  int64_t **rs = malloc(sizeof(int64_t *) * (ns ? ns : 1));
  int64_t *rsn = malloc(sizeof(int64_t) * (ns ? ns : 1));
  int64_t *tbuf = malloc(sizeof(int64_t) * (n + 1));
  uint8_t *mark = calloc(n ? n : 1, 1);
  for (int64_t sn = 0; sn < ns; ++sn) {
    int64_t first = s->sn_first[sn], nc = s->sn_ncol[sn], own_last = first + nc, nt = 0;
    for (int64_t j = first; j < own_last; ++j)
      for (int64_t k = s->pp_colp[j]; k < s->pp_colp[j + 1]; ++k) {
        int64_t r = s->pp_rowi[k];
        if (r >= own_last && !mark[r]) { mark[r] = 1; tbuf[nt++] = r; }
      }
    for (int64_t c = s->ch_ptr[sn]; c < s->ch_ptr[sn + 1]; ++c) {
      int64_t ch = s->ch_idx[c], nck = s->sn_ncol[ch];
      for (int64_t q = nck; q < rsn[ch]; ++q) {
        int64_t r = rs[ch][q];
        if (r >= own_last && !mark[r]) { mark[r] = 1; tbuf[nt++] = r; }
      }
    }
    /* sort trailing rows */
    {
      int64_t gap;
      for (gap = nt / 2; gap > 0; gap /= 2)
        for (int64_t a = gap; a < nt; ++a) {
          int64_t v = tbuf[a], b = a;
          while (b >= gap && tbuf[b - gap] > v) { tbuf[b] = tbuf[b - gap]; b -= gap; }
          tbuf[b] = v;
        }
    }
    rsn[sn] = nc + nt;
    rs[sn] = malloc(sizeof(int64_t) * (nc + nt + 1));
    for (int64_t j = 0; j < nc; ++j) rs[sn][j] = first + j;
    for (int64_t j = 0; j < nt; ++j) { rs[sn][nc + j] = tbuf[j]; mark[tbuf[j]] = 0; }
  }
  /* update lists (ascending updater index) */
  int64_t *col2s = malloc(sizeof(int64_t) * (n ? n : 1));
  for (int64_t sn = 0; sn < ns; ++sn)
    for (int64_t c = 0; c < s->sn_ncol[sn]; ++c) col2s[s->sn_first[sn] + c] = sn;
  int64_t *ucnt = calloc(ns + 1, sizeof(int64_t));
  for (int64_t k = 0; k < ns; ++k) {
    int64_t last = -1;
    for (int64_t q = s->sn_ncol[k]; q < rsn[k]; ++q) {
      int64_t t = col2s[rs[k][q]];
      if (t != last) { ucnt[t + 1]++; last = t; }
    }
  }
  for (int64_t t = 0; t < ns; ++t) ucnt[t + 1] += ucnt[t];
  int64_t *uidx = malloc(sizeof(int64_t) * (ucnt[ns] ? ucnt[ns] : 1));
  int64_t *upos = malloc(sizeof(int64_t) * (ns ? ns : 1));
  for (int64_t t = 0; t < ns; ++t) upos[t] = ucnt[t];
  for (int64_t k = 0; k < ns; ++k) {
    int64_t last = -1;
    for (int64_t q = s->sn_ncol[k]; q < rsn[k]; ++q) {
      int64_t t = col2s[rs[k][q]];
      if (t != last) { uidx[upos[t]++] = k; last = t; }
    }
  }

  /* e numbering and emit maps */
  int64_t *e_off = malloc(sizeof(int64_t) * (ns ? ns : 1));
  { int64_t acc = 0; for (int64_t sn = 0; sn < ns; ++sn) { e_off[sn] = acc; acc += s->sn_ncol[sn]; } }
  int64_t *e_of_g = malloc(sizeof(int64_t) * (n ? n : 1));
  T *d_out = calloc(n ? n : 1, sizeof(T));
  T *ds_out = calloc(n ? n : 1, sizeof(T));
  uint8_t *two_out = calloc(n ? n : 1, 1);
  int32_t *perm_out = calloc(n ? n : 1, sizeof(int32_t));
  int64_t ipos = 0, ineg = 0, izero = 0, nperturbed = 0;

  FN(node) *nodes = calloc(ns ? ns : 1, sizeof(FN(node)));
  int status = 0;

  /* Levels of the supernode tree (leaves 0): a supernode's updaters are its
   * descendants, so the supernodes of one level are independent. Wide levels
   * run their supernodes on all threads; narrow ones run them one at a time
   * with the loops inside each supernode on all threads. Every entry is
   * computed by the same operations in the same order as in the sequential
   * postorder, so the factor is the same bit for bit. */
  int64_t *lev = calloc(ns ? ns : 1, sizeof(int64_t)), nlev = 0;
  for (int64_t sn = 0; sn < ns; ++sn) {
    int64_t l = 0;
    for (int64_t c = s->ch_ptr[sn]; c < s->ch_ptr[sn + 1]; ++c)
      if (lev[s->ch_idx[c]] + 1 > l) l = lev[s->ch_idx[c]] + 1;
    lev[sn] = l;
    if (l + 1 > nlev) nlev = l + 1;
  }
  int64_t *lptr = calloc(nlev + 2, sizeof(int64_t)), *lord = malloc(sizeof(int64_t) * (ns ? ns : 1));
  for (int64_t sn = 0; sn < ns; ++sn) lptr[lev[sn] + 2]++;
  for (int64_t l = 0; l < nlev; ++l) lptr[l + 2] += lptr[l + 1];
  for (int64_t sn = 0; sn < ns; ++sn) lord[lptr[lev[sn] + 1]++] = sn;
  /* small matrices stay on one thread: starting parallel regions costs more
   * than their work */
  const int nthr = n >= 20000 ? ora_threads() : 1;
  FN(work) *works = calloc(nthr, sizeof(FN(work)));
  for (int t = 0; t < nthr; ++t) {
    works[t].gloc = malloc(sizeof(int64_t) * (n ? n : 1));
    for (int64_t i = 0; i < n; ++i) works[t].gloc[i] = -1;
  }
  FN(ctx) X = {s, n, ns, panel_nb, have_floor, floor, pcnt, prow, pval, rs, rsn, ucnt, uidx, e_off,
               e_of_g, d_out, ds_out, two_out, perm_out, nodes, 0};
  for (int64_t l = 0; l < nlev && status == 0; ++l) {
    const int64_t b = lptr[l], e = lptr[l + 1];
    int lstat = 0;
    if (nthr > 1 && e - b >= nthr) {
      X.inner = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthr) reduction(max : lstat)
      for (int64_t q = b; q < e; ++q) {
        int st = FN(factor_node)(&X, &works[ora_thread()], lord[q]);
        if (st > lstat) lstat = st;
      }
    } else {
      X.inner = nthr > 1;
      for (int64_t q = b; q < e && lstat == 0; ++q) lstat = FN(factor_node)(&X, &works[0], lord[q]);
    }
    status = lstat;
  }
  for (int t = 0; t < nthr; ++t) {
    ipos += works[t].cnt[0]; ineg += works[t].cnt[1]; izero += works[t].cnt[2]; nperturbed += works[t].cnt[3];
    free(works[t].gloc); free(works[t].gbuf); free(works[t].ubuf); free(works[t].fbuf);
  }
  free(works); free(lev); free(lptr); free(lord);

  if (status == 0) {
    /* compact L per supernode, concatenated in supernode order */
    int64_t cap = 0;
    for (int64_t sn = 0; sn < ns; ++sn) cap += rsn[sn] * s->sn_ncol[sn];
    out->l_col_ptr = malloc(sizeof(int64_t) * (n + 1));
    out->l_row_idx = malloc(sizeof(int32_t) * (cap ? cap : 1));
    T *lv = malloc(sizeof(T) * (cap ? cap : 1));
    int64_t *ci = malloc(sizeof(int64_t) * (n + 1));
    T *cv = malloc(sizeof(T) * (n + 1));
    int64_t nz = 0, col = 0;
    out->l_col_ptr[0] = 0;
    for (int64_t sn = 0; sn < ns; ++sn) {
      const FN(node) *K = &nodes[sn];
      int64_t ncol = s->sn_ncol[sn], nrow = rsn[sn];
      for (int64_t p = 0; p < ncol; ++p) {
        int64_t m = 0, diag_e = e_of_g[rs[sn][K->lperm[p]]];
        ci[m] = diag_e; cv[m] = ONE; m++;
        int64_t i0 = K->two[p] ? p + 2 : p + 1;
        double colmax = 0.0;
        for (int64_t i = i0; i < nrow; ++i) {
          T v = K->panel[i + p * nrow];
          if (!IS_ZERO(v)) {
            ci[m] = e_of_g[rs[sn][K->lperm[i]]]; cv[m] = v; m++;
            if (MAG(v) > colmax) colmax = MAG(v);
          }
        }
        if (drop_tol >= 0.0) {
          double thr = drop_tol * colmax;
          int64_t w = 1;
          for (int64_t q = 1; q < m; ++q)
            if (MAG(cv[q]) >= thr) { ci[w] = ci[q]; cv[w] = cv[q]; w++; }
          m = w;
        }
        for (int64_t a = 1; a < m; ++a) { /* insertion sort by row */
          int64_t r = ci[a]; T v = cv[a]; int64_t b = a - 1;
          while (b >= 0 && ci[b] > r) { ci[b + 1] = ci[b]; cv[b + 1] = cv[b]; --b; }
          ci[b + 1] = r; cv[b + 1] = v;
        }
        for (int64_t q = 0; q < m; ++q) { out->l_row_idx[nz] = (int32_t)ci[q]; lv[nz] = cv[q]; nz++; }
        out->l_col_ptr[++col] = nz;
      }
    }
    free(ci); free(cv);
    out->nnzl = nz;
    out->l_values = (double *)lv;
    out->d_diag = (double *)d_out;
    out->d_subdiag = (double *)ds_out;
    out->two_by_two = two_out;
    out->perm = perm_out;
    out->n_perturbed = nperturbed;
    out->inertia[0] = ipos; out->inertia[1] = ineg; out->inertia[2] = izero;
  } else {
    free(d_out); free(ds_out); free(two_out); free(perm_out);
  }
  out->status = status;
  for (int64_t sn = 0; sn < ns; ++sn) {
    free(nodes[sn].panel); free(nodes[sn].d); free(nodes[sn].dsub);
    free(nodes[sn].two); free(nodes[sn].lperm); free(rs[sn]);
  }
  free(nodes); free(rs); free(rsn); free(tbuf); free(mark); free(col2s); free(ucnt);
  free(uidx); free(upos); free(e_off); free(e_of_g);
  free(pcnt); free(prow); free(pval); free(pos);
  return status;
}

/* solve_ldlt_block (ldlt_generic.rs:618-714); b, x row-major n x nrhs */
static int FN(solve_many)(const ora_factors *f, const T *b, int64_t nrhs, T *x) {
  const int64_t n = f->n;
  const T *lv = (const T *)f->l_values, *dd = (const T *)f->d_diag, *ds = (const T *)f->d_subdiag;
  T *y = malloc(sizeof(T) * (n * nrhs > 0 ? n * nrhs : 1));
  T *row = malloc(sizeof(T) * (nrhs ? nrhs : 1));
  for (int64_t i = 0; i < n; ++i)
    for (int64_t c = 0; c < nrhs; ++c) y[i * nrhs + c] = b[(int64_t)f->perm[i] * nrhs + c];
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t c = 0; c < nrhs; ++c) row[c] = y[j * nrhs + c];
    for (int64_t k = f->l_col_ptr[j] + 1; k < f->l_col_ptr[j + 1]; ++k) {
      int64_t i = f->l_row_idx[k];
      T nl = ZERO - lv[k];
      for (int64_t c = 0; c < nrhs; ++c) y[i * nrhs + c] = FMADD(nl, row[c], y[i * nrhs + c]);
    }
  }
  for (int64_t k = 0; k < n;) {
    if (f->two_by_two[k]) {
      T d11 = dd[k], d21 = ds[k], d22 = dd[k + 1];
      T det = d11 * d22 - d21 * d21;
      if (IS_ZERO(det)) { free(y); free(row); return 1; }
      T detinv = RECIP(det);
      for (int64_t c = 0; c < nrhs; ++c) {
        T z0 = y[k * nrhs + c], z1 = y[(k + 1) * nrhs + c];
        y[k * nrhs + c] = (d22 * z0 - d21 * z1) * detinv;
        y[(k + 1) * nrhs + c] = (d11 * z1 - d21 * z0) * detinv;
      }
      k += 2;
    } else {
      if (IS_ZERO(dd[k])) { free(y); free(row); return 1; }
      T dinv = RECIP(dd[k]);
      for (int64_t c = 0; c < nrhs; ++c) y[k * nrhs + c] = y[k * nrhs + c] * dinv;
      k += 1;
    }
  }
  for (int64_t j = n - 1; j >= 0; --j) {
    for (int64_t c = 0; c < nrhs; ++c) row[c] = y[j * nrhs + c];
    for (int64_t k = f->l_col_ptr[j] + 1; k < f->l_col_ptr[j + 1]; ++k) {
      int64_t i = f->l_row_idx[k];
      T nl = ZERO - lv[k];
      for (int64_t c = 0; c < nrhs; ++c) row[c] = FMADD(nl, y[i * nrhs + c], row[c]);
    }
    for (int64_t c = 0; c < nrhs; ++c) y[j * nrhs + c] = row[c];
  }
  for (int64_t i = 0; i < n; ++i)
    for (int64_t c = 0; c < nrhs; ++c) x[(int64_t)f->perm[i] * nrhs + c] = y[i * nrhs + c];
  free(y); free(row);
  return 0;
}

static void FN(onepass_scale)(int64_t n, const int64_t *colp, const int32_t *rowi, const T *val,
                              double *scale) {
  for (int64_t i = 0; i < n; ++i) scale[i] = 0.0;
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = colp[j]; k < colp[j + 1]; ++k) {
      int64_t i = rowi[k];
      double m = MAG(val[k]);
      if (m > scale[i]) scale[i] = m;
      if (i != j && m > scale[j]) scale[j] = m;
    }
  for (int64_t i = 0; i < n; ++i) {
    double m = scale[i], out = 1.0;
    if (m > 0.0) {
      double cand = 1.0 / sqrt(m);
      if (isfinite(cand) && cand > 0.0) out = cand;
    }
    scale[i] = out;
  }
}

static void FN(symv)(int64_t n, const int64_t *colp, const int32_t *rowi, const T *val,
                     const T *x, T *y) {
  for (int64_t i = 0; i < n; ++i) y[i] = ZERO;
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = colp[j]; k < colp[j + 1]; ++k) {
      int64_t i = rowi[k];
      T v = val[k];
      y[i] = y[i] + v * x[j];
      if (i != j) y[j] = y[j] + v * x[i];
    }
}

static double FN(norm_inf)(int64_t n, const int64_t *colp, const int32_t *rowi, const T *val) {
  double *rsum = calloc(n ? n : 1, sizeof(double)), best = 0.0;
  for (int64_t c = 0; c < n; ++c)
    for (int64_t k = colp[c]; k < colp[c + 1]; ++k) {
      int64_t r = rowi[k];
      double m = MAG(val[k]);
      rsum[r] += m;
      if (r != c) rsum[c] += m;
    }
  for (int64_t i = 0; i < n; ++i) if (rsum[i] > best) best = rsum[i];
  free(rsum);
  return best;
}

#undef FN
#undef FN1
#undef FN2
