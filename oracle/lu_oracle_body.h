/*
 * Field-generic body of the LU oracle, included once per field by
 * lu_oracle.c with T, SFX and the scalar helpers defined. TEST
 * INFRASTRUCTURE ONLY. The reference functions each part follows are listed
 * in lu_oracle.h.
 */

#define LFN2(a, b) a##_##b
#define LFN1(a, b) LFN2(a, b)
#define LFN(name) LFN1(name, SFX)

/* perturb_pivot (multifrontal_ldlt.rs:786-795) */
static T LFN(lu_perturb)(T d, double floor) {
  double mag = MAG(d);
  if (mag >= floor) return d;
  if (mag == 0.0) return FROM_REAL(floor);
  return d * FROM_REAL(floor / mag);
}

typedef struct {
  T *l;           /* nrow x ncol, column-major: L below the diagonal, U11 above */
  T *u;           /* ncol x (nrow - ncol), column-major: U12 */
  int64_t *rperm; /* row-structure index physically at each panel row */
} LFN(lu_node);

typedef struct {
  int64_t r;
  T v;
} LFN(lu_ent);

static int LFN(lu_ent_cmp)(const void *a, const void *b) {
  int64_t x = ((const LFN(lu_ent) *)a)->r, y = ((const LFN(lu_ent) *)b)->r;
  return x < y ? -1 : x > y;
}

/* CSC of the n x n matrix with entries (row[k], col[k], val[k]), rows sorted
 * within a column (GeneralCsc::from_triplets; no duplicates arise here). */
static void LFN(lu_csc)(int64_t n, int64_t nnz, const int64_t *row, const int64_t *col, const T *val,
                        int64_t **cp_out, int64_t **ri_out, T **v_out) {
  int64_t *cp = calloc(n + 1, sizeof(int64_t));
  for (int64_t k = 0; k < nnz; ++k) cp[col[k] + 1]++;
  for (int64_t j = 0; j < n; ++j) cp[j + 1] += cp[j];
  int64_t *pos = malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t j = 0; j < n; ++j) pos[j] = cp[j];
  int64_t *ri = malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  T *v = malloc(sizeof(T) * (nnz ? nnz : 1));
  for (int64_t k = 0; k < nnz; ++k) {
    int64_t q = pos[col[k]]++;
    ri[q] = row[k];
    v[q] = val[k];
  }
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = cp[j] + 1; k < cp[j + 1]; ++k) {
      int64_t r = ri[k];
      T x = v[k];
      int64_t q = k - 1;
      while (q >= cp[j] && ri[q] > r) { ri[q + 1] = ri[q]; v[q + 1] = v[q]; --q; }
      ri[q + 1] = r;
      v[q + 1] = x;
    }
  free(pos);
  *cp_out = cp;
  *ri_out = ri;
  *v_out = v;
}

static int LFN(lu_factor)(const ora_sym *s, const int64_t *a_colp, const int32_t *a_rowi, const T *a_val,
                          int pivot, double abs_floor, double pivot_u, double drop_tol, oru_factors *out) {
  const int64_t n = s->n, ns = s->nsuper, nnz = a_colp[n];
  /* ZeroPivotAction (multifrontal_lu.rs:2716-2723): ForceAccept's floor is
   * taken over the unscaled values */
  int have_floor = 0;
  double floor = 0.0;
  if (pivot == 1) { have_floor = 1; floor = abs_floor > 0 ? abs_floor : 0.0; }
  if (pivot == 2) {
    double amax = 0.0;
    for (int64_t k = 0; k < nnz; ++k) if (MAG(a_val[k]) > amax) amax = MAG(a_val[k]);
    have_floor = 1;
    floor = (amax > 1.0 ? amax : 1.0) * 2.220446049250313e-16;
  }
  if (pivot_u < 0.0) pivot_u = 0.0;
  if (pivot_u > 1.0) pivot_u = 1.0;
  const double thresh_sq = pivot_u * pivot_u;

  /* two-sided equilibration from the original matrix (:2733-2758) */
  double *d_row = malloc(sizeof(double) * (n ? n : 1)), *d_col = malloc(sizeof(double) * (n ? n : 1));
  {
    double *rmax = calloc(n ? n : 1, sizeof(double)), *cmax = calloc(n ? n : 1, sizeof(double));
    for (int64_t j = 0; j < n; ++j)
      for (int64_t k = a_colp[j]; k < a_colp[j + 1]; ++k) {
        double m = MAG(a_val[k]);
        if (m > rmax[a_rowi[k]]) rmax[a_rowi[k]] = m;
        if (m > cmax[j]) cmax[j] = m;
      }
    for (int64_t i = 0; i < n; ++i) {
      double cr = rmax[i] > 0.0 ? 1.0 / sqrt(rmax[i]) : 1.0;
      d_row[i] = (isfinite(cr) && cr > 0.0) ? cr : 1.0;
      double cc = cmax[i] > 0.0 ? 1.0 / sqrt(cmax[i]) : 1.0;
      d_col[i] = (isfinite(cc) && cc > 0.0) ? cc : 1.0;
    }
    free(rmax);
    free(cmax);
  }
  /* A_perm = P^T (D_r A D_c) P and its transpose (:2760-2778) */
  int64_t *ap_cp, *ap_ri, *at_cp, *at_ri;
  T *ap_v, *at_v;
  {
    int64_t *tr = malloc(sizeof(int64_t) * (nnz ? nnz : 1)), *tc = malloc(sizeof(int64_t) * (nnz ? nnz : 1));
    T *tv = malloc(sizeof(T) * (nnz ? nnz : 1));
    int64_t q = 0;
    for (int64_t j = 0; j < n; ++j)
      for (int64_t k = a_colp[j]; k < a_colp[j + 1]; ++k, ++q) {
        int64_t i = a_rowi[k];
        tr[q] = s->perm_inv[i];
        tc[q] = s->perm_inv[j];
        tv[q] = a_val[k] * FROM_REAL(d_row[i] * d_col[j]);
      }
    LFN(lu_csc)(n, nnz, tr, tc, tv, &ap_cp, &ap_ri, &ap_v);
    LFN(lu_csc)(n, nnz, tc, tr, tv, &at_cp, &at_ri, &at_v);
    free(tr);
    free(tc);
    free(tv);
  }

  /* supernode row structures: own columns, then the sorted trailing rows of
   * the own columns' pattern and of the children's trailing rows */
  int64_t **rs = malloc(sizeof(int64_t *) * (ns ? ns : 1));
  int64_t *rsn = malloc(sizeof(int64_t) * (ns ? ns : 1));
  {
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
      for (int64_t gap = nt / 2; gap > 0; gap /= 2)
        for (int64_t a = gap; a < nt; ++a) {
          int64_t v = tbuf[a], b = a;
          while (b >= gap && tbuf[b - gap] > v) { tbuf[b] = tbuf[b - gap]; b -= gap; }
          tbuf[b] = v;
        }
      rsn[sn] = nc + nt;
      rs[sn] = malloc(sizeof(int64_t) * (nc + nt + 1));
      for (int64_t j = 0; j < nc; ++j) rs[sn][j] = first + j;
      for (int64_t j = 0; j < nt; ++j) { rs[sn][nc + j] = tbuf[j]; mark[tbuf[j]] = 0; }
    }
    free(tbuf);
    free(mark);
  }
  /* update lists: descendants k whose trailing rows hit s's columns,
   * ascending k (factor_lu_left_looking :2444-2455) */
  int64_t *col2s = malloc(sizeof(int64_t) * (n ? n : 1));
  for (int64_t sn = 0; sn < ns; ++sn)
    for (int64_t c = 0; c < s->sn_ncol[sn]; ++c) col2s[s->sn_first[sn] + c] = sn;
  int64_t *uptr = calloc(ns + 1, sizeof(int64_t));
  for (int64_t k = 0; k < ns; ++k) {
    int64_t last = -1;
    for (int64_t q = s->sn_ncol[k]; q < rsn[k]; ++q) {
      int64_t t = col2s[rs[k][q]];
      if (t != last) { uptr[t + 1]++; last = t; }
    }
  }
  for (int64_t sn = 0; sn < ns; ++sn) uptr[sn + 1] += uptr[sn];
  int64_t *uidx = malloc(sizeof(int64_t) * (uptr[ns] ? uptr[ns] : 1));
  {
    int64_t *fill = malloc(sizeof(int64_t) * (ns ? ns : 1));
    for (int64_t sn = 0; sn < ns; ++sn) fill[sn] = uptr[sn];
    for (int64_t k = 0; k < ns; ++k) {
      int64_t last = -1;
      for (int64_t q = s->sn_ncol[k]; q < rsn[k]; ++q) {
        int64_t t = col2s[rs[k][q]];
        if (t != last) { uidx[fill[t]++] = k; last = t; }
      }
    }
    free(fill);
  }

  LFN(lu_node) *nodes = calloc(ns ? ns : 1, sizeof(LFN(lu_node)));
  int64_t *gloc = malloc(sizeof(int64_t) * (n ? n : 1));
  for (int64_t i = 0; i < n; ++i) gloc[i] = -1;
  int64_t *row_pos = malloc(sizeof(int64_t) * (n ? n : 1));
  int32_t *perm = malloc(sizeof(int32_t) * (n ? n : 1)), *perm_row = malloc(sizeof(int32_t) * (n ? n : 1));
  int64_t n_pert = 0;
  int status = 0;
  T pinv_blk[128];

  for (int64_t sn = 0; sn < ns && !status; ++sn) {
    const int64_t first = s->sn_first[sn], ncol = s->sn_ncol[sn], nrow = rsn[sn], cnrow = nrow - ncol;
    const int64_t nl_el = nrow * ncol, nu_el = ncol * cnrow;
    T *lb = calloc(nl_el > 0 ? (size_t)nl_el : 1, sizeof(T));
    T *ub = calloc(nu_el > 0 ? (size_t)nu_el : 1, sizeof(T));
    int64_t *rperm = malloc(sizeof(int64_t) * nrow);
    for (int64_t i = 0; i < nrow; ++i) { rperm[i] = i; gloc[rs[sn][i]] = i; }
    /* assembly: the columns of s into lb, the U12 rows into ub (:1670-1686) */
    for (int64_t p = 0; p < ncol; ++p) {
      const int64_t c = first + p;
      for (int64_t k = ap_cp[c]; k < ap_cp[c + 1]; ++k) {
        int64_t li = gloc[ap_ri[k]];
        if (li >= 0) lb[p * nrow + li] = lb[p * nrow + li] + ap_v[k];
      }
      for (int64_t k = at_cp[c]; k < at_cp[c + 1]; ++k) {
        int64_t lc = gloc[at_ri[k]];
        if (lc >= ncol) ub[p + (lc - ncol) * ncol] = ub[p + (lc - ncol) * ncol] + at_v[k];
      }
    }
    /* cmod from every updater, ascending (:1881-2002, the scalar loop's order) */
    for (int64_t u = uptr[sn]; u < uptr[sn + 1]; ++u) {
      const int64_t kk = uidx[u], nck = s->sn_ncol[kk], nrk = rsn[kk];
      const int64_t *ok = rs[kk] + nck;
      const int64_t nok = nrk - nck;
      int64_t p0 = 0, p1;
      while (p0 < nok && ok[p0] < first) ++p0;
      p1 = p0;
      while (p1 < nok && ok[p1] < first + ncol) ++p1;
      const int64_t npk = p1 - p0, mrows = nok - p0, ntrail = nok - p1;
      if (npk == 0) continue;
      const T *lk = nodes[kk].l, *uk = nodes[kk].u;
      for (int64_t jj = 0; jj < npk; ++jj) {
        const int64_t tcol = ok[p0 + jj] - first;
        for (int64_t i = 0; i < mrows; ++i) {
          T acc = ZERO;
          for (int64_t ck = 0; ck < nck; ++ck) acc = acc + lk[(nck + p0 + i) + ck * nrk] * uk[ck + (p0 + jj) * nck];
          const int64_t trow = gloc[ok[p0 + i]];
          lb[tcol * nrow + trow] = lb[tcol * nrow + trow] - acc;
        }
      }
      for (int64_t jj = 0; jj < ntrail; ++jj) {
        const int64_t tu = gloc[ok[p1 + jj]] - ncol;
        for (int64_t i = 0; i < npk; ++i) {
          T acc = ZERO;
          for (int64_t ck = 0; ck < nck; ++ck) acc = acc + lk[(nck + p0 + i) + ck * nrk] * uk[ck + (p1 + jj) * nck];
          const int64_t ur = ok[p0 + i] - first;
          ub[ur + tu * ncol] = ub[ur + tu * ncol] - acc;
        }
      }
    }
    /* cdiv: blocked panel LU with threshold partial pivoting restricted to
     * the fully-summed rows [0, ncol) (:2005-2283) */
    const int64_t nb = ncol >= 512 ? 128 : 32;
    for (int64_t kb = 0; kb < ncol && !status; kb += nb) {
      const int64_t ke = kb + nb < ncol ? kb + nb : ncol, pw = ke - kb;
      for (int64_t k = kb; k < ke; ++k) {
        if (thresh_sq > 0.0) {
          int64_t p = k;
          double best = MAGSQ(lb[k * nrow + k]);
          for (int64_t i = k + 1; i < ncol; ++i) {
            double m = MAGSQ(lb[k * nrow + i]);
            if (m > best) { best = m; p = i; }
          }
          double diag_sq = MAGSQ(lb[k * nrow + k]);
          if (p != k && diag_sq < thresh_sq * best) {
            for (int64_t c = 0; c < ncol; ++c) { T t = lb[c * nrow + k]; lb[c * nrow + k] = lb[c * nrow + p]; lb[c * nrow + p] = t; }
            for (int64_t t = 0; t < cnrow; ++t) { T x = ub[k + t * ncol]; ub[k + t * ncol] = ub[p + t * ncol]; ub[p + t * ncol] = x; }
            int64_t x = rperm[k]; rperm[k] = rperm[p]; rperm[p] = x;
          }
        }
        T piv = lb[k * nrow + k];
        if (have_floor && MAG(piv) < floor) {
          piv = LFN(lu_perturb)(piv, floor);
          n_pert++;
        } else if (!have_floor && IS_ZERO(piv)) {
          status = 1;
          break;
        }
        lb[k * nrow + k] = piv;
        T pinv = RECIP(piv);
        pinv_blk[k - kb] = pinv;
        for (int64_t i = k + 1; i < ncol; ++i) lb[k * nrow + i] = lb[k * nrow + i] * pinv;
        for (int64_t j = k + 1; j < ke; ++j) {
          T ukj = lb[j * nrow + k];
          if (!IS_ZERO(ukj))
            for (int64_t i = k + 1; i < ncol; ++i) lb[j * nrow + i] = lb[j * nrow + i] - lb[k * nrow + i] * ukj;
        }
      }
      if (status) break;
      /* deep rows [ncol, nrow): apply_panel_trailing (:1357-1387) */
      for (int64_t q = 0; q < pw; ++q) {
        const int64_t k = kb + q;
        T *ck = lb + k * nrow;
        for (int64_t i = ncol; i < nrow; ++i) ck[i] = ck[i] * pinv_blk[q];
        for (int64_t jj = q + 1; jj < pw; ++jj) {
          T ukj = lb[(kb + jj) * nrow + k];
          if (!IS_ZERO(ukj)) {
            T *cj = lb + (kb + jj) * nrow;
            for (int64_t i = ncol; i < nrow; ++i) cj[i] = cj[i] - ck[i] * ukj;
          }
        }
      }
      /* TRSM: rows [kb, ke) of the later panel columns and of U12 */
      for (int64_t j = ke; j < ncol; ++j)
        for (int64_t r = kb + 1; r < ke; ++r) {
          T acc = lb[j * nrow + r];
          for (int64_t i = kb; i < r; ++i) acc = acc - lb[i * nrow + r] * lb[j * nrow + i];
          lb[j * nrow + r] = acc;
        }
      for (int64_t t = 0; t < cnrow; ++t) {
        T *col = ub + t * ncol;
        for (int64_t r = kb + 1; r < ke; ++r) {
          T acc = col[r];
          for (int64_t i = kb; i < r; ++i) acc = acc - lb[i * nrow + r] * col[i];
          col[r] = acc;
        }
      }
      /* GEMMs: lb[ke.., ke..ncol) -= L[ke.., kb..ke) U[kb..ke, ke..ncol);
       * ub[ke..ncol, :] -= L[ke..ncol, kb..ke) U12[kb..ke, :] */
      for (int64_t c = ke; c < ncol; ++c)
        for (int64_t r = ke; r < nrow; ++r) {
          T acc = ZERO;
          for (int64_t q = kb; q < ke; ++q) acc = acc + lb[q * nrow + r] * lb[c * nrow + q];
          lb[c * nrow + r] = lb[c * nrow + r] - acc;
        }
      for (int64_t t = 0; t < cnrow; ++t)
        for (int64_t r = ke; r < ncol; ++r) {
          T acc = ZERO;
          for (int64_t q = kb; q < ke; ++q) acc = acc + lb[q * nrow + r] * ub[q + t * ncol];
          ub[r + t * ncol] = ub[r + t * ncol] - acc;
        }
    }
    for (int64_t i = 0; i < nrow; ++i) gloc[rs[sn][i]] = -1;
    /* index maps (:2328-2344); the column of position e is e itself */
    for (int64_t p = 0; p < ncol; ++p) {
      const int64_t g_col = first + p, g_row = rs[sn][rperm[p]];
      row_pos[g_row] = first + p;
      perm[first + p] = s->perm[g_col];
      perm_row[first + p] = s->perm[g_row];
    }
    nodes[sn].l = lb;
    nodes[sn].u = ub;
    nodes[sn].rperm = rperm;
  }

  out->status = status;
  out->n_perturbed = n_pert;
  out->perm = perm;
  out->perm_row = perm_row;
  out->d_row = d_row;
  out->d_col = d_col;
  if (!status) {
    /* emit (emit_and_free :1525-1618): L columns and U rows of each
     * supernode in elimination order */
    int64_t cap_l = n + 1, cap_u = n + 1;
    for (int64_t sn = 0; sn < ns; ++sn) {
      int64_t nc = s->sn_ncol[sn], nr = rsn[sn];
      cap_l += nc * nr;
      cap_u += nc * nr;
    }
    int64_t *lcp = malloc(sizeof(int64_t) * (n + 1)), *ucp = malloc(sizeof(int64_t) * (n + 1));
    int32_t *lri = malloc(sizeof(int32_t) * cap_l), *uci = malloc(sizeof(int32_t) * cap_u);
    T *lv = malloc(sizeof(T) * cap_l), *uv = malloc(sizeof(T) * cap_u);
    LFN(lu_ent) *buf = malloc(sizeof(LFN(lu_ent)) * (n + 1));
    int64_t nl = 0, nu = 0;
    lcp[0] = ucp[0] = 0;
    for (int64_t sn = 0; sn < ns; ++sn) {
      const int64_t first = s->sn_first[sn], ncol = s->sn_ncol[sn], nrow = rsn[sn], cnrow = nrow - ncol;
      const T *lb = nodes[sn].l, *ub = nodes[sn].u;
      const int64_t *rperm = nodes[sn].rperm;
      for (int64_t p = 0; p < ncol; ++p) {
        const int64_t de = first + p;
        int64_t m = 0;
        buf[m].r = de; buf[m].v = ONE; ++m;
        for (int64_t i = p + 1; i < nrow; ++i) {
          T v = lb[p * nrow + i];
          if (!IS_ZERO(v)) { buf[m].r = row_pos[rs[sn][rperm[i]]]; buf[m].v = v; ++m; }
        }
        if (drop_tol >= 0.0) {
          double mx = 0.0;
          for (int64_t q = 0; q < m; ++q) if (buf[q].r != de && MAG(buf[q].v) > mx) mx = MAG(buf[q].v);
          double thr = drop_tol * mx;
          int64_t w = 0;
          for (int64_t q = 0; q < m; ++q) if (buf[q].r == de || MAG(buf[q].v) >= thr) buf[w++] = buf[q];
          m = w;
        }
        qsort(buf, m, sizeof(buf[0]), LFN(lu_ent_cmp));
        for (int64_t q = 0; q < m; ++q) { lri[nl] = (int32_t)buf[q].r; lv[nl] = buf[q].v; ++nl; }
        lcp[de + 1] = nl;
        m = 0;
        buf[m].r = de; buf[m].v = lb[p * nrow + p]; ++m;
        for (int64_t j = p + 1; j < ncol; ++j) {
          T v = lb[j * nrow + p];
          if (!IS_ZERO(v)) { buf[m].r = first + j; buf[m].v = v; ++m; }
        }
        for (int64_t t = 0; t < cnrow; ++t) {
          T v = ub[p + t * ncol];
          if (!IS_ZERO(v)) { buf[m].r = rs[sn][ncol + t]; buf[m].v = v; ++m; }
        }
        if (drop_tol >= 0.0) {
          double mx = 0.0;
          for (int64_t q = 0; q < m; ++q) if (buf[q].r != de && MAG(buf[q].v) > mx) mx = MAG(buf[q].v);
          double thr = drop_tol * mx;
          int64_t w = 0;
          for (int64_t q = 0; q < m; ++q) if (buf[q].r == de || MAG(buf[q].v) >= thr) buf[w++] = buf[q];
          m = w;
        }
        qsort(buf, m, sizeof(buf[0]), LFN(lu_ent_cmp));
        for (int64_t q = 0; q < m; ++q) { uci[nu] = (int32_t)buf[q].r; uv[nu] = buf[q].v; ++nu; }
        ucp[de + 1] = nu;
      }
    }
    free(buf);
    out->nnz_l = nl;
    out->nnz_u = nu;
    out->l_col_ptr = lcp;
    out->l_row_idx = lri;
    out->l_values = (double *)lv;
    out->u_row_ptr = ucp;
    out->u_col_idx = uci;
    out->u_values = (double *)uv;
  }
  for (int64_t sn = 0; sn < ns; ++sn) {
    free(nodes[sn].l);
    free(nodes[sn].u);
    free(nodes[sn].rperm);
    free(rs[sn]);
  }
  free(nodes);
  free(rs);
  free(rsn);
  free(col2s);
  free(uptr);
  free(uidx);
  free(gloc);
  free(row_pos);
  free(ap_cp);
  free(ap_ri);
  free(ap_v);
  free(at_cp);
  free(at_ri);
  free(at_v);
  return status;
}

/* solve_lu_block (multifrontal_lu.rs:3209-3283) */
static void LFN(lu_solve)(const oru_factors *f, const T *b, int64_t nrhs, T *x) {
  const int64_t n = f->n;
  const T *lv = (const T *)f->l_values, *uv = (const T *)f->u_values;
  const int64_t ny = n * nrhs;
  T *y = malloc(sizeof(T) * (ny > 0 ? (size_t)ny : 1));
  T *row = malloc(sizeof(T) * (nrhs ? nrhs : 1));
  for (int64_t e = 0; e < n; ++e) {
    const int64_t orig = f->perm_row[e];
    for (int64_t c = 0; c < nrhs; ++c) y[e * nrhs + c] = b[orig * nrhs + c] * FROM_REAL(f->d_row[orig]);
  }
  for (int64_t e = 0; e < n; ++e) {
    for (int64_t c = 0; c < nrhs; ++c) row[c] = y[e * nrhs + c];
    for (int64_t k = f->l_col_ptr[e] + 1; k < f->l_col_ptr[e + 1]; ++k) {
      const int64_t i = f->l_row_idx[k];
      const T nl = ZERO - lv[k];
      for (int64_t c = 0; c < nrhs; ++c) y[i * nrhs + c] = FMADD(nl, row[c], y[i * nrhs + c]);
    }
  }
  for (int64_t e = n - 1; e >= 0; --e) {
    for (int64_t c = 0; c < nrhs; ++c) row[c] = y[e * nrhs + c];
    const int64_t s0 = f->u_row_ptr[e];
    for (int64_t k = s0 + 1; k < f->u_row_ptr[e + 1]; ++k) {
      const T nu = ZERO - uv[k];
      const int64_t cb = f->u_col_idx[k];
      for (int64_t c = 0; c < nrhs; ++c) row[c] = FMADD(nu, y[cb * nrhs + c], row[c]);
    }
    const T dinv = RECIP(uv[s0]);
    for (int64_t c = 0; c < nrhs; ++c) y[e * nrhs + c] = row[c] * dinv;
  }
  for (int64_t e = 0; e < n; ++e) {
    const int64_t orig = f->perm[e];
    for (int64_t c = 0; c < nrhs; ++c) x[orig * nrhs + c] = y[e * nrhs + c] * FROM_REAL(f->d_col[orig]);
  }
  free(y);
  free(row);
}

#undef LFN
#undef LFN1
#undef LFN2
