// Approximate Minimum Degree ordering (Amestoy, Davis & Duff, SIMAX 17(4)
// 1996; ACM TOMS 30(3) 2004), quotient-graph form with aggressive absorption,
// mass elimination, supervariable detection and dense-row deferral, followed by
// the big-child-last assembly-tree postorder.
//
// rslab's ordering (crates/rslab-amd/src/lib.rs:108-136 driving
// crates/rslab-ordering-core/src/quotient_graph/{workspace.rs:132-299,
// algo.rs:59-632, algo.rs:1362-1560}) reproduces SuiteSparse AMD's amd_2
// step for step; the permutation therefore depends on every tie-break
// (LIFO degree lists, hash buckets, list layout), and this file follows the
// same published sequence of steps so the output permutation is identical.
// Workspace sizing (iwlen = nzaat + nzaat/5 + n) follows workspace.rs:172-181,
// which only changes how often the list storage is compacted (ncmpa).
#include <cmath>
#include <cstdint>
#include <stdexcept>

#include "symbolic.hpp"

namespace rsl {

namespace {
constexpr int32_t NONE = -1;
inline int32_t flip(int32_t x) { return -2 - x; }

inline int32_t clear_flag(int32_t wflg, int32_t wbig, std::vector<int32_t>& w) {
  if (wflg < 2 || wflg >= wbig) {
    for (auto& x : w)
      if (x != 0) x = 1;
    return 2;
  }
  return wflg;
}
}  // namespace

std::vector<idx> amd_order(const Pattern& full, AmdStats* stats) {
  const int32_t n = full.n;
  std::vector<idx> perm(n);
  if (n == 0) return perm;

  // ---- workspace (workspace.rs:132-299) ----
  std::vector<int32_t> len(n, 0);
  int64_t nzaat = 0;
  for (int32_t j = 0; j < n; ++j) {
    int32_t c = 0;
    for (off k = full.colp[j]; k < full.colp[j + 1]; ++k)
      if (full.rowi[k] != j) ++c;
    len[j] = c;
    nzaat += c;
  }
  int64_t iwlen64 = nzaat + nzaat / 5 + n;
  if (iwlen64 > INT32_MAX) throw std::runtime_error("amd: pattern too large");
  const int32_t iwlen = (int32_t)iwlen64;
  std::vector<int32_t> iw(iwlen > 0 ? iwlen : 1, 0), pe(n, 0);
  int32_t pfree = 0;
  for (int32_t j = 0; j < n; ++j) {
    pe[j] = pfree;
    for (off k = full.colp[j]; k < full.colp[j + 1]; ++k)
      if (full.rowi[k] != j) iw[pfree++] = full.rowi[k];
  }
  int64_t dense = (int64_t)(10.0 * std::sqrt((double)n));
  if (dense < 16) dense = 16;
  if (dense > n) dense = n;

  std::vector<int32_t> nv(n, 1), elen(n, 0), w(n, 1), degree(len), head(n, NONE),
      next(n, NONE), last(n, NONE);
  const int32_t wbig = INT32_MAX - n;
  int32_t wflg = 0, lemax = 0, mindeg = 0, ncmpa = 0, nel = 0, ndense = 0;
  for (int32_t i = 0; i < n; ++i) {
    int32_t deg = degree[i];
    if (deg == 0) {
      elen[i] = flip(1);
      nel++;
      pe[i] = NONE;
      w[i] = 0;
    } else if (deg > dense) {
      ndense++;
      nv[i] = 0;
      elen[i] = NONE;
      pe[i] = NONE;
      nel++;
    } else {
      int32_t inext = head[deg];
      if (inext != NONE) last[inext] = i;
      next[i] = inext;
      head[deg] = i;
    }
  }

  double lnz = 0, ndiv = 0, nms_lu = 0, nms_ldl = 0, dmax = 1;

  // ---- elimination loop ----
  while (nel < n) {
    // select pivot of minimum approximate degree
    int32_t deg = mindeg, me = NONE;
    for (; deg < n; ++deg) {
      me = head[deg];
      if (me != NONE) break;
    }
    if (me == NONE) break;
    mindeg = deg;
    {
      int32_t inext = next[me];
      if (inext != NONE) last[inext] = NONE;
      head[deg] = inext;
    }
    const int32_t elenme = elen[me];
    int32_t nvpiv = nv[me];
    nel += nvpiv;

    // construct the new element in place or at the end of iw
    nv[me] = -nvpiv;
    int32_t degme = 0, pme1, pme2;
    auto remove_from_degree_list = [&](int32_t i) {
      int32_t ilast = last[i], inext = next[i];
      if (inext != NONE) last[inext] = ilast;
      if (ilast != NONE) next[ilast] = inext;
      else head[degree[i]] = inext;
    };
    if (elenme == 0) {
      pme1 = pe[me];
      pme2 = pme1 - 1;
      for (int32_t p = pme1; p <= pme1 + len[me] - 1; ++p) {
        int32_t i = iw[p], nvi = nv[i];
        if (nvi > 0) {
          degme += nvi;
          nv[i] = -nvi;
          iw[++pme2] = i;
          remove_from_degree_list(i);
        }
      }
    } else {
      int32_t p = pe[me];
      pme1 = pfree;
      const int32_t slenme = len[me] - elenme;
      for (int32_t knt1 = 1; knt1 <= elenme + 1; ++knt1) {
        int32_t e, pj, ln;
        if (knt1 > elenme) {
          e = me;
          pj = p;
          ln = slenme;
        } else {
          e = iw[p++];
          pj = pe[e];
          ln = len[e];
        }
        for (int32_t knt2 = 1; knt2 <= ln; ++knt2) {
          int32_t i = iw[pj++], nvi = nv[i];
          if (nvi > 0) {
            if (pfree >= iwlen) {
              // compress iw: list heads are tagged by flipping their
              // first entry, then every live list is slid to the front
              pe[me] = p;
              len[me] -= knt1;
              if (len[me] == 0) pe[me] = NONE;
              pe[e] = pj;
              len[e] = ln - knt2;
              if (len[e] == 0) pe[e] = NONE;
              ncmpa++;
              for (int32_t j = 0; j < n; ++j) {
                int32_t pn = pe[j];
                if (pn >= 0) {
                  pe[j] = iw[pn];
                  iw[pn] = flip(j);
                }
              }
              int32_t psrc = 0, pdst = 0, pend = pme1 - 1;
              while (psrc <= pend) {
                int32_t j = flip(iw[psrc++]);
                if (j >= 0) {
                  iw[pdst] = pe[j];
                  pe[j] = pdst++;
                  int32_t lenj = len[j];
                  for (int32_t knt3 = 0; knt3 <= lenj - 2; ++knt3) iw[pdst++] = iw[psrc++];
                }
              }
              int32_t p1 = pdst;
              for (psrc = pme1; psrc <= pfree - 1; ++psrc) iw[pdst++] = iw[psrc];
              pme1 = p1;
              pfree = pdst;
              pj = pe[e];
              p = pe[me];
            }
            degme += nvi;
            nv[i] = -nvi;
            iw[pfree++] = i;
            remove_from_degree_list(i);
          }
        }
        if (e != me) {
          pe[e] = flip(me);  // e absorbed into me
          w[e] = 0;
        }
      }
      pme2 = pfree - 1;
    }
    degree[me] = degme;
    pe[me] = pme1;
    len[me] = pme2 - pme1 + 1;
    elen[me] = flip(nvpiv + degme);
    wflg = clear_flag(wflg, wbig, w);

    // |Le \ Lme| for every element e adjacent to a variable of me
    for (int32_t pme = pme1; pme <= pme2; ++pme) {
      int32_t i = iw[pme], eln = elen[i];
      if (eln > 0) {
        int32_t nvi = -nv[i], wnvi = wflg - nvi;
        for (int32_t p = pe[i]; p <= pe[i] + eln - 1; ++p) {
          int32_t e = iw[p], we = w[e];
          if (we >= wflg) we -= nvi;
          else if (we != 0) we = degree[e] + wnvi;
          w[e] = we;
        }
      }
    }

    // degree update and (aggressive) element absorption
    for (int32_t pme = pme1; pme <= pme2; ++pme) {
      int32_t i = iw[pme];
      int32_t p1 = pe[i], p2 = p1 + elen[i] - 1, pn = p1;
      uint64_t hash = 0;
      int32_t dg = 0;
      for (int32_t p = p1; p <= p2; ++p) {
        int32_t e = iw[p], we = w[e];
        if (we != 0) {
          int32_t dext = we - wflg;
          if (dext > 0) {
            dg += dext;
            iw[pn++] = e;
            hash += (uint64_t)e;
          } else {
            pe[e] = flip(me);
            w[e] = 0;
          }
        }
      }
      elen[i] = pn - p1 + 1;
      int32_t p3 = pn, p4 = p1 + len[i];
      for (int32_t p = p2 + 1; p < p4; ++p) {
        int32_t j = iw[p], nvj = nv[j];
        if (nvj > 0) {
          dg += nvj;
          iw[pn++] = j;
          hash += (uint64_t)j;
        }
      }
      if (elen[i] == 1 && p3 == pn) {
        // mass elimination: i is indistinguishable from me
        pe[i] = flip(me);
        int32_t nvi = -nv[i];
        degme -= nvi;
        nvpiv += nvi;
        nel += nvi;
        nv[i] = 0;
        elen[i] = NONE;
      } else {
        if (dg < degree[i]) degree[i] = dg;
        iw[pn] = iw[p3];
        iw[p3] = iw[p1];
        iw[p1] = me;
        len[i] = pn - p1 + 1;
        int32_t h = (int32_t)(hash % (uint64_t)n);
        int32_t j = head[h];
        if (j <= NONE) {
          next[i] = flip(j);
          head[h] = flip(i);
        } else {
          next[i] = last[j];
          last[j] = i;
        }
        last[i] = h;
      }
    }
    degree[me] = degme;
    if (degme > lemax) lemax = degme;
    wflg += lemax;
    wflg = clear_flag(wflg, wbig, w);

    // supervariable detection
    for (int32_t pme = pme1; pme <= pme2; ++pme) {
      int32_t i = iw[pme];
      if (nv[i] < 0) {
        int32_t h = last[i];
        int32_t j = head[h];
        if (j == NONE) {
          i = NONE;
        } else if (j < NONE) {
          i = flip(j);
          head[h] = NONE;
        } else {
          i = last[j];
          last[j] = NONE;
        }
        while (i != NONE && next[i] != NONE) {
          int32_t ln = len[i], eln = elen[i];
          for (int32_t p = pe[i] + 1; p <= pe[i] + ln - 1; ++p) w[iw[p]] = wflg;
          int32_t jlast = i;
          j = next[i];
          while (j != NONE) {
            bool ok = len[j] == ln && elen[j] == eln;
            for (int32_t p = pe[j] + 1; ok && p <= pe[j] + ln - 1; ++p)
              if (w[iw[p]] != wflg) ok = false;
            if (ok) {
              pe[j] = flip(i);
              nv[i] += nv[j];
              nv[j] = 0;
              elen[j] = NONE;
              j = next[j];
              next[jlast] = j;
            } else {
              jlast = j;
              j = next[j];
            }
          }
          wflg++;
          i = next[i];
        }
      }
    }

    // restore degree lists, compact me's variable list
    int32_t p = pme1;
    const int32_t nleft = n - nel;
    for (int32_t pme = pme1; pme <= pme2; ++pme) {
      int32_t i = iw[pme], nvi = -nv[i];
      if (nvi > 0) {
        nv[i] = nvi;
        int32_t d = degree[i] + degme - nvi;
        if (d > nleft - nvi) d = nleft - nvi;
        int32_t inext = head[d];
        if (inext != NONE) last[inext] = i;
        next[i] = inext;
        last[i] = NONE;
        head[d] = i;
        if (d < mindeg) mindeg = d;
        degree[i] = d;
        iw[p++] = i;
      }
    }
    nv[me] = nvpiv;
    len[me] = p - pme1;
    if (len[me] == 0) {
      pe[me] = NONE;
      w[me] = 0;
    }
    if (elenme != 0) pfree = p;

    double f = nvpiv, r = (double)degme + ndense;
    if (f + r > dmax) dmax = f + r;
    double lnzme = f * r + (f - 1) * f / 2;
    lnz += lnzme;
    ndiv += lnzme;
    double s = f * r * r + r * (f - 1) * f + (f - 1) * f * (2 * f - 1) / 6;
    nms_lu += s;
    nms_ldl += (s + lnzme) / 2;
  }
  if (ndense > 0) {
    double f = ndense;
    if (f > dmax) dmax = f;
    double lnzme = (f - 1) * f / 2;
    lnz += lnzme;
    ndiv += lnzme;
    double s = (f - 1) * f * (2 * f - 1) / 6;
    nms_lu += s;
    nms_ldl += (s + lnzme) / 2;
  }

  // ---- final permutation (algo.rs:1362-1560) ----
  for (int32_t i = 0; i < n; ++i) {
    pe[i] = flip(pe[i]);
    elen[i] = flip(elen[i]);
  }
  for (int32_t i = 0; i < n; ++i) {
    if (nv[i] != 0) continue;
    int32_t j = pe[i];
    if (j == NONE) continue;  // dense row
    while (nv[j] == 0) j = pe[j];
    int32_t e = j;
    j = i;
    while (nv[j] == 0) {
      int32_t jn = pe[j];
      pe[j] = e;
      j = jn;
    }
  }
  // assembly-tree postorder: child lists ascending, largest front moved last
  std::vector<int32_t>& child = head;
  std::vector<int32_t>& sib = next;
  std::vector<int32_t>& stack = last;
  std::vector<int32_t>& order = w;
  for (int32_t j = 0; j < n; ++j) {
    child[j] = NONE;
    sib[j] = NONE;
  }
  for (int32_t j = n - 1; j >= 0; --j) {
    if (nv[j] > 0) {
      int32_t par = pe[j];
      if (par >= 0 && par < n) {
        sib[j] = child[par];
        child[par] = j;
      }
    }
  }
  for (int32_t i = 0; i < n; ++i) {
    if (nv[i] > 0 && child[i] != NONE) {
      int32_t fprev = NONE, bigfprev = NONE, bigf = NONE, maxfr = NONE;
      for (int32_t f = child[i]; f != NONE; f = sib[f]) {
        int32_t fr = elen[f];
        if (fr >= maxfr) {
          maxfr = fr;
          bigfprev = fprev;
          bigf = f;
        }
        fprev = f;
      }
      int32_t fnext = sib[bigf];
      if (fnext != NONE) {
        if (bigfprev == NONE) child[i] = fnext;
        else sib[bigfprev] = fnext;
        sib[bigf] = NONE;
        sib[fprev] = bigf;
      }
    }
  }
  for (int32_t j = 0; j < n; ++j) order[j] = NONE;
  int32_t k = 0;
  for (int32_t root = 0; root < n; ++root) {
    if (pe[root] != NONE || nv[root] <= 0) continue;
    int32_t top = 1;
    stack[0] = root;
    while (top > 0) {
      int32_t i = stack[top - 1];
      if (child[i] != NONE) {
        int32_t cnt = 0;
        for (int32_t f = child[i]; f != NONE; f = sib[f]) cnt++;
        int32_t t = top + cnt;
        for (int32_t f = child[i]; f != NONE; f = sib[f]) stack[--t] = f;
        top += cnt;
        child[i] = NONE;
      } else {
        top--;
        order[i] = k++;
      }
    }
  }
  std::vector<int32_t> pos(n, NONE), start(n, NONE);
  for (int32_t e = 0; e < n; ++e)
    if (order[e] != NONE) pos[order[e]] = e;
  int32_t cnt = 0;
  for (int32_t q = 0; q < n; ++q) {
    int32_t e = pos[q];
    if (e == NONE) break;
    start[e] = cnt;
    cnt += nv[e];
  }
  for (int32_t i = 0; i < n; ++i) {
    if (nv[i] == 0) {
      int32_t e = pe[i];
      if (e != NONE) {
        start[i] = start[e];
        start[e]++;
      } else {
        start[i] = cnt++;
      }
    }
  }
  for (int32_t i = 0; i < n; ++i) perm[start[i]] = i;

  if (stats) {
    stats->ncmpa = ncmpa;
    stats->ndense = ndense;
    stats->lnz = lnz;
    stats->ndiv = ndiv;
    stats->nms_lu = nms_lu;
    stats->nms_ldl = nms_ldl;
    stats->dmax = dmax;
  }
  return perm;
}

}  // namespace rsl
