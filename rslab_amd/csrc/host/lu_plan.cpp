// Execution plan of the LU twin: storage offsets, the A -> L / U12 scatter
// map, the cmod target tiles of both regions with their updater tasks, the
// per-level panel steps, and the solve's gather lists (the solve's block
// steps depend on the field and are laid out by the factor).
//
// Panel width: the reference blocks its cdiv in panels of 32 columns, 128
// from 512 columns up (multifrontal_lu.rs:2020). The pivot search spans every
// own row below the diagonal, so each column's rank-1 update inside a panel
// sweeps all those rows for every later column of the panel; on the device
// one workgroup runs a panel, and 16-column panels keep that sweep short
// while the trailing updates run as multi-workgroup GEMMs. The blocking
// changes only the order of the sums, not the pivot rule.
#include "lu_plan.hpp"

#include <algorithm>
#include <stdexcept>

namespace rsl {

namespace {
constexpr int TILE = 64;

struct Run {
  int32_t tile, lo, hi;
};

// consecutive indices [lo, hi) grouped by key(i) / TILE (keys ascend with i)
template <class Key>
void runs_by_tile(int32_t lo, int32_t hi, Key key, std::vector<Run>& out) {
  out.clear();
  for (int32_t i = lo; i < hi; ++i) {
    const int32_t t = key(i) / TILE;
    if (out.empty() || out.back().tile != t) out.push_back({t, i, i + 1});
    else out.back().hi = i + 1;
  }
}
}  // namespace

void symmetrized_lower(int32_t n, const int64_t* colp, const int32_t* rowi, std::vector<int64_t>& scp,
                       std::vector<int32_t>& sri) {
  std::vector<int64_t> cnt(n + 1, 0);
  for (int32_t j = 0; j < n; ++j)
    for (int64_t k = colp[j]; k < colp[j + 1]; ++k) cnt[std::min(rowi[k], j) + 1]++;
  for (int32_t j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
  std::vector<int32_t> hi(cnt[n]);
  std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
  for (int32_t j = 0; j < n; ++j)
    for (int64_t k = colp[j]; k < colp[j + 1]; ++k) {
      const int32_t i = rowi[k];
      hi[pos[std::min(i, j)]++] = std::max(i, j);
    }
  scp.assign(n + 1, 0);
  sri.clear();
  sri.reserve(hi.size());
  for (int32_t j = 0; j < n; ++j) {
    std::sort(hi.begin() + cnt[j], hi.begin() + cnt[j + 1]);
    int32_t last = -1;
    for (int64_t k = cnt[j]; k < cnt[j + 1]; ++k)
      if (hi[k] != last) {
        sri.push_back(hi[k]);
        last = hi[k];
      }
    scp[j + 1] = (int64_t)sri.size();
  }
}

void build_lu_plan(LuPlan& P, const int64_t* colp, const int32_t* rowi) {
  const Symbolic& S = P.sym;
  const int32_t n = S.n, ns = (int32_t)S.snodes.size();
  P.n = n;
  P.ns = ns;
  P.nnz = colp[n];
  P.colp.assign(colp, colp + n + 1);
  P.rowi.assign(rowi, rowi + P.nnz);
  std::vector<std::vector<idx>> rs = supernode_row_structures(S);
  std::vector<std::vector<idx>> ul = update_lists(S, rs);

  P.sn_first.resize(ns);
  P.sn_ncol.resize(ns);
  P.sn_nrow.resize(ns);
  P.nb.resize(ns);
  P.rs_off.assign(ns + 1, 0);
  P.l_off.resize(ns);
  P.u_off.resize(ns);
  P.s_off.resize(ns);
  std::vector<int32_t> level(ns, 0);
  int32_t maxlev = 0;
  P.l_total = P.u_total = P.s_total = 0;
  P.factor_flops = 0;
  P.l_entries = P.u_entries = 0;
  for (int32_t s = 0; s < ns; ++s) {
    const Supernode& sn = S.snodes[s];
    const int64_t h = (int64_t)rs[s].size(), w = sn.ncol;
    P.sn_first[s] = sn.first_col;
    P.sn_ncol[s] = sn.ncol;
    P.sn_nrow[s] = (int32_t)h;
    P.nb[s] = LU_PANEL;
    P.rs_off[s + 1] = P.rs_off[s] + h;
    P.l_off[s] = P.l_total;
    P.l_total += h * w;
    P.u_off[s] = P.u_total;
    P.u_total += w * (h - w);
    P.s_off[s] = P.s_total;
    P.s_total += h - w;
    for (int64_t k = 0; k < w; ++k) {
      const double r = (double)(h - k - 1);
      P.factor_flops += 2.0 * r * r + r;
      P.l_entries += h - k;
      P.u_entries += h - k;
    }
    int32_t lv = 0;
    for (idx c : sn.children) lv = std::max(lv, level[c] + 1);
    level[s] = lv;
    maxlev = std::max(maxlev, lv);
  }
  P.rs_idx.resize(P.rs_off[ns]);
  for (int32_t s = 0; s < ns; ++s) std::copy(rs[s].begin(), rs[s].end(), P.rs_idx.begin() + P.rs_off[s]);
  P.rperm_init.resize(n);
  for (int32_t s = 0; s < ns; ++s)
    for (int32_t p = 0; p < P.sn_ncol[s]; ++p) P.rperm_init[P.sn_first[s] + p] = p;

  // A -> storage: entry (i, j) in the permuted numbering lands in the L panel
  // of the supernode owning column j when i is in its rows (the lower part
  // and U11), otherwise in the U12 block of the supernode owning row i
  std::vector<int32_t> col2s(n);
  for (int32_t s = 0; s < ns; ++s)
    for (int32_t c = 0; c < P.sn_ncol[s]; ++c) col2s[P.sn_first[s] + c] = s;
  auto row_pos = [&](int32_t s, int32_t g) -> int64_t {  // position of g in rs[s], -1 if absent
    const int32_t* b = P.rs_idx.data() + P.rs_off[s];
    const int32_t* e = P.rs_idx.data() + P.rs_off[s + 1];
    const int32_t* it = std::lower_bound(b, e, g);
    return (it != e && *it == g) ? (int64_t)(it - b) : -1;
  };
  P.amap.resize(P.nnz);
  for (int32_t j = 0; j < n; ++j)
    for (int64_t k = colp[j]; k < colp[j + 1]; ++k) {
      const int32_t gi = S.perm_inv[rowi[k]], gj = S.perm_inv[j];
      const int32_t sj = col2s[gj];
      const int64_t li = row_pos(sj, gi);
      if (li >= 0) {
        P.amap[k] = P.l_off[sj] + (int64_t)(gj - P.sn_first[sj]) * P.sn_nrow[sj] + li;
        continue;
      }
      const int32_t si = col2s[gi];
      const int64_t lc = row_pos(si, gj);
      if (lc < P.sn_ncol[si]) throw std::runtime_error("lu plan: entry outside the symbolic structure");
      P.amap[k] = -1 - (P.u_off[si] + (gi - P.sn_first[si]) + (lc - P.sn_ncol[si]) * P.sn_ncol[si]);
    }

  // CSR of A in the user numbering, for the refinement's residuals
  P.csr_ptr.assign((size_t)n + 1, 0);
  for (int64_t k = 0; k < P.nnz; ++k) P.csr_ptr[rowi[k] + 1]++;
  for (int32_t i = 0; i < n; ++i) P.csr_ptr[i + 1] += P.csr_ptr[i];
  P.csr_col.resize(P.nnz);
  P.csr_src.resize(P.nnz);
  {
    std::vector<int64_t> fill(P.csr_ptr.begin(), P.csr_ptr.end() - 1);
    for (int32_t j = 0; j < n; ++j)
      for (int64_t k = colp[j]; k < colp[j + 1]; ++k) {
        const int64_t q = fill[rowi[k]]++;
        P.csr_col[q] = j;
        P.csr_src[q] = k;
      }
  }

  // levels
  P.levels.assign(ns ? maxlev + 1 : 0, LuLevelPlan());
  for (int32_t s = 0; s < ns; ++s) P.levels[level[s]].all.push_back(s);
  P.level_lists.clear();
  P.level_off.assign(P.levels.size() + 1, 0);
  for (size_t l = 0; l < P.levels.size(); ++l) {
    P.level_off[l] = (int64_t)P.level_lists.size();
    P.level_lists.insert(P.level_lists.end(), P.levels[l].all.begin(), P.levels[l].all.end());
  }
  P.level_off[P.levels.size()] = (int64_t)P.level_lists.size();

  // cmod: per target, L-region tiles (rows over the h positions, columns over
  // the w own positions) and U12 tiles (rows over the own positions, columns
  // over the h - w trailing positions); a task covers the updater rows that
  // land in one tile row window and one tile column window
  P.items.clear();
  P.tasks.clear();
  std::vector<int32_t> gloc(n, -1);
  std::vector<std::vector<CTask>> tl, tu;
  std::vector<Run> cr, rr, tr;
  for (LuLevelPlan& L : P.levels) {
    L.item_beg = (int64_t)P.items.size();
    for (int32_t s : L.all) {
      if (ul[s].empty()) continue;
      const int32_t h = P.sn_nrow[s], w = P.sn_ncol[s], first = P.sn_first[s];
      const int32_t* rss = P.rs_idx.data() + P.rs_off[s];
      for (int32_t i = 0; i < h; ++i) gloc[rss[i]] = i;
      const int32_t nrt = (h + TILE - 1) / TILE, nct = (w + TILE - 1) / TILE;
      const int32_t nut = (h - w + TILE - 1) / TILE;
      tl.assign((size_t)nrt * nct, {});
      tu.assign((size_t)nct * std::max(nut, 1), {});
      for (idx kk : ul[s]) {
        const int32_t k = (int32_t)kk, nck = P.sn_ncol[k];
        const int32_t* ok = P.rs_idx.data() + P.rs_off[k] + nck;
        const int32_t nok = P.sn_nrow[k] - nck;
        const int32_t p0 = (int32_t)(std::lower_bound(ok, ok + nok, first) - ok);
        const int32_t p1 = (int32_t)(std::lower_bound(ok, ok + nok, first + w) - ok);
        if (p0 == p1) continue;
        for (int32_t a = p0; a < nok; ++a)
          if (gloc[ok[a]] < 0) throw std::runtime_error("lu plan: updater row outside target structure");
        runs_by_tile(p0, p1, [&](int32_t b) { return ok[b] - first; }, cr);
        runs_by_tile(p0, nok, [&](int32_t a) { return gloc[ok[a]]; }, rr);
        runs_by_tile(p1, nok, [&](int32_t b) { return gloc[ok[b]] - w; }, tr);
        for (const Run& c : cr) {
          uint64_t cm = 0;
          for (int32_t b = c.lo; b < c.hi; ++b) cm |= 1ull << (ok[b] - first - c.tile * TILE);
          for (const Run& r : rr) {
            CTask tk{k, r.lo, r.hi, c.lo, c.hi, 0, 0, cm, 0};
            for (int32_t a = r.lo; a < r.hi; ++a) tk.rmask |= 1ull << (gloc[ok[a]] - r.tile * TILE);
            tl[(size_t)r.tile * nct + c.tile].push_back(tk);
            L.cmod_flops += 2.0 * nck * (double)(r.hi - r.lo) * (c.hi - c.lo);
          }
          for (const Run& t : tr) {
            CTask tk{k, c.lo, c.hi, t.lo, t.hi, 0, cm, 0, 0};
            for (int32_t b = t.lo; b < t.hi; ++b) tk.cmask |= 1ull << (gloc[ok[b]] - w - t.tile * TILE);
            tu[(size_t)c.tile * nut + t.tile].push_back(tk);
            L.cmod_flops += 2.0 * nck * (double)(c.hi - c.lo) * (t.hi - t.lo);
          }
        }
      }
      auto emit_items = [&](std::vector<std::vector<CTask>>& tt, int32_t nr, int32_t nc, int region) {
        for (int32_t rb = 0; rb < nr; ++rb)
          for (int32_t cb = 0; cb < nc; ++cb) {
            auto& v = tt[(size_t)rb * nc + cb];
            if (v.empty()) continue;
            LuItem it{s, rb * TILE, cb * TILE, (int32_t)P.tasks.size(), 0, region};
            P.tasks.insert(P.tasks.end(), v.begin(), v.end());
            it.t_end = (int32_t)P.tasks.size();
            P.items.push_back(it);
          }
      };
      emit_items(tl, nrt, nct, 0);
      if (nut > 0) emit_items(tu, nct, nut, 1);
      for (int32_t i = 0; i < h; ++i) gloc[rss[i]] = -1;
    }
    L.item_end = (int64_t)P.items.size();
  }
  if (P.tasks.size() > (size_t)INT32_MAX || P.items.size() > (size_t)INT32_MAX)
    throw std::runtime_error("lu plan: too many cmod tasks");

  // panel steps per level
  P.units.clear();
  for (LuLevelPlan& L : P.levels) {
    L.steps.clear();
    int maxp = 0;
    for (int32_t s : L.all) maxp = std::max(maxp, (P.sn_ncol[s] + P.nb[s] - 1) / P.nb[s]);
    for (int p = 0; p < maxp; ++p) {
      LuStepPlan st;
      st.g_beg = (int64_t)P.units.size();
      for (int32_t s : L.all) {
        const int32_t kb = p * P.nb[s];
        if (kb < P.sn_ncol[s]) P.units.push_back({s, kb, 0, 0});
      }
      st.g_end = st.a_beg = (int64_t)P.units.size();
      for (int32_t s : L.all) {
        const int32_t kb = p * P.nb[s], w = P.sn_ncol[s], h = P.sn_nrow[s];
        if (kb >= w) continue;
        const int32_t ke = std::min(kb + P.nb[s], w);
        const int64_t ncols = (int64_t)kb + (w - ke) + (h - w);
        for (int64_t c0 = 0; c0 < ncols; c0 += LU_COLS) P.units.push_back({s, kb, (int32_t)c0, 0});
        for (int64_t r0 = 0; r0 < h - w; r0 += LU_COLS) P.units.push_back({s, kb, (int32_t)r0, 1});
      }
      st.a_end = st.m_beg = (int64_t)P.units.size();
      for (int32_t s : L.all) {
        const int32_t kb = p * P.nb[s], w = P.sn_ncol[s], h = P.sn_nrow[s];
        if (kb >= w) continue;
        const int32_t ke = std::min(kb + P.nb[s], w), pw = ke - kb;
        const int32_t mt = h - ke, nt = w - ke, mu = w - ke, nu = h - w;
        if (nt > 0)
          for (int32_t rb = 0; rb * TILE < mt; ++rb)
            for (int32_t cb = 0; cb * TILE < nt; ++cb) P.units.push_back({s, kb, rb, cb});
        if (mu > 0 && nu > 0)
          for (int32_t rb = 0; rb * TILE < mu; ++rb)
            for (int32_t cb = 0; cb * TILE < nu; ++cb) P.units.push_back({s, kb, -1 - rb, cb});
        st.gemm_flops += 2.0 * pw * ((double)mt * nt + (double)mu * nu);
      }
      st.m_end = (int64_t)P.units.size();
      L.steps.push_back(st);
    }
  }
  if (P.units.size() > (size_t)INT32_MAX) throw std::runtime_error("lu plan: too many panel units");

  // solve: the scratch rows that update each permuted row (ascending updater)
  P.gat_ptr.assign((size_t)n + 1, 0);
  for (int32_t k = 0; k < ns; ++k)
    for (int64_t q = P.rs_off[k] + P.sn_ncol[k]; q < P.rs_off[k + 1]; ++q) P.gat_ptr[P.rs_idx[q] + 1]++;
  for (int32_t g = 0; g < n; ++g) P.gat_ptr[g + 1] += P.gat_ptr[g];
  P.gat_idx.assign(P.gat_ptr[n], 0);
  {
    std::vector<int64_t> fill(P.gat_ptr.begin(), P.gat_ptr.end() - 1);
    for (int32_t k = 0; k < ns; ++k)
      for (int64_t q = P.rs_off[k] + P.sn_ncol[k]; q < P.rs_off[k + 1]; ++q)
        P.gat_idx[fill[P.rs_idx[q]]++] = P.s_off[k] + (q - P.rs_off[k] - P.sn_ncol[k]);
  }
}

}  // namespace rsl
