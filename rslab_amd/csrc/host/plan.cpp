// Execution plan: panel layout, update lists, tree levels, cmod tile tasks and
// the A -> panel scatter map. Everything here depends on the pattern only, so
// a plan is reused across numeric refactorizations (LdltSymbolic reuse,
// src/numeric/sparse_solver.rs:457-488).
#include "plan.hpp"

#include <algorithm>
#include <stdexcept>

namespace rsl {

namespace {
constexpr int TILE = 64;
}

void build_plan(Plan& P, const int64_t* colp, const int32_t* rowi, const PlanParams& prm) {
  const Symbolic& S = P.sym;
  P.params = prm;
  P.n = S.n;
  const int32_t ns = (int32_t)S.snodes.size();
  P.ns = ns;
  std::vector<std::vector<idx>> rs = supernode_row_structures(S);
  std::vector<std::vector<idx>> ul = update_lists(S, rs);

  P.sn_first.resize(ns);
  P.sn_ncol.resize(ns);
  P.sn_nrow.resize(ns);
  P.e_off.resize(ns);
  P.nb_eff.resize(ns);
  P.rs_off.resize(ns + 1);
  P.level.assign(ns, 0);
  int32_t acc = 0, maxlev = 0;
  int64_t roff = 0;
  for (int32_t s = 0; s < ns; ++s) {
    const Supernode& sn = S.snodes[s];
    P.sn_first[s] = sn.first_col;
    P.sn_ncol[s] = sn.ncol;
    P.sn_nrow[s] = (int32_t)rs[s].size();
    P.e_off[s] = acc;
    acc += sn.ncol;
    int nb = prm.panel_nb;
    if (sn.ncol >= 512 && nb < 128) nb = 128;
    P.nb_eff[s] = nb;
    P.rs_off[s] = roff;
    roff += (int64_t)rs[s].size();
    int32_t lv = 0;
    for (idx c : sn.children) lv = std::max(lv, P.level[c] + 1);
    P.level[s] = lv;
    maxlev = std::max(maxlev, lv);
  }
  P.rs_off[ns] = roff;
  P.rs_idx.resize(roff);
  for (int32_t s = 0; s < ns; ++s)
    std::copy(rs[s].begin(), rs[s].end(), P.rs_idx.begin() + P.rs_off[s]);
  P.upd_ptr.assign(ns + 1, 0);
  for (int32_t s = 0; s < ns; ++s) P.upd_ptr[s + 1] = P.upd_ptr[s] + (int32_t)ul[s].size();
  P.upd_idx.clear();
  P.upd_idx.reserve(P.upd_ptr[ns]);
  for (int32_t s = 0; s < ns; ++s) P.upd_idx.insert(P.upd_idx.end(), ul[s].begin(), ul[s].end());

  // levels; panels laid out level by level so a level's panels are contiguous
  P.levels.assign(ns ? maxlev + 1 : 0, LevelPlan());
  for (int32_t s = 0; s < ns; ++s) P.levels[P.level[s]].all.push_back(s);
  P.panel_off.assign(ns, 0);
  int64_t poff = 0;
  P.factor_flops = P.geom_flops = 0;
  P.nnzl_panel = 0;
  for (LevelPlan& L : P.levels) {
    for (int32_t s : L.all) {
      const int64_t h = P.sn_nrow[s], w = P.sn_ncol[s];
      P.panel_off[s] = poff;
      poff += h * w;
      if ((double)h * w * w > prm.wide_flops && w > 32) {
        L.wide.push_back(s);
      } else {
        L.small.push_back(s);
        for (int64_t k = 0; k < w; ++k) L.small_flops += (double)(h - k) * (h - k) - (double)(h - w) * (h - w);
      }
      L.tri_entries += 0.5 * (double)w * (w + 1);
      L.off_entries += (double)(h - w) * w;
      L.ncols += w;
      for (int64_t k = 0; k < w; ++k) {
        P.factor_flops += (double)(h - k) * (double)(h - k);
        P.nnzl_panel += h - k;
      }
      P.geom_flops += (double)h * h * w;
    }
  }
  P.panel_total = poff;
  P.g_off.assign(ns, 0);
  P.g_total = 0;
  for (int32_t s = 0; s < ns; ++s) {
    P.g_off[s] = P.g_total;
    P.g_total += (int64_t)(P.sn_nrow[s] - P.sn_ncol[s]) * P.sn_ncol[s];
  }

  // cmod tiles and tasks
  P.items.clear();
  P.tasks.clear();
  P.ptiles.clear();
  P.c_elems_max = 0;
  std::vector<int32_t> gloc(S.n, -1);
  std::vector<std::vector<CTask>> tt;
  struct Run {
    int32_t tile, lo, hi;
  };
  std::vector<Run> cruns, rruns;
  // The cmod work of a level: (target s, its updaters k ascending) groups.
  // Left-looking: a level's targets with all their updaters (pulled before
  // the level's cdiv). Right-looking: every target of the level's updaters
  // with the updaters of that level (pushed after the level's cdiv).
  const int32_t nlev = (int32_t)P.levels.size();
  std::vector<std::vector<std::pair<int32_t, std::vector<int32_t>>>> groups(nlev);
  if (prm.schedule == 0) {
    for (int32_t l = 0; l < nlev; ++l)
      for (int32_t s : P.levels[l].all) {
        if (P.upd_ptr[s] == P.upd_ptr[s + 1]) continue;
        groups[l].push_back({s, std::vector<int32_t>(P.upd_idx.begin() + P.upd_ptr[s],
                                                      P.upd_idx.begin() + P.upd_ptr[s + 1])});
      }
  } else {
    for (int32_t s = 0; s < ns; ++s)
      for (int32_t u = P.upd_ptr[s]; u < P.upd_ptr[s + 1]; ++u) {
        const int32_t k = P.upd_idx[u], l = P.level[k];
        auto& g = groups[l];
        if (g.empty() || g.back().first != s) g.push_back({s, {}});
        g.back().second.push_back(k);  // ascending k within the level
      }
  }
  for (int32_t l = 0; l < nlev; ++l) {
    LevelPlan& L = P.levels[l];
    L.item_beg = (int64_t)P.items.size();
    L.ptile_beg = (int64_t)P.ptiles.size();
    int64_t coff = 0;  // product blocks of the level, one after the other
    for (const auto& grp : groups[l]) {
      const int32_t s = grp.first;
      const int32_t h = P.sn_nrow[s], w = P.sn_ncol[s], first = P.sn_first[s];
      const int32_t* rss = P.rs_idx.data() + P.rs_off[s];
      for (int32_t i = 0; i < h; ++i) gloc[rss[i]] = i;
      const int32_t nrt = (h + TILE - 1) / TILE, nct = (w + TILE - 1) / TILE;
      tt.assign((size_t)nrt * nct, {});
      for (const int32_t k : grp.second) {
        const int32_t nck = P.sn_ncol[k];
        const int32_t* ok = P.rs_idx.data() + P.rs_off[k] + nck;
        const int32_t nok = P.sn_nrow[k] - nck;
        const int32_t p0 = (int32_t)(std::lower_bound(ok, ok + nok, first) - ok);
        const int32_t p1 = (int32_t)(std::lower_bound(ok, ok + nok, first + w) - ok);
        if (p0 == p1) continue;
        // product block of the pair: rows p0..nok, columns p0..p1 (row-major);
        // tiles wholly above its diagonal are never read and not computed
        const int32_t pm = nok - p0, pn = p1 - p0;
        for (int32_t nt = 0; nt * TILE < pn; ++nt)
          for (int32_t mt = nt; mt * TILE < pm; ++mt) P.ptiles.push_back({k, p0, pm, pn, mt, nt, coff});
        const int64_t cbase = coff - (int64_t)p0 * pn - p0;
        coff += (int64_t)pm * pn;
        cruns.clear();
        for (int32_t b = p0; b < p1; ++b) {
          int32_t cb = (ok[b] - first) / TILE;
          if (cruns.empty() || cruns.back().tile != cb) cruns.push_back({cb, b, b + 1});
          else cruns.back().hi = b + 1;
        }
        rruns.clear();
        for (int32_t a = p0; a < nok; ++a) {
          int32_t li = gloc[ok[a]];
          if (li < 0) throw std::runtime_error("plan: updater row outside target structure");
          int32_t rb = li / TILE;
          if (rruns.empty() || rruns.back().tile != rb) rruns.push_back({rb, a, a + 1});
          else rruns.back().hi = a + 1;
        }
        for (const Run& c : cruns)
          for (const Run& r : rruns) {
            if (r.tile < c.tile) continue;
            int32_t a0 = std::max(r.lo, c.lo);
            if (a0 >= r.hi) continue;
            // the kernel maps these ranges into a 64x64 tile: keep them inside it
            if (r.hi - a0 > TILE || c.hi - c.lo > TILE ||
                ok[c.lo] - first < c.tile * TILE || ok[c.hi - 1] - first >= std::min(w, (c.tile + 1) * TILE))
              throw std::runtime_error("plan: cmod task outside its tile");
            CTask tk{k, a0, r.hi, c.lo, c.hi, pn, 0, 0, cbase};
            for (int32_t a = a0; a < r.hi; ++a) tk.rmask |= 1ull << (gloc[ok[a]] - r.tile * TILE);
            for (int32_t b = c.lo; b < c.hi; ++b) tk.cmask |= 1ull << (ok[b] - first - c.tile * TILE);
            tt[(size_t)r.tile * nct + c.tile].push_back(tk);
            double pairs = 0;
            for (int32_t cc = c.lo; cc < c.hi; ++cc) pairs += std::max(0, r.hi - std::max(a0, cc));
            L.cmod_flops += 2.0 * nck * pairs;
          }
      }
      for (int32_t rb = 0; rb < nrt; ++rb)
        for (int32_t cb = 0; cb < nct; ++cb) {
          auto& v = tt[(size_t)rb * nct + cb];
          if (v.empty()) continue;
          CItem it;
          it.s = s;
          it.r0 = rb * TILE;
          it.c0 = cb * TILE;
          it.t_beg = (int32_t)P.tasks.size();
          P.tasks.insert(P.tasks.end(), v.begin(), v.end());
          it.t_end = (int32_t)P.tasks.size();
          P.items.push_back(it);
        }
      for (int32_t i = 0; i < h; ++i) gloc[rss[i]] = -1;
    }
    L.item_end = (int64_t)P.items.size();
    L.ptile_end = (int64_t)P.ptiles.size();
    if (L.ptile_end - L.ptile_beg > INT32_MAX) throw std::runtime_error("plan: too many cmod product tiles");
    L.c_elems = coff;
    P.c_elems_max = std::max(P.c_elems_max, coff);
  }
  if ((int64_t)P.tasks.size() > INT32_MAX) throw std::runtime_error("plan: too many cmod tasks");

  // batched wide path: per level, one step per panel index
  P.witems.clear();
  for (LevelPlan& L : P.levels) {
    L.wsteps.clear();
    int maxp = 0;
    for (int32_t s : L.wide) maxp = std::max(maxp, (P.sn_ncol[s] + P.nb_eff[s] - 1) / P.nb_eff[s]);
    for (int p = 0; p < maxp; ++p) {
      WideStep ws;
      ws.g_beg = (int64_t)P.witems.size();
      for (int32_t s : L.wide) {
        const int32_t kb = p * P.nb_eff[s];
        if (kb < P.sn_ncol[s]) {
          P.witems.push_back({s, kb, 0, 0});
          ws.max_kb = std::max(ws.max_kb, kb);
          const double pw = std::min(P.nb_eff[s], P.sn_ncol[s] - kb);
          ws.getf2_flops += pw * pw * pw / 3.0;
        }
      }
      ws.g_end = ws.r_beg = (int64_t)P.witems.size();
      for (int32_t s : L.wide) {
        const int32_t kb = p * P.nb_eff[s];
        if (kb >= P.sn_ncol[s]) continue;
        const int32_t ke = std::min(kb + P.nb_eff[s], P.sn_ncol[s]);
        const int64_t mt = (int64_t)P.sn_nrow[s] - ke;
        for (int64_t r0 = 0; r0 < mt; r0 += REPLAY_ROWS) P.witems.push_back({s, kb, (int32_t)r0, 0});
        const double pw = ke - kb;
        ws.replay_flops += (double)mt * pw * (pw - 1);
      }
      ws.r_end = ws.t_beg = (int64_t)P.witems.size();
      for (int32_t s : L.wide) {
        const int32_t kb = p * P.nb_eff[s];
        if (kb >= P.sn_ncol[s]) continue;
        const int32_t ke = std::min(kb + P.nb_eff[s], P.sn_ncol[s]);
        const int64_t mt = (int64_t)P.sn_nrow[s] - ke;
        const int32_t cw = P.sn_ncol[s] - ke;
        if (cw <= 0 || mt <= 0) continue;
        const int32_t nrt = (int32_t)((mt + TILE - 1) / TILE), nct = (cw + TILE - 1) / TILE;
        for (int32_t cb = 0; cb < nct; ++cb)
          for (int32_t rb = cb; rb < nrt; ++rb) P.witems.push_back({s, kb, rb, cb});
        const double pw = ke - kb;
        ws.schur_flops += 2.0 * pw * ((double)cw * mt - 0.5 * (double)cw * (cw - 1));
      }
      ws.t_end = (int64_t)P.witems.size();
      L.wsteps.push_back(ws);
    }
  }

  // solve plan: scratch rows, gather lists, per-level work items
  P.u_off.assign(ns, 0);
  P.u_total = 0;
  for (int32_t s = 0; s < ns; ++s) {
    P.u_off[s] = P.u_total;
    P.u_total += P.sn_nrow[s] - P.sn_ncol[s];
  }
  P.gat_ptr.assign((size_t)S.n + 1, 0);
  for (int32_t k = 0; k < ns; ++k)
    for (int64_t q = P.rs_off[k] + P.sn_ncol[k]; q < P.rs_off[k + 1]; ++q) P.gat_ptr[P.rs_idx[q] + 1]++;
  for (int32_t g = 0; g < S.n; ++g) P.gat_ptr[g + 1] += P.gat_ptr[g];
  P.gat_idx.assign(P.gat_ptr[S.n], 0);
  {
    std::vector<int64_t> pos(P.gat_ptr.begin(), P.gat_ptr.end() - 1);
    for (int32_t k = 0; k < ns; ++k) {
      const int64_t base = P.rs_off[k] + P.sn_ncol[k];
      for (int64_t q = base; q < P.rs_off[k + 1]; ++q) P.gat_idx[pos[P.rs_idx[q]]++] = P.u_off[k] + (q - base);
    }
  }
  P.fwd_items.clear();
  P.bwd_items.clear();
  P.gat_items.clear();
  P.sitems.clear();
  P.bitems.clear();
  P.zitems.clear();
  for (LevelPlan& L : P.levels) {
    L.fwd_beg = (int64_t)P.fwd_items.size();
    L.bwd_beg = (int64_t)P.bwd_items.size();
    L.gat_beg = (int64_t)P.gat_items.size();
    L.snarrow.clear();
    L.swide.clear();
    L.ssteps.clear();
    for (int32_t s : L.all) {
      const int32_t off = P.sn_nrow[s] - P.sn_ncol[s];
      for (int32_t r0 = 0; r0 < off; r0 += 64) P.fwd_items.push_back({s, r0});
      if (off > 0)
        for (int32_t p0 = 0; p0 < P.sn_ncol[s]; p0 += 16) P.bwd_items.push_back({s, p0});
      bool any = false;
      for (int32_t c = 0; c < P.sn_ncol[s] && !any; ++c)
        any = P.gat_ptr[P.sn_first[s] + c + 1] > P.gat_ptr[P.sn_first[s] + c];
      if (any)
        for (int32_t r0 = 0; r0 < P.sn_ncol[s]; r0 += 64) P.gat_items.push_back({s, r0});
      if (P.sn_ncol[s] > prm.solve_wide) L.swide.push_back(s);
      else L.snarrow.push_back(s);
    }
    L.fwd_end = (int64_t)P.fwd_items.size();
    L.bwd_end = (int64_t)P.bwd_items.size();
    L.gat_end = (int64_t)P.gat_items.size();
    int32_t maxb = 0;
    for (int32_t s : L.swide) maxb = std::max(maxb, (P.sn_ncol[s] + 63) / 64);
    for (int32_t b = 0; b < maxb; ++b) {
      SolveStep st;
      // forward: row units (s, j0, r0) for the 64-row chunks below block b,
      // then one diagonal unit (s, j0) per supernode, all in one launch
      st.u_beg = (int64_t)P.sitems.size();
      for (int32_t s : L.swide) {
        const int32_t j0 = b * 64, w = P.sn_ncol[s];
        if (j0 < w)
          for (int32_t r0 = j0 + 64; r0 < w; r0 += 64) P.sitems.push_back({s, j0, r0, 0});
      }
      st.u_end = st.d_beg = (int64_t)P.sitems.size();
      for (int32_t s : L.swide)
        if (b * 64 < P.sn_ncol[s]) P.sitems.push_back({s, b * 64, -1, 0});
      st.d_end = (int64_t)P.sitems.size();
      L.ssteps.push_back(st);
    }
    // backward: block b (last first) is solved by every unit of its step; the
    // diagonal unit stores it, unit (s, j0, i0) pushes it into block i0 < j0
    L.bsteps.clear();
    for (int32_t b = maxb - 1; b >= 0; --b) {
      SolveStep st;
      st.d_beg = (int64_t)P.bitems.size();
      for (int32_t s : L.swide)
        if (b * 64 < P.sn_ncol[s]) P.bitems.push_back({s, b * 64, -1, 0});
      st.d_end = st.u_beg = (int64_t)P.bitems.size();
      for (int32_t s : L.swide)
        if (b * 64 < P.sn_ncol[s])
          for (int32_t i0 = 0; i0 < b * 64; i0 += 64) P.bitems.push_back({s, b * 64, i0, 0});
      st.u_end = (int64_t)P.bitems.size();
      L.bsteps.push_back(st);
    }
    L.z_beg = (int64_t)P.zitems.size();
    for (int32_t s : L.swide)
      for (int32_t j0 = 0; j0 < P.sn_ncol[s]; j0 += 64) P.zitems.push_back({s, j0, 0, 0});
    L.z_end = (int64_t)P.zitems.size();
  }

  // dependency-driven solve: every unit waits only on units taken before it
  // (updaters and ancestors sit in earlier / later levels, own blocks in order)
  P.dfwd_items.clear();
  P.dfwd_chain.clear();
  P.dbwd_items.clear();
  // Split level: below it the solve runs level launches (many small
  // supernodes), from it up one dependency-driven launch per sweep. Automatic:
  // the first level with at most 128 supernodes.
  {
    const int32_t nl = (int32_t)P.levels.size();
    int32_t sp = prm.solve_split;
    if (sp < 0) {
      sp = nl;
      for (int32_t l = 0; l < nl; ++l)
        if ((int64_t)P.levels[l].all.size() <= 128) {
          sp = l;
          break;
        }
    }
    P.solve_split = std::min(sp, nl);
  }
  // Within a level the units that start chains come first (forward: own
  // blocks; backward: the off units the own blocks wait for), so workgroups
  // waiting on a chain do not hold the slots its links need.
  P.fu_init.assign(ns, 0);
  P.dep_tri = P.dep_off = 0;
  for (size_t l = 0; l < P.levels.size(); ++l) {
    const LevelPlan& L = P.levels[l];
    if ((int32_t)l < P.solve_split) {
      for (int32_t s : L.all) P.fu_init[s] = (P.sn_nrow[s] - P.sn_ncol[s] + 63) / 64;
      continue;
    }
    for (int32_t s : L.all) {
      const int32_t w = P.sn_ncol[s], off = P.sn_nrow[s] - w;
      for (int32_t c = 0; c < (w + 63) / 64; ++c) P.dfwd_items.push_back({s, c});
      // a chain's helpers follow it: only the newest chain can wait on a
      // helper not yet started, and every unit before it finishes without it
      P.dfwd_chain.push_back({s, -1});
      for (int32_t c = 1; c < (w + 63) / 64; ++c) P.dfwd_chain.push_back({s, c});
      P.dep_tri += 0.5 * (double)w * (w + 1);
      P.dep_off += (double)off * w;
    }
    for (int32_t s : L.all) {
      const int32_t w = P.sn_ncol[s], nblk = (w + 63) / 64, nuch = (P.sn_nrow[s] - w + 63) / 64;
      for (int32_t c = nblk; c < nblk + nuch; ++c) {
        P.dfwd_items.push_back({s, c});
        P.dfwd_chain.push_back({s, c});
      }
    }
  }
  for (size_t l = P.levels.size(); l-- > (size_t)P.solve_split;) {
    // off units (s, -1 - (q + ng r)) for 16-column chunks q of row chunk r,
    // then own blocks, last first
    for (int32_t s : P.levels[l].all)
      if (P.sn_nrow[s] > P.sn_ncol[s]) {
        const int32_t w = P.sn_ncol[s], ng = (w + 15) / 16, nch = bwd_row_chunks(w, P.sn_nrow[s] - w);
        for (int32_t r = 0; r < nch; ++r)
          for (int32_t q = 0; q < ng; ++q) P.dbwd_items.push_back({s, -1 - (q + ng * r)});
      }
    for (int32_t s : P.levels[l].all)
      for (int32_t c = (P.sn_ncol[s] + 63) / 64; c-- > 0;) P.dbwd_items.push_back({s, c});
  }
  {
    std::vector<int32_t> owner(S.n, 0);
    for (int32_t s = 0; s < ns; ++s)
      for (int32_t c = 0; c < P.sn_ncol[s]; ++c) owner[P.sn_first[s] + c] = s;
    P.anc_ptr.assign((size_t)ns + 1, 0);
    P.anc_idx.clear();
    for (int32_t s = 0; s < ns; ++s) {
      // off-diagonal rows ascend and every owner holds a contiguous column
      // range, so equal owners are adjacent
      int32_t last = -1;
      for (int64_t q = P.rs_off[s] + P.sn_ncol[s]; q < P.rs_off[s + 1]; ++q) {
        const int32_t a = owner[P.rs_idx[q]];
        if (a != last) P.anc_idx.push_back(a);
        last = a;
      }
      P.anc_ptr[s + 1] = (int32_t)P.anc_idx.size();
    }
  }
  // The device units wait only on units taken before them; a plan that broke
  // this would make a unit wait forever, so it is checked here. Forward: a
  // supernode's own blocks come after every off-diagonal chunk of each of its
  // updaters. Backward: after the first own block (the last unit) of every
  // ancestor owning its off-diagonal rows. Updaters below the split have no
  // units and their chunk counters start full. Within a supernode the order
  // is fixed by construction.
  {
    std::vector<int64_t> fu_last(ns, -1), f_first(ns, -1), b_last(ns, -1), b_first(ns, -1);
    for (int64_t i = 0; i < (int64_t)P.dfwd_items.size(); ++i) {
      const Int2 it = P.dfwd_items[i];
      if (f_first[it.x] < 0) f_first[it.x] = i;
      fu_last[it.x] = i;
    }
    // the chain list holds the same units per supernode in the same places
    // (a chain and its helpers where the own blocks were)
    if (P.dfwd_chain.size() != P.dfwd_items.size())
      throw std::runtime_error("plan: chain and block solve lists differ in length");
    for (size_t i = 0; i < P.dfwd_items.size(); ++i)
      if (P.dfwd_chain[i].x != P.dfwd_items[i].x)
        throw std::runtime_error("plan: chain and block solve lists order supernodes differently");
    for (int64_t i = 0; i < (int64_t)P.dbwd_items.size(); ++i) {
      const Int2 it = P.dbwd_items[i];
      if (b_first[it.x] < 0) b_first[it.x] = i;
      b_last[it.x] = i;
    }
    for (int32_t s = 0; s < ns; ++s) {
      if (f_first[s] < 0) continue;  // below the split: no units
      for (int32_t q = P.upd_ptr[s]; q < P.upd_ptr[s + 1]; ++q) {
        const int32_t k = P.upd_idx[q];
        if (fu_last[k] < 0 ? P.fu_init[k] != (P.sn_nrow[k] - P.sn_ncol[k] + 63) / 64 : fu_last[k] >= f_first[s])
          throw std::runtime_error("plan: solve unit ordered before a unit it waits on");
      }
      for (int32_t q = P.anc_ptr[s]; q < P.anc_ptr[s + 1]; ++q)
        if (b_last[P.anc_idx[q]] < 0 || b_last[P.anc_idx[q]] >= b_first[s])
          throw std::runtime_error("plan: solve unit ordered before a unit it waits on");
    }
  }

  // A scatter map: user lower entry (i >= j) -> slot of P^T A P in its panel
  P.nnz = colp[S.n];
  P.amap.resize(P.nnz);
  P.a_row.resize(P.nnz);
  P.a_col.resize(P.nnz);
  std::vector<int32_t> col2s(S.n, 0);
  for (int32_t s = 0; s < ns; ++s)
    for (int32_t c = 0; c < P.sn_ncol[s]; ++c) col2s[P.sn_first[s] + c] = s;
  for (int32_t j = 0; j < S.n; ++j)
    for (int64_t q = colp[j]; q < colp[j + 1]; ++q) {
      const int32_t i = rowi[q];
      const int32_t gi = S.perm_inv[i], gj = S.perm_inv[j];
      const int32_t r = std::max(gi, gj), c = std::min(gi, gj);
      const int32_t s = col2s[c];
      const int32_t* rss = P.rs_idx.data() + P.rs_off[s];
      const int32_t h = P.sn_nrow[s];
      const int32_t li = (int32_t)(std::lower_bound(rss, rss + h, r) - rss);
      if (li >= h || rss[li] != r) throw std::runtime_error("plan: entry outside panel structure");
      P.amap[q] = P.panel_off[s] + li + (int64_t)(c - P.sn_first[s]) * h;
      P.a_row[q] = i;
      P.a_col[q] = j;
    }
}

}  // namespace rsl
