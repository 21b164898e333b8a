// Multilevel nested-dissection ordering (Karypis & Kumar, SISC 20(1) 1998;
// George, SINUM 10(2) 1973) with the exact step sequence of rslab's
// clean-room rslab-metis crate, so the permutation handed to the numeric
// factor is the reference's:
//   driver            crates/rslab-metis/src/node_nd.rs:30-273
//   coarsening        crates/rslab-metis/src/coarsen.rs (SHEM + 2-hop, contraction)
//   initial bisection crates/rslab-metis/src/initial_partition.rs (GGP / BFS)
//   edge FM           crates/rslab-metis/src/fm_refine.rs
//   node separator    crates/rslab-metis/src/separator.rs (Konig cover)
//   node FM, balance  crates/rslab-metis/src/node_refine.rs
//   PRNG              crates/rslab-metis/src/rng.rs (SplitMix64)
// Defaults as MetisOptions::default (crates/rslab-metis/src/lib.rs:140-153):
// seed 1, 7 initial trials, coarsen floor 120, AMD leaves at <= 200 vertices,
// 2-hop threshold 0.85, imbalance 0.20, 10 FM passes.
// Every random draw, stable sort, heap tie-break and work-stack order below is
// part of the result; they follow the reference one for one.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <queue>
#include <stdexcept>
#include <tuple>
#include <vector>

#include "symbolic.hpp"

namespace rsl {

namespace {

constexpr uint8_t PA = 0, PB = 1, PS = 2;

struct Opts {
  uint64_t seed = 1;
  int niparts = 7;
  int coarsen_floor = 120;
  int amd_switch = 200;
  double two_hop = 0.85;
  double imbalance = 0.20;
  int fm_passes = 10;
};

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed + 0x9E3779B97F4A7C15ULL) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  uint64_t range(uint64_t n) {
    const uint64_t zone = UINT64_MAX - (UINT64_MAX % n);
    for (;;) {
      uint64_t r = next();
      if (r < zone) return r % n;
    }
  }
  template <class V>
  void shuffle(V& v) {
    const size_t n = v.size();
    if (n < 2) return;
    for (size_t i = n - 1; i >= 1; --i) {
      size_t j = (size_t)range(i + 1);
      std::swap(v[i], v[j]);
    }
  }
};

struct Graph {
  int32_t n = 0;
  std::vector<int32_t> xadj, adj, vwgt, adjwgt;
  int32_t deg(int32_t v) const { return xadj[v + 1] - xadj[v]; }
};

inline int32_t sat_add(int32_t a, int32_t b) {
  int64_t s = (int64_t)a + b;
  if (s > INT32_MAX) return INT32_MAX;
  if (s < INT32_MIN) return INT32_MIN;
  return (int32_t)s;
}

Graph from_pattern(const Pattern& p) {
  Graph g;
  g.n = p.n;
  g.xadj.push_back(0);
  for (int32_t j = 0; j < p.n; ++j) {
    int32_t last = -1;
    for (off k = p.colp[j]; k < p.colp[j + 1]; ++k) {
      int32_t r = p.rowi[k];
      if (r == j || r == last) continue;
      g.adj.push_back(r);
      last = r;
    }
    g.xadj.push_back((int32_t)g.adj.size());
  }
  g.vwgt.assign(g.n, 1);
  g.adjwgt.assign(g.adj.size(), 1);
  return g;
}

struct Coarse {
  Graph g;
  std::vector<int32_t> cmap;
};

int32_t two_hop_pass(const Graph& f, std::vector<int32_t>& match, std::vector<int32_t>& cmap) {
  for (int32_t v = 0; v < f.n; ++v) {
    if (match[v] != v) continue;
    int32_t partner = -1;
    for (int32_t k = f.xadj[v]; k < f.xadj[v + 1] && partner < 0; ++k) {
      int32_t mid = f.adj[k];
      for (int32_t kk = f.xadj[mid]; kk < f.xadj[mid + 1]; ++kk) {
        int32_t w = f.adj[kk];
        if (w == v) continue;
        if (match[w] == w) {
          partner = w;
          break;
        }
      }
    }
    if (partner >= 0) {
      match[v] = partner;
      match[partner] = v;
    }
  }
  std::fill(cmap.begin(), cmap.end(), -1);
  int32_t cn = 0;
  for (int32_t v = 0; v < f.n; ++v) {
    if (cmap[v] >= 0) continue;
    cmap[v] = cn;
    int32_t u = match[v];
    if (u != v) cmap[u] = cn;
    cn++;
  }
  return cn;
}

Graph contract(const Graph& f, const std::vector<int32_t>& cmap, int32_t cn) {
  Graph c;
  c.n = cn;
  c.vwgt.assign(cn, 0);
  for (int32_t v = 0; v < f.n; ++v) c.vwgt[cmap[v]] = sat_add(c.vwgt[cmap[v]], f.vwgt[v]);
  std::vector<int32_t> head(cn, -1), next(f.n, -1), marker(cn, -1), wto(cn, 0), touched;
  for (int32_t v = 0; v < f.n; ++v) {
    next[v] = head[cmap[v]];
    head[cmap[v]] = v;
  }
  c.xadj.push_back(0);
  for (int32_t ci = 0; ci < cn; ++ci) {
    touched.clear();
    for (int32_t v = head[ci]; v >= 0; v = next[v]) {
      for (int32_t k = f.xadj[v]; k < f.xadj[v + 1]; ++k) {
        int32_t c2 = cmap[f.adj[k]];
        if (c2 == ci) continue;
        if (marker[c2] != ci) {
          marker[c2] = ci;
          wto[c2] = f.adjwgt[k];
          touched.push_back(c2);
        } else {
          wto[c2] = sat_add(wto[c2], f.adjwgt[k]);
        }
      }
    }
    for (int32_t t : touched) {
      c.adj.push_back(t);
      c.adjwgt.push_back(wto[t]);
    }
    c.xadj.push_back((int32_t)c.adj.size());
  }
  return c;
}

Coarse coarsen_level(const Graph& f, SplitMix& rng, double two_hop) {
  const int32_t n = f.n;
  std::vector<int32_t> match(n, -1), cmap(n, -1), order(n);
  for (int32_t i = 0; i < n; ++i) order[i] = i;
  rng.shuffle(order);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return f.deg(a) < f.deg(b); });
  int32_t cn = 0;
  for (int32_t v : order) {
    if (match[v] != -1) continue;
    int32_t best = -1, best_w = -1;
    for (int32_t k = f.xadj[v]; k < f.xadj[v + 1]; ++k) {
      int32_t u = f.adj[k];
      if (match[u] != -1) continue;
      int32_t w = f.adjwgt[k];
      if (w > best_w || (w == best_w && (best < 0 || u < best))) {
        best = u;
        best_w = w;
      }
    }
    if (best >= 0) {
      match[v] = best;
      match[best] = v;
      cmap[v] = cn;
      cmap[best] = cn;
    } else {
      match[v] = v;
      cmap[v] = cn;
    }
    cn++;
  }
  double ratio = (double)cn / (double)std::max(n, 1);
  if (ratio > two_hop && n >= 4) cn = two_hop_pass(f, match, cmap);
  Coarse c;
  c.g = contract(f, cmap, cn);
  c.cmap = std::move(cmap);
  return c;
}

std::vector<Coarse> coarsen(const Graph& g, const Opts& o, SplitMix& rng) {
  std::vector<Coarse> levels;
  int32_t prev = g.n;
  const Graph* cur = &g;
  for (;;) {
    if (cur->n <= o.coarsen_floor) break;
    Coarse lv = coarsen_level(*cur, rng, o.two_hop);
    int32_t nn = lv.g.n;
    if (nn == 0 || (double)nn > 0.95 * (double)prev) {
      if (nn > 0 && nn < prev) levels.push_back(std::move(lv));
      break;
    }
    prev = nn;
    levels.push_back(std::move(lv));
    cur = &levels.back().g;
  }
  return levels;
}

int32_t cut_size(const Graph& g, const std::vector<uint8_t>& lab) {
  int64_t cut = 0;
  for (int32_t v = 0; v < g.n; ++v)
    for (int32_t k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      int32_t u = g.adj[k];
      if (u > v && lab[u] != lab[v]) cut += g.adjwgt[k];
    }
  return (int32_t)std::min<int64_t>(cut, INT32_MAX);
}

int64_t part_weight(const Graph& g, const std::vector<uint8_t>& lab, uint8_t p) {
  int64_t s = 0;
  for (int32_t v = 0; v < g.n; ++v)
    if (lab[v] == p) s += g.vwgt[v];
  return s;
}

std::vector<uint8_t> bisect_ggp(const Graph& g, SplitMix& rng, int64_t target) {
  const int32_t n = g.n;
  std::vector<uint8_t> lab(n, PB);
  if (n == 0) return lab;
  const int32_t seed = (int32_t)rng.range((uint64_t)n);
  lab[seed] = PA;
  int64_t aw = g.vwgt[seed];
  std::vector<int32_t> gain(n, 0), wtot(n, 0), boundary;
  std::vector<char> inb(n, 0);
  for (int32_t v = 0; v < n; ++v)
    for (int32_t k = g.xadj[v]; k < g.xadj[v + 1]; ++k) wtot[v] = sat_add(wtot[v], g.adjwgt[k]);
  auto push = [&](int32_t v) {
    for (int32_t k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      int32_t u = g.adj[k];
      if (lab[u] != PB) continue;
      gain[u] = sat_add(gain[u], g.adjwgt[k]);
      if (!inb[u]) {
        inb[u] = 1;
        boundary.push_back(u);
      }
    }
  };
  push(seed);
  while (aw < target && !boundary.empty()) {
    size_t bi = 0;
    int64_t bg = INT64_MIN;
    for (size_t i = 0; i < boundary.size(); ++i) {
      int32_t v = boundary[i];
      if (lab[v] != PB) continue;
      int64_t gg = 2 * (int64_t)gain[v] - wtot[v];
      if (gg > bg) {
        bg = gg;
        bi = i;
      }
    }
    int32_t v = boundary[bi];
    boundary[bi] = boundary.back();
    boundary.pop_back();
    if (lab[v] != PB) continue;
    lab[v] = PA;
    inb[v] = 0;
    aw += g.vwgt[v];
    push(v);
  }
  return lab;
}

std::vector<uint8_t> bisect_bfs(const Graph& g, SplitMix& rng, int64_t target) {
  const int32_t n = g.n;
  std::vector<uint8_t> lab(n, PB);
  if (n == 0) return lab;
  const int32_t seed = (int32_t)rng.range((uint64_t)n);
  std::vector<int32_t> q;
  size_t qh = 0;
  q.push_back(seed);
  std::vector<char> vis(n, 0);
  vis[seed] = 1;
  int64_t aw = 0;
  while (qh < q.size()) {
    int32_t v = q[qh++];
    if (aw >= target) break;
    lab[v] = PA;
    aw += g.vwgt[v];
    for (int32_t k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      int32_t u = g.adj[k];
      if (!vis[u]) {
        vis[u] = 1;
        q.push_back(u);
      }
    }
  }
  return lab;
}

int32_t refine_bisection(const Graph& g, std::vector<uint8_t>& lab, double imb, int passes) {
  const int32_t n = g.n;
  if (n < 2) return cut_size(g, lab);
  int64_t total = 0;
  for (int32_t w : g.vwgt) total += w;
  const int64_t max_side = (int64_t)std::ceil((1.0 + imb) * (double)total / 2.0);
  int32_t pass_cut = cut_size(g, lab);
  std::vector<int32_t> gain(n), moves;
  std::vector<char> locked(n);
  for (int p = 0; p < passes; ++p) {
    const int32_t before = pass_cut;
    for (int32_t v = 0; v < n; ++v) {
      int32_t ed = 0, id = 0;
      for (int32_t k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
        if (lab[g.adj[k]] == lab[v]) id = sat_add(id, g.adjwgt[k]);
        else ed = sat_add(ed, g.adjwgt[k]);
      }
      gain[v] = ed - id;
    }
    std::fill(locked.begin(), locked.end(), 0);
    // max-heap on (gain, -vertex, stamp): larger gain first, then smaller vertex
    std::priority_queue<std::tuple<int32_t, int32_t, int32_t>> heap;
    for (int32_t v = 0; v < n; ++v) heap.push({gain[v], -v, gain[v]});
    int32_t cur = pass_cut;
    moves.clear();
    int64_t aw = part_weight(g, lab, PA), bw = total - aw;
    bool have_best = std::max(aw, bw) <= max_side;
    size_t best_prefix = 0;
    int32_t best_cut = pass_cut;
    int no_improve = 0;
    while (!heap.empty()) {
      auto [gg, nv, stamp] = heap.top();
      heap.pop();
      (void)gg;
      const int32_t v = -nv;
      if (locked[v] || stamp != gain[v]) continue;
      const uint8_t from = lab[v], to = from == PA ? PB : PA;
      int64_t na = to == PA ? aw + g.vwgt[v] : aw - g.vwgt[v];
      int64_t nb = to == PA ? bw - g.vwgt[v] : bw + g.vwgt[v];
      const int64_t side_max = std::max(na, nb);
      lab[v] = to;
      aw = na;
      bw = nb;
      cur -= gain[v];
      locked[v] = 1;
      moves.push_back(v);
      for (int32_t k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
        int32_t u = g.adj[k];
        if (locked[u]) continue;
        if (lab[u] == from) gain[u] += 2 * g.adjwgt[k];
        else gain[u] -= 2 * g.adjwgt[k];
        heap.push({gain[u], -u, gain[u]});
      }
      if (side_max <= max_side) {
        if (!have_best || cur < best_cut) {
          best_cut = cur;
          best_prefix = moves.size();
          have_best = true;
          no_improve = 0;
        } else {
          no_improve++;
        }
      } else {
        no_improve++;
      }
      if (no_improve >= 50) break;
    }
    if (have_best) {
      for (size_t i = best_prefix; i < moves.size(); ++i) {
        int32_t v = moves[i];
        lab[v] = lab[v] == PA ? PB : PA;
      }
      pass_cut = best_cut;
    } else {
      pass_cut = cur;
    }
    if (pass_cut == before) break;
  }
  return pass_cut;
}

bool augment(int32_t u, const std::vector<std::vector<int32_t>>& adj, std::vector<int32_t>& ma,
             std::vector<int32_t>& mb, std::vector<char>& vis) {
  for (int32_t v : adj[u]) {
    if (vis[v]) continue;
    vis[v] = 1;
    if (mb[v] < 0 || augment(mb[v], adj, ma, mb, vis)) {
      ma[u] = v;
      mb[v] = u;
      return true;
    }
  }
  return false;
}

void construct_separator(const Graph& g, std::vector<uint8_t>& lab) {
  const int32_t n = g.n;
  std::vector<char> ba(n, 0), bb(n, 0);
  for (int32_t v = 0; v < n; ++v) {
    const uint8_t lv = lab[v];
    if (lv != PA && lv != PB) continue;
    for (int32_t k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      const uint8_t lu = lab[g.adj[k]];
      if (lu != lv && (lu == PA || lu == PB)) {
        (lv == PA ? ba : bb)[v] = 1;
        break;
      }
    }
  }
  std::vector<int32_t> al, bl, aix(n, -1), bix(n, -1);
  for (int32_t v = 0; v < n; ++v) {
    if (ba[v]) {
      aix[v] = (int32_t)al.size();
      al.push_back(v);
    }
    if (bb[v]) {
      bix[v] = (int32_t)bl.size();
      bl.push_back(v);
    }
  }
  if (al.empty() || bl.empty()) return;
  const size_t na = al.size(), nb = bl.size();
  std::vector<std::vector<int32_t>> adj(na);
  for (size_t ai = 0; ai < na; ++ai) {
    int32_t v = al[ai];
    for (int32_t k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      int32_t bi = bix[g.adj[k]];
      if (bi >= 0) adj[ai].push_back(bi);
    }
  }
  std::vector<int32_t> ma(na, -1), mb(nb, -1);
  for (size_t ai = 0; ai < na; ++ai) {
    std::vector<char> vis(nb, 0);
    augment((int32_t)ai, adj, ma, mb, vis);
  }
  std::vector<char> za(na, 0), zb(nb, 0);
  std::vector<int32_t> st;
  for (size_t ai = 0; ai < na; ++ai)
    if (ma[ai] < 0) {
      za[ai] = 1;
      st.push_back((int32_t)ai);
    }
  while (!st.empty()) {
    int32_t ai = st.back();
    st.pop_back();
    for (int32_t bi : adj[ai]) {
      if (!zb[bi]) {
        zb[bi] = 1;
        int32_t m = mb[bi];
        if (m >= 0 && !za[m]) {
          za[m] = 1;
          st.push_back(m);
        }
      }
    }
  }
  for (size_t ai = 0; ai < na; ++ai)
    if (!za[ai]) lab[al[ai]] = PS;
  for (size_t bi = 0; bi < nb; ++bi)
    if (zb[bi]) lab[bl[bi]] = PS;
}

// Indexed binary max-heap with the reference's sift rules (ties stay put).
struct MaxHeap {
  std::vector<int32_t> heap, pos;
  std::vector<int64_t> key;
  explicit MaxHeap(int32_t n) : pos(n, -1), key(n, 0) {}
  void clear() {
    for (int32_t v : heap) pos[v] = -1;
    heap.clear();
  }
  bool contains(int32_t v) const { return pos[v] >= 0; }
  void swap_at(size_t i, size_t j) {
    std::swap(heap[i], heap[j]);
    pos[heap[i]] = (int32_t)i;
    pos[heap[j]] = (int32_t)j;
  }
  void up(size_t i) {
    while (i > 0) {
      size_t p = (i - 1) / 2;
      if (key[heap[i]] > key[heap[p]]) {
        swap_at(i, p);
        i = p;
      } else {
        break;
      }
    }
  }
  void down(size_t i) {
    const size_t n = heap.size();
    for (;;) {
      size_t l = 2 * i + 1, r = 2 * i + 2, m = i;
      if (l < n && key[heap[l]] > key[heap[m]]) m = l;
      if (r < n && key[heap[r]] > key[heap[m]]) m = r;
      if (m == i) break;
      swap_at(i, m);
      i = m;
    }
  }
  void insert(int32_t v, int64_t k) {
    key[v] = k;
    pos[v] = (int32_t)heap.size();
    heap.push_back(v);
    up(heap.size() - 1);
  }
  void update(int32_t v, int64_t k) {
    int64_t old = key[v];
    key[v] = k;
    size_t i = (size_t)pos[v];
    if (k > old) up(i);
    else if (k < old) down(i);
  }
  bool pop(int32_t& out) {
    if (heap.empty()) return false;
    out = heap[0];
    pos[out] = -1;
    int32_t last = heap.back();
    heap.pop_back();
    if (!heap.empty()) {
      heap[0] = last;
      pos[last] = 0;
      down(0);
    }
    return true;
  }
};

struct NodeState {
  int64_t pw[3] = {0, 0, 0};
  std::vector<std::array<int64_t, 2>> edeg;
  std::vector<int32_t> bndind, bndptr;
  NodeState(const Graph& g, const std::vector<uint8_t>& lab) : edeg(g.n, {0, 0}), bndptr(g.n, -1) {
    for (int32_t v = 0; v < g.n; ++v) {
      const uint8_t l = lab[v];
      pw[l] += g.vwgt[v];
      if (l == PS) {
        bndptr[v] = (int32_t)bndind.size();
        bndind.push_back(v);
        std::array<int64_t, 2> e{0, 0};
        for (int32_t k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
          int32_t u = g.adj[k];
          if (lab[u] != PS) e[lab[u]] += g.vwgt[u];
        }
        edeg[v] = e;
      }
    }
  }
  void bnd_insert(int32_t v) {
    bndptr[v] = (int32_t)bndind.size();
    bndind.push_back(v);
  }
  void bnd_delete(int32_t v) {
    int32_t p = bndptr[v];
    int32_t last = bndind.back();
    bndind[p] = last;
    bndptr[last] = p;
    bndind.pop_back();
    bndptr[v] = -1;
  }
};

void node_refine_1sided(const Graph& g, std::vector<uint8_t>& lab, double imb, int niter,
                        SplitMix& rng) {
  const int32_t n = g.n;
  NodeState st(g, lab);
  const int64_t total = st.pw[0] + st.pw[1] + st.pw[2];
  const int64_t badmax = (int64_t)(0.5 * (1.0 + imb) * (double)total);
  MaxHeap q(n);
  std::vector<int32_t> swaps, mind, order;
  std::vector<size_t> mptr;
  int to = st.pw[0] < st.pw[1] ? 1 : 0;
  int pass = 0;
  for (;;) {
    if (pass >= 2 * niter) break;
    const int other = to;
    to = 1 - to;
    q.clear();
    int64_t mincutorder = -1;
    const int64_t initcut = st.pw[2];
    int64_t mincut = st.pw[2];
    order = st.bndind;
    rng.shuffle(order);
    for (int32_t v : order) q.insert(v, (int64_t)g.vwgt[v] - st.edeg[v][other]);
    const int64_t nbnd0 = (int64_t)st.bndind.size();
    const int64_t limit = std::min<int64_t>(3 * nbnd0, 300);
    swaps.clear();
    mind.clear();
    mptr.clear();
    mptr.push_back(0);
    int64_t mindiff = std::llabs(st.pw[0] - st.pw[1]);
    int32_t hv;
    while (q.pop(hv)) {
      const int64_t hw = g.vwgt[hv];
      if ((int64_t)mind.size() + g.deg(hv) >= 2 * (int64_t)n - 1) break;
      if (st.pw[to] + hw > badmax) break;
      const int64_t gain = hw - st.edeg[hv][other];
      st.pw[2] -= gain;
      const int64_t newdiff = std::llabs(st.pw[to] + hw - (st.pw[other] - st.edeg[hv][other]));
      const int64_t ns = (int64_t)swaps.size();
      if (st.pw[2] < mincut || (st.pw[2] == mincut && newdiff < mindiff)) {
        mincut = st.pw[2];
        mincutorder = ns;
        mindiff = newdiff;
      } else if ((ns - mincutorder) > 3 * limit ||
                 ((ns - mincutorder) > limit && (double)st.pw[2] > 1.10 * (double)mincut)) {
        st.pw[2] += gain;
        break;
      }
      st.bnd_delete(hv);
      st.pw[to] += hw;
      lab[hv] = (uint8_t)to;
      swaps.push_back(hv);
      for (int32_t k = g.xadj[hv]; k < g.xadj[hv + 1]; ++k) {
        const int32_t u = g.adj[k];
        if (lab[u] == PS) {
          st.edeg[u][to] += hw;
        } else if (lab[u] == (uint8_t)other) {
          st.bnd_insert(u);
          mind.push_back(u);
          lab[u] = PS;
          st.pw[other] -= g.vwgt[u];
          std::array<int64_t, 2> e{0, 0};
          for (int32_t kk = g.xadj[u]; kk < g.xadj[u + 1]; ++kk) {
            const int32_t w = g.adj[kk];
            if (lab[w] != PS) {
              e[lab[w]] += g.vwgt[w];
            } else {
              st.edeg[w][other] -= g.vwgt[u];
              if (q.contains(w)) q.update(w, (int64_t)g.vwgt[w] - st.edeg[w][other]);
            }
          }
          st.edeg[u] = e;
          q.insert(u, (int64_t)g.vwgt[u] - e[other]);
        }
      }
      mptr.push_back(mind.size());
    }
    while ((int64_t)swaps.size() > mincutorder + 1) {
      const int32_t h = swaps.back();
      swaps.pop_back();
      const int64_t hw = g.vwgt[h];
      st.pw[2] += hw;
      st.pw[to] -= hw;
      lab[h] = PS;
      st.bnd_insert(h);
      std::array<int64_t, 2> e{0, 0};
      for (int32_t k = g.xadj[h]; k < g.xadj[h + 1]; ++k) {
        const int32_t u = g.adj[k];
        if (lab[u] == PS) st.edeg[u][to] -= hw;
        else e[lab[u]] += g.vwgt[u];
      }
      st.edeg[h] = e;
      const size_t s = swaps.size();
      for (size_t q2 = mptr[s]; q2 < mptr[s + 1]; ++q2) {
        const int32_t u = mind[q2];
        lab[u] = (uint8_t)other;
        st.pw[other] += g.vwgt[u];
        st.pw[2] -= g.vwgt[u];
        st.bnd_delete(u);
        for (int32_t kk = g.xadj[u]; kk < g.xadj[u + 1]; ++kk) {
          const int32_t w = g.adj[kk];
          if (lab[w] == PS) st.edeg[w][other] += g.vwgt[u];
        }
      }
      mptr.pop_back();
    }
    pass++;
    if (pass % 2 == 0 && (mincutorder == -1 || mincut >= initcut)) break;
  }
}

void node_balance(const Graph& g, std::vector<uint8_t>& lab, double imb, SplitMix& rng) {
  const int32_t n = g.n;
  NodeState st(g, lab);
  const double mult = 0.5 * (1.0 + imb);
  int64_t badmax = (int64_t)(mult * (double)(st.pw[0] + st.pw[1]));
  if (std::max(st.pw[0], st.pw[1]) < badmax) return;
  const int64_t total = st.pw[0] + st.pw[1] + st.pw[2];
  if (std::llabs(st.pw[0] - st.pw[1]) < 3 * total / std::max<int64_t>(n, 1)) return;
  const int to = st.pw[0] < st.pw[1] ? 0 : 1;
  const int other = 1 - to;
  MaxHeap q(n);
  std::vector<char> moved(n, 0);
  std::vector<int32_t> order = st.bndind;
  rng.shuffle(order);
  for (int32_t v : order) q.insert(v, (int64_t)g.vwgt[v] - st.edeg[v][other]);
  int32_t hv;
  while (q.pop(hv)) {
    const int64_t hw = g.vwgt[hv];
    moved[hv] = 1;
    const int64_t gain = hw - st.edeg[hv][other];
    badmax = (int64_t)(mult * (double)(st.pw[0] + st.pw[1]));
    if (st.pw[to] > st.pw[other]) break;
    if (gain < 0 && st.pw[other] < badmax) break;
    if (st.pw[to] + hw > badmax) continue;
    st.pw[2] -= gain;
    st.bnd_delete(hv);
    st.pw[to] += hw;
    lab[hv] = (uint8_t)to;
    for (int32_t k = g.xadj[hv]; k < g.xadj[hv + 1]; ++k) {
      const int32_t u = g.adj[k];
      if (lab[u] == PS) {
        st.edeg[u][to] += hw;
      } else if (lab[u] == (uint8_t)other) {
        st.bnd_insert(u);
        lab[u] = PS;
        st.pw[other] -= g.vwgt[u];
        std::array<int64_t, 2> e{0, 0};
        for (int32_t kk = g.xadj[u]; kk < g.xadj[u + 1]; ++kk) {
          const int32_t w = g.adj[kk];
          if (lab[w] != PS) {
            e[lab[w]] += g.vwgt[w];
          } else {
            st.edeg[w][other] -= g.vwgt[u];
            if (!moved[w] && q.contains(w)) q.update(w, (int64_t)g.vwgt[w] - st.edeg[w][other]);
          }
        }
        st.edeg[u] = e;
        q.insert(u, (int64_t)g.vwgt[u] - e[other]);
      }
    }
  }
}

std::vector<uint8_t> node_bisection(const Graph& g, const Opts& o, SplitMix& rng) {
  std::vector<Coarse> levels = coarsen(g, o, rng);
  const Graph& cg = levels.empty() ? g : levels.back().g;
  int64_t total = 0;
  for (int32_t w : cg.vwgt) total += w;
  const int64_t target = total / 2;
  std::vector<uint8_t> best(cg.n, PA);
  int32_t best_cut = INT32_MAX;
  for (int t = 0; t < o.niparts; ++t) {
    std::vector<uint8_t> lab = (t % 2 == 0) ? bisect_ggp(cg, rng, target) : bisect_bfs(cg, rng, target);
    int32_t cut = refine_bisection(cg, lab, o.imbalance, o.fm_passes);
    if (cut < best_cut) {
      best_cut = cut;
      best = std::move(lab);
    }
  }
  std::vector<uint8_t> lab = std::move(best);
  construct_separator(cg, lab);
  node_refine_1sided(cg, lab, o.imbalance, o.fm_passes, rng);
  for (size_t li = levels.size(); li-- > 0;) {
    const Graph& prev = li == 0 ? g : levels[li - 1].g;
    std::vector<uint8_t> proj(prev.n, PA);
    for (int32_t v = 0; v < prev.n; ++v) proj[v] = lab[levels[li].cmap[v]];
    lab = std::move(proj);
    node_balance(prev, lab, o.imbalance, rng);
    node_refine_1sided(prev, lab, o.imbalance, o.fm_passes, rng);
  }
  return lab;
}

std::vector<int32_t> components(const Graph& g, int32_t& ncc) {
  std::vector<int32_t> cc(g.n, -1), stack;
  ncc = 0;
  for (int32_t s = 0; s < g.n; ++s) {
    if (cc[s] >= 0) continue;
    cc[s] = ncc;
    stack.clear();
    stack.push_back(s);
    while (!stack.empty()) {
      int32_t v = stack.back();
      stack.pop_back();
      for (int32_t k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
        int32_t u = g.adj[k];
        if (cc[u] < 0) {
          cc[u] = ncc;
          stack.push_back(u);
        }
      }
    }
    ncc++;
  }
  return cc;
}

Graph induced(const Graph& g, const std::vector<int32_t>& vmap, const std::vector<int32_t>& loc) {
  Graph s;
  s.n = (int32_t)vmap.size();
  s.xadj.push_back(0);
  for (int32_t v : vmap) {
    s.vwgt.push_back(g.vwgt[v]);
    for (int32_t k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      int32_t lu = loc[g.adj[k]];
      if (lu >= 0) {
        s.adj.push_back(lu);
        s.adjwgt.push_back(g.adjwgt[k]);
      }
    }
    s.xadj.push_back((int32_t)s.adj.size());
  }
  return s;
}

struct Item {
  Graph g;
  std::vector<int32_t> map;
  int64_t off;
};

void amd_leaf(const Graph& g, const std::vector<int32_t>& map, int64_t base,
              std::vector<int32_t>& iperm) {
  if (g.n == 0) return;
  Pattern p;  // graph -> full CSC with the diagonal spliced in, rows sorted
  p.n = g.n;
  p.colp.push_back(0);
  for (int32_t v = 0; v < g.n; ++v) {
    bool diag = false;
    for (int32_t k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      int32_t u = g.adj[k];
      if (!diag && u > v) {
        p.rowi.push_back(v);
        diag = true;
      }
      p.rowi.push_back(u);
    }
    if (!diag) p.rowi.push_back(v);
    p.colp.push_back((off)p.rowi.size());
  }
  std::vector<int32_t> lp = amd_order(p, nullptr);
  for (int32_t pos = 0; pos < g.n; ++pos) iperm[map[lp[pos]]] = (int32_t)(base + pos);
}

}  // namespace

std::vector<idx> nd_order(const Pattern& full, uint64_t seed) {
  Opts o;
  o.seed = seed;
  Graph g = from_pattern(full);
  const int32_t n = g.n;
  std::vector<int32_t> iperm(n, -1);
  SplitMix rng(o.seed);
  std::vector<Item> work;
  {
    int32_t ncc;
    std::vector<int32_t> cc = components(g, ncc);
    int64_t off = 0;
    std::vector<std::vector<int32_t>> maps(ncc);
    for (int32_t v = 0; v < n; ++v) maps[cc[v]].push_back(v);
    std::vector<int32_t> loc(n, -1);
    for (int32_t c = 0; c < ncc; ++c) {
      for (size_t i = 0; i < maps[c].size(); ++i) loc[maps[c][i]] = (int32_t)i;
      Graph s = induced(g, maps[c], loc);
      for (int32_t v : maps[c]) loc[v] = -1;
      const int64_t cnt = (int64_t)maps[c].size();
      if (cnt > 0) work.push_back({std::move(s), maps[c], off});
      off += cnt;
    }
  }
  while (!work.empty()) {
    Item it = std::move(work.back());
    work.pop_back();
    const int32_t m = it.g.n;
    if (m == 0) continue;
    if (m == 1) {
      iperm[it.map[0]] = (int32_t)it.off;
      continue;
    }
    int32_t ncc;
    std::vector<int32_t> cc = components(it.g, ncc);
    if (ncc > 1) {
      int64_t off = it.off;
      std::vector<std::vector<int32_t>> locals(ncc);
      for (int32_t v = 0; v < m; ++v) locals[cc[v]].push_back(v);
      std::vector<int32_t> loc(m, -1);
      for (int32_t c = 0; c < ncc; ++c) {
        for (size_t i = 0; i < locals[c].size(); ++i) loc[locals[c][i]] = (int32_t)i;
        Graph s = induced(it.g, locals[c], loc);
        for (int32_t v : locals[c]) loc[v] = -1;
        std::vector<int32_t> mp(locals[c].size());
        for (size_t i = 0; i < locals[c].size(); ++i) mp[i] = it.map[locals[c][i]];
        const int64_t cnt = (int64_t)locals[c].size();
        if (cnt > 0) work.push_back({std::move(s), std::move(mp), off});
        off += cnt;
      }
      continue;
    }
    if (m <= o.amd_switch) {
      amd_leaf(it.g, it.map, it.off, iperm);
      continue;
    }
    std::vector<uint8_t> lab = node_bisection(it.g, o, rng);
    std::vector<int32_t> av, bv, sv;
    for (int32_t v = 0; v < m; ++v) (lab[v] == PA ? av : lab[v] == PB ? bv : sv).push_back(v);
    const size_t big = std::max(av.size(), bv.size());
    if (av.empty() || bv.empty() || (double)big >= 0.9 * (double)m) {
      amd_leaf(it.g, it.map, it.off, iperm);
      continue;
    }
    const int64_t na = (int64_t)av.size(), nb = (int64_t)bv.size();
    for (size_t i = 0; i < sv.size(); ++i) iperm[it.map[sv[i]]] = (int32_t)(it.off + na + nb + (int64_t)i);
    std::vector<int32_t> loc(m, -1);
    for (size_t i = 0; i < av.size(); ++i) loc[av[i]] = (int32_t)i;
    Graph ga = induced(it.g, av, loc);
    for (int32_t v : av) loc[v] = -1;
    for (size_t i = 0; i < bv.size(); ++i) loc[bv[i]] = (int32_t)i;
    Graph gb = induced(it.g, bv, loc);
    std::vector<int32_t> ma(av.size()), mb(bv.size());
    for (size_t i = 0; i < av.size(); ++i) ma[i] = it.map[av[i]];
    for (size_t i = 0; i < bv.size(); ++i) mb[i] = it.map[bv[i]];
    const int64_t off = it.off;
    work.push_back({std::move(ga), std::move(ma), off});
    work.push_back({std::move(gb), std::move(mb), off + na});
  }
  std::vector<idx> perm(n, -1);
  for (int32_t v = 0; v < n; ++v) {
    if (iperm[v] < 0 || iperm[v] >= n || perm[iperm[v]] >= 0)
      throw std::runtime_error("nested dissection produced an invalid permutation");
    perm[iperm[v]] = v;
  }
  return perm;
}

}  // namespace rsl
