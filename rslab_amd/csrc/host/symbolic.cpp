// Symbolic pipeline (see symbolic.hpp for the reference map).
#include "symbolic.hpp"

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <stdexcept>

namespace rsl {

std::string validate_lower_csc(idx n, const off* colp, const idx* rowi, off nnz) {
  // Same checks and order as CscMatrix::validate (csc.rs:157-238).
  if (n < 0) return "negative dimension";
  if (colp[0] != 0) return "col_ptr[0] must be 0";
  if (colp[n] != nnz) return "col_ptr[n] != nnz";
  for (idx j = 0; j < n; ++j)
    if (colp[j + 1] < colp[j]) return "col_ptr not monotonically non-decreasing";
  for (idx j = 0; j < n; ++j) {
    for (off k = colp[j]; k < colp[j + 1]; ++k) {
      if (rowi[k] < 0 || rowi[k] >= n) return "row index out of bounds";
      if (rowi[k] < j) return "row index is upper-triangle";
    }
    for (off k = colp[j] + 1; k < colp[j + 1]; ++k)
      if (rowi[k] <= rowi[k - 1]) return "row indices not sorted";
  }
  return "";
}

Pattern symmetric_pattern(idx n, const off* colp, const idx* rowi) {
  // csc.rs:240-289: count, place (lower entry then transpose), sort columns.
  Pattern p;
  p.n = n;
  std::vector<off> cnt(n, 0);
  for (idx j = 0; j < n; ++j)
    for (off k = colp[j]; k < colp[j + 1]; ++k) {
      idx i = rowi[k];
      cnt[j]++;
      if (i != j) cnt[i]++;
    }
  p.colp.assign(n + 1, 0);
  for (idx j = 0; j < n; ++j) p.colp[j + 1] = p.colp[j] + cnt[j];
  p.rowi.assign(p.colp[n], 0);
  std::vector<off> pos(p.colp.begin(), p.colp.end() - 1);
  for (idx j = 0; j < n; ++j)
    for (off k = colp[j]; k < colp[j + 1]; ++k) {
      idx i = rowi[k];
      p.rowi[pos[j]++] = i;
      if (i != j) p.rowi[pos[i]++] = j;
    }
  for (idx j = 0; j < n; ++j)
    std::sort(p.rowi.begin() + p.colp[j], p.rowi.begin() + p.colp[j + 1]);
  return p;
}

Pattern permute_pattern(const Pattern& a, const std::vector<idx>& perm) {
  // amd.rs:21-77: counting-sort relayout then per-column sort.
  const idx n = a.n;
  std::vector<idx> inv(n);
  for (idx k = 0; k < n; ++k) inv[perm[k]] = k;
  Pattern p;
  p.n = n;
  p.colp.assign(n + 1, 0);
  for (idx j = 0; j < n; ++j) p.colp[inv[j] + 1] = a.colp[j + 1] - a.colp[j];
  for (idx j = 0; j < n; ++j) p.colp[j + 1] += p.colp[j];
  p.rowi.assign(p.colp[n], 0);
  std::vector<off> pos(p.colp.begin(), p.colp.end() - 1);
  for (idx j = 0; j < n; ++j) {
    idx nj = inv[j];
    for (off k = a.colp[j]; k < a.colp[j + 1]; ++k) p.rowi[pos[nj]++] = inv[a.rowi[k]];
  }
  for (idx j = 0; j < n; ++j)
    std::sort(p.rowi.begin() + p.colp[j], p.rowi.begin() + p.colp[j + 1]);
  return p;
}

std::vector<idx> etree_from_pattern(const Pattern& a) {
  // elimination_tree.rs:28-63 (Liu, union-find with path compression).
  const idx n = a.n;
  std::vector<idx> parent(n, -1), anc(n);
  for (idx j = 0; j < n; ++j) {
    anc[j] = j;
    for (off k = a.colp[j]; k < a.colp[j + 1]; ++k) {
      idx i = a.rowi[k];
      if (i >= j) continue;
      idx r = i;
      while (anc[r] != r) r = anc[r];
      idx v = i;
      while (v != r) {
        idx nx = anc[v];
        anc[v] = r;
        v = nx;
      }
      if (r != j) {
        parent[r] = j;
        anc[r] = j;
      }
    }
  }
  return parent;
}

namespace {

struct Tree {
  std::vector<std::vector<idx>> children;  // ascending index
  std::vector<idx> size;                   // subtree sizes
  std::vector<idx> roots;                  // ascending index
};

Tree build_tree(const std::vector<idx>& parent) {
  const idx n = (idx)parent.size();
  Tree t;
  t.children.assign(n, {});
  t.size.assign(n, 1);
  for (idx j = 0; j < n; ++j) {
    if (parent[j] >= 0) t.children[parent[j]].push_back(j);
    else t.roots.push_back(j);
  }
  for (idx j = 0; j < n; ++j)
    if (parent[j] >= 0) t.size[parent[j]] += t.size[j];
  return t;
}

// Ascending subtree size; ties keep ascending index. (Rust's slice
// sort_unstable_by_key is an insertion sort -- hence stable -- on the short
// child lists these trees have.)
void sort_by_size(std::vector<idx>& v, const std::vector<idx>& size) {
  std::stable_sort(v.begin(), v.end(), [&](idx a, idx b) { return size[a] < size[b]; });
}

template <class ChildOrder>
void dfs_postorder(const Tree& t, idx n, ChildOrder order, std::vector<idx>& post,
                   std::vector<idx>& post_inv) {
  post.clear();
  post.reserve(n);
  std::vector<idx> roots = t.roots;
  sort_by_size(roots, t.size);
  struct Frame {
    idx node;
    std::vector<idx> kids;
    size_t ci;
  };
  std::vector<Frame> st;
  for (idx r : roots) {
    st.push_back({r, order(t.children[r]), 0});
    while (!st.empty()) {
      Frame& f = st.back();
      if (f.ci < f.kids.size()) {
        idx c = f.kids[f.ci++];
        st.push_back({c, order(t.children[c]), 0});
      } else {
        post.push_back(f.node);
        st.pop_back();
      }
    }
  }
  post_inv.assign(n, 0);
  for (idx k = 0; k < n; ++k) post_inv[post[k]] = k;
}

}  // namespace

void postorder(const std::vector<idx>& parent, std::vector<idx>& post,
               std::vector<idx>& post_inv) {
  // postorder.rs:22-94
  const idx n = (idx)parent.size();
  Tree t = build_tree(parent);
  dfs_postorder(
      t, n,
      [&](const std::vector<idx>& ch) {
        std::vector<idx> v = ch;
        sort_by_size(v, t.size);
        return v;
      },
      post, post_inv);
}

void biased_postorder(const std::vector<idx>& parent, const std::vector<char>& bias,
                      std::vector<idx>& post, std::vector<idx>& post_inv) {
  // postorder.rs:98-172 + merge_bias_partition (:281-290): unbiased children
  // first, then biased ones, each group by ascending subtree size.
  const idx n = (idx)parent.size();
  Tree t = build_tree(parent);
  dfs_postorder(
      t, n,
      [&](const std::vector<idx>& ch) {
        std::vector<idx> early, late;
        for (idx c : ch) (bias[c] ? late : early).push_back(c);
        sort_by_size(early, t.size);
        sort_by_size(late, t.size);
        early.insert(early.end(), late.begin(), late.end());
        return early;
      },
      post, post_inv);
}

std::vector<idx> column_counts(const Pattern& a, const std::vector<idx>& parent) {
  // Exact column counts of L by the row-subtree / least-common-ancestor
  // method of Gilbert, Ng & Peyton (SIMAX 15(4), 1994), as the reference's
  // column_counts_gnp (column_counts.rs:97-183) computes them.
  const idx n = a.n;
  std::vector<idx> cc(n, 0);
  if (n == 0) return cc;
  // Plain DFS postorder (children ascending) and first descendants.
  std::vector<idx> head(n, -1), next(n, -1), post;
  post.reserve(n);
  for (idx j = n - 1; j >= 0; --j)
    if (parent[j] >= 0) {
      next[j] = head[parent[j]];
      head[parent[j]] = j;
    }
  {
    std::vector<idx> stack;
    for (idx r = 0; r < n; ++r) {
      if (parent[r] >= 0) continue;
      stack.push_back(r);
      while (!stack.empty()) {
        idx v = stack.back();
        idx c = head[v];
        if (c >= 0) {
          head[v] = next[c];
          stack.push_back(c);
        } else {
          post.push_back(v);
          stack.pop_back();
        }
      }
    }
  }
  std::vector<idx> first(n, -1);
  std::vector<int64_t> delta(n, 0);
  for (idx k = 0; k < n; ++k) {
    idx j = post[k];
    delta[j] = (first[j] == -1) ? 1 : 0;  // leaf of the etree
    for (; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
  }
  std::vector<idx> maxfirst(n, -1), prevleaf(n, -1), anc(n);
  for (idx i = 0; i < n; ++i) anc[i] = i;
  for (idx k = 0; k < n; ++k) {
    idx j = post[k];
    if (parent[j] >= 0) delta[parent[j]]--;
    for (off p = a.colp[j]; p < a.colp[j + 1]; ++p) {
      idx i = a.rowi[p];
      if (i <= j || first[j] <= maxfirst[i]) continue;  // j not a leaf of row subtree i
      maxfirst[i] = first[j];
      idx jprev = prevleaf[i];
      prevleaf[i] = j;
      delta[j]++;
      if (jprev != -1) {
        idx q = jprev;
        while (q != anc[q]) q = anc[q];
        for (idx s = jprev, sp; s != q; s = sp) {
          sp = anc[s];
          anc[s] = q;
        }
        delta[q]--;  // q = lca(jprev, j)
      }
    }
    if (parent[j] >= 0) anc[j] = parent[j];
  }
  for (idx k = 0; k < n; ++k) {
    idx j = post[k];
    if (parent[j] >= 0) delta[parent[j]] += delta[j];
  }
  for (idx j = 0; j < n; ++j) cc[j] = (idx)delta[j];
  return cc;
}

Amalgamation pick_amalgamation_strategy(const std::vector<idx>& parent) {
  // supernode.rs:222-247, threshold AUTO_MULTI_CHILD_FRAC_THRESHOLD = 0.05.
  const idx n = (idx)parent.size();
  if (n == 0) return Amalgamation::Adjacency;
  std::vector<idx> cnt(n, 0);
  for (idx p : parent)
    if (p >= 0) cnt[p]++;
  idx leaves = 0, multi = 0;
  for (idx c : cnt) {
    if (c == 0) leaves++;
    if (c >= 2) multi++;
  }
  idx internal = n - leaves;
  if (internal == 0) return Amalgamation::Adjacency;
  double frac = (double)multi / (double)internal;
  return frac < 0.05 ? Amalgamation::Adjacency : Amalgamation::Renumber;
}

namespace {
struct Fundamental {
  std::vector<idx> start, ncol, parent;  // parent = -1 for roots
};

Fundamental find_fundamental(const std::vector<idx>& parent, const std::vector<idx>& cc) {
  // supernode.rs:605-660
  const idx n = (idx)parent.size();
  Fundamental f;
  if (n == 0) return f;
  std::vector<idx> nch(n, 0), id(n, 0);
  for (idx j = 0; j < n; ++j)
    if (parent[j] >= 0) nch[parent[j]]++;
  f.start.push_back(0);
  for (idx j = 1; j < n; ++j) {
    bool same = parent[j - 1] == j && cc[j] + 1 == cc[j - 1] && nch[j] == 1;
    if (same) {
      id[j] = id[j - 1];
    } else {
      id[j] = (idx)f.start.size();
      f.start.push_back(j);
    }
  }
  idx ns = (idx)f.start.size();
  f.ncol.assign(ns, 0);
  f.parent.assign(ns, -1);
  for (idx j = 0; j < n; ++j) f.ncol[id[j]]++;
  for (idx s = 0; s < ns; ++s) {
    idx last = f.start[s] + f.ncol[s] - 1;
    if (parent[last] >= 0) f.parent[s] = id[parent[last]];
  }
  return f;
}
}  // namespace

std::vector<char> predict_merges(const std::vector<idx>& parent, const std::vector<idx>& cc,
                                 int nemin) {
  // supernode.rs:678-722
  const idx n = (idx)parent.size();
  std::vector<char> bias(n, 0);
  if (n == 0) return bias;
  Fundamental f = find_fundamental(parent, cc);
  for (idx s = 0; s < (idx)f.start.size(); ++s) {
    idx p = f.parent[s];
    if (p < 0) continue;
    idx cn = f.ncol[s], pn = f.ncol[p];
    idx sf = f.start[s], pf = f.start[p];
    idx child_last = sf + cn - 1;
    bool trivial = pn == 1 && cc[pf] + 1 == cc[child_last];
    bool size_based = cn < nemin && pn < nemin;
    if (trivial || size_based)
      for (idx c = sf; c < sf + cn; ++c) bias[c] = 1;
  }
  return bias;
}

std::vector<Supernode> find_supernodes(const std::vector<idx>& parent,
                                       const std::vector<idx>& cc, const SymbolicParams& prm,
                                       Amalgamation strategy) {
  // supernode.rs:267-488
  const idx n = (idx)parent.size();
  std::vector<Supernode> out;
  if (n == 0) return out;
  const idx RELAX_MIN_N = 1024, ROOT_CAP_MIN_N = 1024;
  const bool relax = n >= RELAX_MIN_N && prm.relax;
  Fundamental f = find_fundamental(parent, cc);
  const idx ns = (idx)f.start.size();
  std::vector<idx> merged(ns, -1), first = f.start, ncol = f.ncol;
  auto root_of = [&](idx s) {
    while (merged[s] >= 0) s = merged[s];
    return s;
  };
  const bool reverse = relax || strategy == Amalgamation::Renumber;
  const idx root_cap = n >= ROOT_CAP_MIN_N ? std::min<idx>(n / 20, 2048) : INT32_MAX;
  for (idx it = 0; it < ns; ++it) {
    idx s = reverse ? ns - 1 - it : it;
    idx p = f.parent[s];
    if (p < 0) continue;
    if (root_of(s) != s) continue;
    idx rs = s, rp = root_of(p);
    if (rs == rp) continue;
    idx s_first = first[rs], s_ncol = ncol[rs], p_first = first[rp];
    if (s_first + s_ncol != p_first) continue;  // adjacency
    idx cn = ncol[rs], pn = ncol[rp];
    idx child_last = s_first + s_ncol - 1;
    bool trivial = pn == 1 && cc[p_first] + 1 == cc[child_last];
    bool size_based = cn < prm.nemin && pn < prm.nemin;
    bool parent_is_root = f.parent[rp] < 0;
    idx merged_ncol = cn + pn;
    bool cap_exceeded = parent_is_root && merged_ncol > root_cap;
    bool relaxed = false;
    if (relax) {
      idx extra = cc[child_last] - (1 + cc[p_first]);
      if (extra < 0) extra = 0;  // saturating_sub
      relaxed = merged_ncol <= prm.relax_width && extra <= prm.relax_rows;
    }
    if ((trivial || size_based || relaxed) && !cap_exceeded) {
      merged[rs] = rp;
      ncol[rp] = merged_ncol;
      first[rp] = s_first;
    }
  }
  std::vector<idx> newid(ns, 0);
  for (idx s = 0; s < ns; ++s) {
    if (merged[s] >= 0) continue;
    Supernode sn;
    sn.first_col = first[s];
    sn.ncol = ncol[s];
    sn.nrow = std::max(cc[first[s]], ncol[s]);
    newid[s] = (idx)out.size();
    out.push_back(std::move(sn));
  }
  for (idx s = 0; s < ns; ++s) {
    if (merged[s] >= 0) continue;
    idx p = f.parent[s];
    if (p >= 0) {
      idx rp = root_of(p);
      if (rp != s) out[newid[rp]].children.push_back(newid[s]);
    }
  }
  return out;
}

Symbolic analyze(idx n, const off* colp, const idx* rowi, Ordering method,
                 const SymbolicParams& prm) {
  // symbolic_prefix_with (mod.rs:861-1132) + symbolic_finish (:1138-1216).
  Symbolic S;
  S.n = n;
  S.method = method;
  Pattern full = symmetric_pattern(n, colp, rowi);
  std::vector<idx> ord;
  if (method == Ordering::Amd) {
    ord = amd_order(full, nullptr);
  } else if (method == Ordering::MetisND) {
    ord = nd_order(full, prm.nd_seed);
  } else {
    ord.resize(n);
    std::iota(ord.begin(), ord.end(), 0);
  }
  if ((idx)ord.size() != n) throw std::runtime_error("ordering returned wrong length");
  Pattern ap = permute_pattern(full, ord);
  std::vector<idx> aet = etree_from_pattern(ap);
  std::vector<idx> post, post_inv;
  postorder(aet, post, post_inv);
  S.perm.resize(n);
  for (idx k = 0; k < n; ++k) S.perm[k] = ord[post[k]];
  S.perm_inv.resize(n);
  for (idx k = 0; k < n; ++k) S.perm_inv[S.perm[k]] = k;
  S.permuted = permute_pattern(full, S.perm);
  S.parent.assign(n, -1);
  for (idx k = 0; k < n; ++k) {
    idx op = aet[post[k]];
    S.parent[k] = op >= 0 ? post_inv[op] : -1;
  }
  S.col_counts = column_counts(S.permuted, S.parent);
  Amalgamation strat = prm.amalgamation;
  if (strat == Amalgamation::Auto) strat = pick_amalgamation_strategy(S.parent);
  S.resolved_amalgamation = strat;
  if (strat == Amalgamation::Renumber) {
    std::vector<char> bias = predict_merges(S.parent, S.col_counts, prm.nemin);
    bool any = std::any_of(bias.begin(), bias.end(), [](char b) { return b != 0; });
    if (any) {
      std::vector<idx> p2, p2i;
      biased_postorder(S.parent, bias, p2, p2i);
      std::vector<idx> np(n);
      for (idx k = 0; k < n; ++k) np[k] = S.perm[p2[k]];
      S.perm = np;
      for (idx k = 0; k < n; ++k) S.perm_inv[S.perm[k]] = k;
      S.permuted = permute_pattern(full, S.perm);
      S.parent = etree_from_pattern(S.permuted);
      S.col_counts = column_counts(S.permuted, S.parent);
    }
  }
  S.factor_nnz = 0;
  for (idx c : S.col_counts) S.factor_nnz += c;
  S.snodes = find_supernodes(S.parent, S.col_counts, prm, strat);
  return S;
}

Symbolic compose_block_diagonal(const std::vector<Symbolic>& parts) {
  Symbolic S;
  int64_t n = 0;
  for (const Symbolic& q : parts) n += q.n;
  if (n >= INT32_MAX) throw std::runtime_error("batch dimension out of range");
  S.n = (idx)n;
  S.method = parts.empty() ? Ordering::Amd : parts[0].method;
  S.resolved_amalgamation = parts.empty() ? Amalgamation::Adjacency : parts[0].resolved_amalgamation;
  S.perm.reserve(n);
  S.perm_inv.reserve(n);
  S.parent.reserve(n);
  S.col_counts.reserve(n);
  S.permuted.n = S.n;
  S.permuted.colp.assign(1, 0);
  idx base = 0, sbase = 0;
  for (const Symbolic& q : parts) {
    for (idx k = 0; k < q.n; ++k) {
      S.perm.push_back(q.perm[k] + base);
      S.perm_inv.push_back(q.perm_inv[k] + base);
      S.parent.push_back(q.parent[k] >= 0 ? q.parent[k] + base : -1);
      S.col_counts.push_back(q.col_counts[k]);
    }
    const off pb = S.permuted.colp.back();
    for (idx k = 0; k < q.n; ++k) S.permuted.colp.push_back(pb + q.permuted.colp[k + 1]);
    for (idx r : q.permuted.rowi) S.permuted.rowi.push_back(r + base);
    for (const Supernode& sn : q.snodes) {
      Supernode t = sn;
      t.first_col += base;
      for (idx& c : t.children) c += sbase;
      S.snodes.push_back(std::move(t));
    }
    S.factor_nnz += q.factor_nnz;
    base += q.n;
    sbase += (idx)q.snodes.size();
  }
  return S;
}

Symbolic from_partition(idx n, const off* colp, const idx* rowi, const std::vector<idx>& perm,
                        const std::vector<idx>& sn_first, const std::vector<idx>& sn_ncol,
                        const std::vector<idx>& sn_parent) {
  Symbolic S;
  S.n = n;
  S.method = Ordering::Natural;  // the caller's ordering, applied through perm
  if ((idx)perm.size() != n) throw std::runtime_error("perm must hold n entries");
  S.perm = perm;
  S.perm_inv.assign(n, -1);
  for (idx k = 0; k < n; ++k) {
    const idx o = perm[k];
    if (o < 0 || o >= n || S.perm_inv[o] != -1) throw std::runtime_error("perm is not a permutation of 0..n-1");
    S.perm_inv[o] = k;
  }
  const idx ns = (idx)sn_first.size();
  if ((idx)sn_ncol.size() != ns || (idx)sn_parent.size() != ns)
    throw std::runtime_error("supernode arrays must have nsuper entries");
  if (n > 0 && ns == 0) throw std::runtime_error("no supernodes for a nonempty matrix");
  idx next = 0;
  for (idx s = 0; s < ns; ++s) {
    if (sn_first[s] != next || sn_ncol[s] < 1)
      throw std::runtime_error("supernode " + std::to_string(s) + " does not continue the column partition");
    next += sn_ncol[s];
    if (sn_parent[s] != -1 && (sn_parent[s] <= s || sn_parent[s] >= ns))
      throw std::runtime_error("supernode " + std::to_string(s) + ": parent must be a later supernode or -1");
  }
  if (next != n) throw std::runtime_error("the supernode partition does not cover columns 0..n-1");
  Pattern full = symmetric_pattern(n, colp, rowi);
  S.permuted = permute_pattern(full, S.perm);
  S.parent = etree_from_pattern(S.permuted);
  S.col_counts = column_counts(S.permuted, S.parent);
  for (idx c : S.col_counts) S.factor_nnz += c;
  S.snodes.resize(ns);
  for (idx s = 0; s < ns; ++s) {
    Supernode& sn = S.snodes[s];
    sn.first_col = sn_first[s];
    sn.ncol = sn_ncol[s];
    sn.nrow = std::max(S.col_counts[sn_first[s]], sn_ncol[s]);
  }
  for (idx s = 0; s < ns; ++s)
    if (sn_parent[s] >= 0) S.snodes[sn_parent[s]].children.push_back(s);  // ascending s
  // every off-diagonal row of a supernode must lie in one of its ancestors:
  // the left-looking schedule then finishes every updater of a supernode
  // (a descendant) before it, as the reference's recursion does
  std::vector<std::vector<idx>> rs = supernode_row_structures(S);
  std::vector<idx> col2s(n, 0);
  for (idx s = 0; s < ns; ++s)
    for (idx j = 0; j < sn_ncol[s]; ++j) col2s[sn_first[s] + j] = s;
  for (idx s = 0; s < ns; ++s) {
    idx last_t = -1;
    for (size_t q = (size_t)sn_ncol[s]; q < rs[s].size(); ++q) {
      const idx t = col2s[rs[s][q]];
      if (t == last_t) continue;
      last_t = t;
      idx a = sn_parent[s];
      while (a != -1 && a < t) a = sn_parent[a];
      if (a != t)
        throw std::runtime_error("supernode " + std::to_string(s) + " has rows in supernode " + std::to_string(t) +
                                 ", which is not one of its ancestors");
    }
  }
  return S;
}

std::vector<std::vector<idx>> supernode_row_structures(const Symbolic& S) {
  // multifrontal_ldlt.rs:2121-2152
  const idx ns = (idx)S.snodes.size();
  std::vector<std::vector<idx>> rs(ns);
  std::vector<idx> trailing;
  for (idx s = 0; s < ns; ++s) {
    const Supernode& sn = S.snodes[s];
    idx own_last = sn.first_col + sn.ncol;
    trailing.clear();
    for (idx j = sn.first_col; j < own_last; ++j)
      for (off k = S.permuted.colp[j]; k < S.permuted.colp[j + 1]; ++k) {
        idx r = S.permuted.rowi[k];
        if (r >= own_last) trailing.push_back(r);
      }
    for (idx ch : sn.children) {
      idx nck = S.snodes[ch].ncol;
      for (size_t q = nck; q < rs[ch].size(); ++q)
        if (rs[ch][q] >= own_last) trailing.push_back(rs[ch][q]);
    }
    std::sort(trailing.begin(), trailing.end());
    trailing.erase(std::unique(trailing.begin(), trailing.end()), trailing.end());
    std::vector<idx>& r = rs[s];
    r.reserve(sn.ncol + trailing.size());
    for (idx j = sn.first_col; j < own_last; ++j) r.push_back(j);
    r.insert(r.end(), trailing.begin(), trailing.end());
  }
  return rs;
}

std::vector<std::vector<idx>> update_lists(const Symbolic& S,
                                           const std::vector<std::vector<idx>>& rs) {
  // multifrontal_ldlt.rs:3386-3401
  const idx ns = (idx)S.snodes.size();
  std::vector<idx> col2s(S.n, 0);
  for (idx s = 0; s < ns; ++s)
    for (idx c = 0; c < S.snodes[s].ncol; ++c) col2s[S.snodes[s].first_col + c] = s;
  std::vector<std::vector<idx>> ul(ns);
  for (idx k = 0; k < ns; ++k) {
    idx nck = S.snodes[k].ncol, last = -1;
    for (size_t q = nck; q < rs[k].size(); ++q) {
      idx s = col2s[rs[k][q]];
      if (s != last) {
        ul[s].push_back(k);
        last = s;
      }
    }
  }
  return ul;
}

}  // namespace rsl
