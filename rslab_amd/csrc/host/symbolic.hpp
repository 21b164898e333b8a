// Host-side symbolic analysis for the supernodal LDL^T path.
//
// Restates rslab's value-free symbolic pipeline so that the permutation and the
// supernode partition handed to the device numeric factor are the reference's:
//   symmetric pattern   src/sparse/csc.rs:240-289
//   permute_pattern     src/ordering/amd.rs:21-77
//   elimination tree    src/ordering/elimination_tree.rs:28-63
//   postorder           src/ordering/postorder.rs:22-94 (+ biased :98-172)
//   column counts       src/symbolic/column_counts.rs:97-183 (exact counts)
//   supernodes          src/symbolic/supernode.rs:222-722
//   pipeline            src/symbolic/mod.rs:861-1216
//   row structures      src/numeric/multifrontal_ldlt.rs:2121-2152
//   update lists        src/numeric/multifrontal_ldlt.rs:3386-3401
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace rsl {

using idx = int32_t;  // row / column / supernode index
using off = int64_t;  // offset into nonzero or panel storage

// Full (both-triangle) or lower-triangle compressed-column pattern.
struct Pattern {
  idx n = 0;
  std::vector<off> colp;  // n+1
  std::vector<idx> rowi;  // colp[n]
};

enum class Ordering : int { Amd = 0, MetisND = 1, Natural = 2 };
enum class Amalgamation : int { Adjacency = 0, Renumber = 1, Auto = 2 };

struct SymbolicParams {
  int nemin = 16;                 // SolverSettings default (multifrontal_ldlt.rs:600)
  bool relax = true;              // Some(RelaxAmalgamation{256, 64}) by default
  int relax_width = 256;
  int relax_rows = 64;
  Amalgamation amalgamation = Amalgamation::Auto;
  uint64_t nd_seed = 1;           // rslab-metis MetisOptions::default().seed
};

struct Supernode {
  idx first_col = 0;
  idx ncol = 0;
  idx nrow = 0;                   // col_counts[first_col].max(ncol)
  std::vector<idx> children;      // ascending supernode index
};

struct Symbolic {
  idx n = 0;
  std::vector<idx> perm;          // new -> old
  std::vector<idx> perm_inv;      // old -> new
  std::vector<idx> parent;        // etree of the permuted pattern, -1 = root
  std::vector<idx> col_counts;    // column counts of L (incl. diagonal)
  Pattern permuted;               // full pattern of P A P^T
  std::vector<Supernode> snodes;  // postorder (children before parents)
  Ordering method = Ordering::Amd;
  Amalgamation resolved_amalgamation = Amalgamation::Adjacency;
  int64_t factor_nnz = 0;         // sum of column counts
};

// AMD statistics, the counters SuiteSparse AMD reports (amd_2 Info[]).
struct AmdStats {
  int ncmpa = 0;
  int ndense = 0;
  double lnz = 0, ndiv = 0, nms_lu = 0, nms_ldl = 0, dmax = 0;
};

// ---- pattern utilities ----
// Lower CSC (rows sorted, row >= col) -> full symmetric pattern, rows sorted.
Pattern symmetric_pattern(idx n, const off* colp, const idx* rowi);
// P A P^T of a full pattern, perm is new -> old; rows sorted per column.
Pattern permute_pattern(const Pattern& full, const std::vector<idx>& perm);

// ---- orderings (all return new -> old) ----
std::vector<idx> amd_order(const Pattern& full, AmdStats* stats);
std::vector<idx> nd_order(const Pattern& full, uint64_t seed);

// ---- trees ----
std::vector<idx> etree_from_pattern(const Pattern& full);
// Postorder, children by ascending subtree size (stable), roots likewise.
void postorder(const std::vector<idx>& parent, std::vector<idx>& post,
               std::vector<idx>& post_inv);
void biased_postorder(const std::vector<idx>& parent, const std::vector<char>& bias,
                      std::vector<idx>& post, std::vector<idx>& post_inv);
std::vector<idx> column_counts(const Pattern& full, const std::vector<idx>& parent);

// ---- supernodes ----
Amalgamation pick_amalgamation_strategy(const std::vector<idx>& parent);
std::vector<char> predict_merges(const std::vector<idx>& parent,
                                 const std::vector<idx>& cc, int nemin);
std::vector<Supernode> find_supernodes(const std::vector<idx>& parent,
                                       const std::vector<idx>& cc,
                                       const SymbolicParams& p,
                                       Amalgamation strategy);

// Full pipeline: symbolic_prefix_with + symbolic_finish.
// Throws std::runtime_error on invalid input.
Symbolic analyze(idx n, const off* colp, const idx* rowi, Ordering method,
                 const SymbolicParams& p);

// Symbolic of the block-diagonal matrix diag(A_0, A_1, ...) whose blocks were
// analysed one by one: permutations, trees and supernodes side by side, so
// every block keeps exactly its own analysis.
Symbolic compose_block_diagonal(const std::vector<Symbolic>& parts);

// Symbolic from a caller's own analysis: permutation (new -> old) and the
// supernode partition of the permuted columns in postorder with its tree
// (parent a later supernode, or -1). Throws std::runtime_error when the
// permutation or the partition is malformed, or when a supernode's rows
// reach a supernode that is not one of its ancestors.
Symbolic from_partition(idx n, const off* colp, const idx* rowi, const std::vector<idx>& perm,
                        const std::vector<idx>& sn_first, const std::vector<idx>& sn_ncol,
                        const std::vector<idx>& sn_parent);

// Filled row structure per supernode: own columns, then sorted trailing rows.
std::vector<std::vector<idx>> supernode_row_structures(const Symbolic& s);
// update_list[s]: descendants k whose off-diagonal rows hit s's columns,
// ascending k.
std::vector<std::vector<idx>> update_lists(const Symbolic& s,
                                           const std::vector<std::vector<idx>>& rs);

// Validation of a lower CSC (csc.rs:157-238); returns "" when valid.
std::string validate_lower_csc(idx n, const off* colp, const idx* rowi, off nnz);

}  // namespace rsl
