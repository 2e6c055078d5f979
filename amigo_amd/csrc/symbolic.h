// Host-side symbolic analysis for the multifrontal LDL^T: ordering,
// elimination tree, column counts, fundamental and relaxed supernodes and
// contribution-row patterns.  The outputs are bit-identical to the
// reference's SparseLDL symbolic phase (include/sparse_ldl.h:2204-2358) for
// the AMD and NATURAL orderings; the GPU numeric phase consumes them.
#pragma once

#include <cstdint>
#include <vector>

namespace amigo_amd {

// Same numbering as the reference's OrderingType (ordering_utils.h:21).
enum Ordering : int {
  ORDER_NESTED_DISSECTION = 0,
  ORDER_AMD = 1,
  ORDER_MULTI_COLOR = 2,
  ORDER_NATURAL = 3,
};

// Approximate minimum degree on the quotient graph with supervariable
// detection.  Input pattern must contain the diagonal and sorted rows; it is
// consumed (overwritten).  `mult` lists variables to be ordered last.
// Follows include/block_amd.h:27-452 step for step (tie-breaking included).
// Returns 0 on success.
int block_amd(int n, int* rowp, int* cols, int nmult, int* mult, int* perm);

// Copy a CSR pattern for reordering: AMD adds the diagonal, ND strips it;
// rows are sorted and made unique (ordering_utils.h:313-360, 812-838).
void copy_for_reorder(int order, int n, const int* rowp, const int* cols,
                      std::vector<int>& rowp_out, std::vector<int>& cols_out);

// Nested dissection by recursive graph bisection (this build's own ND: the
// reference's ND is METIS_NodeND, or the identity without METIS).  With
// `groups` (a group id per variable) the quotient graph of the groups is
// dissected and each group's variables stay together in index order.
void nested_dissection(int n, const int* rowp, const int* cols, int* perm,
                       const int* groups = nullptr);

// perm/iperm as SparseLDL computes them (ordering_utils.h:38-56).
void compute_ordering(int order, int n, const int* rowp, const int* cols,
                      std::vector<int>& perm, std::vector<int>& iperm,
                      const int* groups = nullptr);

// Elimination tree, postorder and column counts in the permuted ordering
// (sparse_ldl.h:2375-2518, 2552-2701).  perm/iperm may be null (identity).
void etree(int n, const int* colp, const int* rows, const int* perm,
           const int* iperm, int* parent);
void postorder(int n, const int* parent, int* post);
void column_counts(int n, const int* colp, const int* rows, const int* post,
                   const int* perm, const int* iperm, const int* parent,
                   int* colcount);

struct SymbolicPlan {
  int n = 0;
  int order = ORDER_NATURAL;
  std::vector<int> perm, iperm;  // empty for NATURAL

  int num_nodes = 0;
  std::vector<int> node_ptr;           // K+1
  std::vector<int> node_to_var;        // n, original numbering
  std::vector<int> node_parent;        // K (-1 for roots)
  std::vector<int> children_ptr;       // K+2, roots listed at the end
  std::vector<int> children;           // K
  std::vector<int> contrib_ptr;        // K+1
  std::vector<int> contrib_rows;       // original numbering
  int max_contrib = 0;

  // Statistics used to size the numeric phase.
  int64_t cholesky_factor_nnz = 0;
  int64_t cholesky_int_nnz = 0;
};

// Full symbolic analysis as in SparseLDL's constructor.  The CSR must hold
// the full symmetric pattern with sorted rows.
// With `user_perm` (perm[new] = old) the ordering step is skipped and that
// permutation is used instead (e.g. the concatenated orderings of the
// independent blocks of a block-diagonal matrix).
void symbolic_analysis(int order, int n, const int* rowp, const int* cols,
                       SymbolicPlan& plan, const int* groups = nullptr,
                       const int* user_perm = nullptr);

// Plan of a block-diagonal matrix from the plans of its blocks (block b's
// variables start at var_offset[b]): nodes, variables, contribution rows and
// permutations are concatenated with offsets and the children lists rebuilt,
// so every block keeps exactly the nodes of its own analysis.  (Analysing
// the whole matrix at once need not: the relaxed amalgamation visits
// children in an order keyed by their slot in the global child array.)
void concat_plans(const std::vector<const SymbolicPlan*>& blocks,
                  const std::vector<int>& var_offset, SymbolicPlan& out);

}  // namespace amigo_amd
