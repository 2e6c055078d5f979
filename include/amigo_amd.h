/*
 * amigo_amd — MI355X-native backend for Amigo's interior-point KKT path.
 *
 * C ABI (plain pointers and sizes; device pointers are HIP device memory,
 * `stream` is a hipStream_t passed as void*, NULL = default stream).
 * Every entry point returns 0 on success or a negative AMIGO_ERR_* code;
 * amigo_last_error() gives a message.  No entry point aborts.
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef AMIGO_AMD_H
#define AMIGO_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  AMIGO_OK = 0,
  AMIGO_ERR_ARG = -1,        /* invalid argument / shape */
  AMIGO_ERR_HIP = -2,        /* HIP runtime error */
  AMIGO_ERR_STATE = -3,      /* call out of order (e.g. solve before factor) */
  AMIGO_ERR_CAPACITY = -4,   /* delayed pivots exceeded the front pool */
  AMIGO_ERR_NOMEM = -5,      /* device allocation failed */
  AMIGO_ERR_NUMERIC = -6,    /* non-finite pivot during factorization */
};

/* Orderings, numbered as include/ordering_utils.h:21 OrderingType. */
enum {
  AMIGO_ORDER_NESTED_DISSECTION = 0,
  AMIGO_ORDER_AMD = 1,
  AMIGO_ORDER_MULTI_COLOR = 2,
  AMIGO_ORDER_NATURAL = 3,
};

/* Solver types, as include/sparse_ldl.h:23 SparseLDL::SolverType. */
enum { AMIGO_SOLVER_LDL = 0, AMIGO_SOLVER_CHOLESKY = 1 };

/* Factor precisions (this build's mixed-precision form). */
enum { AMIGO_PREC_FP64 = 0, AMIGO_PREC_MIXED = 1 };

const char* amigo_last_error(void);

/* ---------------- ordering / symbolic analysis (host) ---------------- */

/* OrderingUtils::copy_for_reorder + OrderingUtils::reorder
 * (include/ordering_utils.h:313-360, :38-56): perm[new] = old,
 * iperm[old] = new.  AMD is bit-identical to BlockAMD::amd
 * (include/block_amd.h:27-452); ND is this build's own recursive bisection. */
int amigo_order(int order, int n, const int* rowp, const int* cols, int* perm,
                int* iperm);

/* BlockAMD::amd on a prepared pattern (diagonal included, rows sorted) with
 * multipliers ordered last (include/block_amd.h:27, :77). */
int amigo_block_amd(int n, const int* rowp, const int* cols, int nmult,
                    const int* mult, int* perm);

/* Etree / postorder / column counts in the permuted order
 * (include/sparse_ldl.h:2375-2701).  perm/iperm may be NULL. */
int amigo_etree_post_colcount(int n, const int* rowp, const int* cols,
                              const int* perm, const int* iperm, int* parent,
                              int* post, int* colcount);

/* ---------------- SparseLDL (symbolic on host, numeric on GPU) ---------- */

typedef struct amigo_ldl amigo_ldl;

/* SparseLDL<T>::SparseLDL(mat, solver_type, ustab, pivot_tol, delay_growth,
 * order) (include/sparse_ldl.h:65-118).  Runs the symbolic analysis on the
 * full symmetric CSR pattern (sorted rows) and uploads the front plan. */
int amigo_ldl_create(int n, const int* rowp, const int* cols, int solver_type,
                     double ustab, double pivot_tol, double delay_growth,
                     int order, amigo_ldl** out);
/* Same, with a group id per variable (e.g. the model node it belongs to;
 * NULL = none).  Used by the nested-dissection ordering only: the quotient
 * graph of the groups is dissected and each group's variables are kept
 * together in index order.  No reference counterpart (the reference's ND is
 * METIS on the variable graph). */
int amigo_ldl_create_grouped(int n, const int* rowp, const int* cols,
                             int solver_type, double ustab, double pivot_tol,
                             double delay_growth, int order,
                             const int* var_group, amigo_ldl** out);
/* Same, with the symmetric permutation given (perm[new] = old) instead of
 * computed: the OrderingUtils::reorder step of SparseLDL's constructor
 * (include/sparse_ldl.h:108-117, ordering_utils.h:38-56) is replaced by the
 * caller's permutation; the rest of the symbolic analysis
 * (sparse_ldl.h:2204-2358) is unchanged.  Used for block-diagonal unions
 * of independent KKT matrices, whose permutation is the concatenation of
 * each block's own ordering, so that every block keeps the fronts it has
 * alone. */
int amigo_ldl_create_perm(int n, const int* rowp, const int* cols, int solver_type,
                          double ustab, double pivot_tol, double delay_growth,
                          const int* perm, amigo_ldl** out);
/* Handle for the block-diagonal union of independent KKT matrices whose
 * own handles (same solver parameters) are given in order: block b's
 * variables follow those of blocks 0..b-1, and (rowp, cols) must be the
 * blocks' patterns placed on the diagonal.  The symbolic plan is the
 * concatenation of the blocks' plans (SparseLDL's constructor,
 * include/sparse_ldl.h:65-118, run once per block), so every block is
 * factored with the fronts it has alone while all blocks share each level
 * kernel launch.  The block handles may be destroyed afterwards. */
int amigo_ldl_create_union(int n, const int* rowp, const int* cols, int num_blocks,
                           const amigo_ldl* const* blocks, amigo_ldl** out);
int amigo_ldl_destroy(amigo_ldl* h);

/* Sizes of the symbolic arrays: num_nodes, total contribution rows, number
 * of elimination-tree levels, and the largest static front. */
int amigo_ldl_sizes(const amigo_ldl* h, int* num_nodes, int* contrib_total,
                    int* num_levels, int* max_front);

/* Copies of the symbolic arrays in the reference's layout
 * (sparse_ldl.h:3252-3268): node_ptr[K+1], node_to_var[n],
 * children_ptr[K+2], children[K], contrib_ptr[K+1],
 * contrib_rows[contrib_total], invperm[n] (identity for NATURAL). */
int amigo_ldl_get_symbolic(const amigo_ldl* h, int* node_ptr, int* node_to_var,
                           int* children_ptr, int* children, int* contrib_ptr,
                           int* contrib_rows, int* invperm);

/* SparseLDL::factor() (include/sparse_ldl.h:136-151): numeric multifrontal
 * LDL^T of the CSR values (device pointer, nnz doubles, same pattern as at
 * create).  precision = AMIGO_PREC_FP64 or AMIGO_PREC_MIXED.  Synchronous:
 * returns after the inertia is known. */
int amigo_ldl_factor(amigo_ldl* h, const double* values_dev, int precision,
                     void* stream);

/* Same as amigo_ldl_factor with the diagonal shift values[diag[i]] +=
 * shift_dev[i] fused into the front assembly (CSRMat::add_diagonal,
 * include/csr_matrix.h:508-528, followed by factor()). shift_dev may be
 * NULL. */
int amigo_ldl_factor_shifted(amigo_ldl* h, const double* values_dev,
                             const double* shift_dev, int precision,
                             void* stream);

/* SparseLDL::get_inertia (include/sparse_ldl.h:198-206).  npos counts zero
 * pivots as positive like the reference; nzero reports them separately. */
int amigo_ldl_get_inertia(const amigo_ldl* h, int* npos, int* nneg,
                          int* nzero);

/* SparseCholesky<T>::SparseCholesky(mat) (include/sparse_cholesky.h:67-136)
 * for a symmetric quasi-definite matrix [A B^T; B -C] whose first
 * sqdef_index variables form the definite block A (CSRMat's sqdef_index,
 * include/csr_matrix.h:689; a value outside [0, n] means the whole matrix
 * is definite, as sparse_cholesky.h:73-76).  Natural order, no pivoting. */
int amigo_chol_create(int n, const int* rowp, const int* cols, int sqdef_index,
                      amigo_ldl** out);
/* SparseCholesky<T>::factor() (include/sparse_cholesky.h:180-306): returns
 * 0 when every pivot has the sign of its block (positive in A, negative in
 * -C), otherwise the 1-based index of the smallest variable whose pivot
 * fails (the reference returns the failing column of its left-looking
 * order; both are positive exactly when the matrix is not quasi-definite
 * in this split), or a negative AMIGO_ERR_* code.  Solve with
 * amigo_ldl_solve (sparse_cholesky.h:311-376). */
int amigo_chol_factor(amigo_ldl* h, const double* values_dev, void* stream);

/* Block-diagonal batches (many independent KKT matrices factored as one):
 * node_seg (host, num_blocks+1 entries, from 0 to num_nodes) gives each
 * block's contiguous node range.  After every factorization the inertia is
 * also counted per block (SparseLDL::get_inertia of each block's own
 * factor, include/sparse_ldl.h:198-206, :1917-1969), and a non-finite pivot
 * is reported in that block's nnan count instead of failing the whole
 * factorization.  get_batch_inertia takes host arrays of num_blocks ints
 * (any may be NULL). */
int amigo_ldl_set_batch(amigo_ldl* h, int num_blocks, const int* node_seg);
int amigo_ldl_get_batch_inertia(const amigo_ldl* h, int* npos, int* nneg, int* nzero,
                                int* nnan);
/* Blocks that the following factorizations and solves touch (host array of
 * num_blocks ints, nonzero = on; NULL = every block).  Each block has its
 * own regions of the front pools, so a block that is left out keeps its
 * last factor: a factorization refactors only the blocks that need it (an
 * inertia-correction retry, the problems still iterating) and a solve runs
 * only theirs.  Blocks' inertia counts are those of their last
 * factorization.  Stream-ordered on `stream`. */
int amigo_ldl_set_batch_mask(amigo_ldl* h, const int* mask, void* stream);

/* Per-node pivot and delayed-pivot counts of the last factorization
 * (MatrixFactor::get_factor, include/sparse_ldl.h:411-445). */
int amigo_ldl_pivot_counts(const amigo_ldl* h, int* num_pivots,
                           int* num_delayed);

/* Symmetric equilibration before each factorization: mode 0 = none (the
 * reference's behaviour, default), 1 = Ruiz scaling (five passes of
 * s_i <- s_i / sqrt(max_j |s_i a_ij s_j|)).  Inertia is unchanged; solves
 * unscale.  No reference counterpart. */
int amigo_ldl_set_scaling(amigo_ldl* h, int mode);

/* HIP-event timing of every factor/solve launch (on = 1; 2 also resets
 * the totals), and the accumulated milliseconds / launch counts per kernel
 * class: [level (assemble + workgroup factor), panel, Schur update,
 * forward solve, backward solve, blocked-front assembly], plus the
 * algorithmic flops of the fronts each factorization path owns.  No
 * reference counterpart. */
int amigo_ldl_set_timing(amigo_ldl* h, int on);
int amigo_ldl_get_timing(const amigo_ldl* h, double* ms6, long long* count6,
                         double* flops_small, double* flops_big);

/* Total delayed pivots and largest front of the last factorization (the
 * reference keeps these per node in MatrixFactor, sparse_ldl.h:477-489). */
int amigo_ldl_get_stats(const amigo_ldl* h, int* num_delayed, int* max_front);

/* SparseLDL::solve(nrhs, x, ldx) (include/sparse_ldl.h:184-190): x is a
 * device column-major n x nrhs block, overwritten in place. */
int amigo_ldl_solve(amigo_ldl* h, int nrhs, double* x_dev, int ldx,
                    void* stream);

/* Flop count of the last factorization per the algorithmic definition
 * sum over fronts sum_k [(m-k-1) + (m-k-1)(m-k)], and bytes streamed by
 * one solve with one right-hand side. */
int amigo_ldl_work(const amigo_ldl* h, double* factor_flops,
                   double* solve_bytes);

/* Replay of HIP graphs (on = 1, the default; 0 = eager launches): once the
 * same value / shift buffers have been factored twice on a stream, the
 * factorization's launch sequence is captured and later factorizations
 * replay it; likewise the solve sequence per right-hand-side buffer.  A
 * replay that finds more delayed pivots than the captured plan holds falls
 * back to eager launches.  No reference counterpart. */
int amigo_ldl_set_graphs(amigo_ldl* h, int on);

/* Predicted device bytes of the numeric phase for this symbolic plan (front
 * pools at delay_growth x the static sizes, index pools, solve and panel
 * workspaces, plan arrays), the counterpart of the memory estimate of
 * SparseLDL's symbolic analysis (include/sparse_ldl.h:3162-3220).  Used to
 * admit solves against the HBM budget. */
int amigo_ldl_memory(const amigo_ldl* h, long long* bytes);

#ifdef __cplusplus
}
#endif

#endif /* AMIGO_AMD_H */
