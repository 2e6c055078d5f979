/*
 * amigo_amd — interior-point vector kernels (C ABI).
 *
 * All pointers are HIP device pointers except the `out*` arguments, which
 * are host arrays that receive reduction results (the call synchronizes the
 * stream for those).  `ws` is a device workspace of
 * amigo_ipm_workspace_doubles() doubles.  Returns 0 or -2 (HIP error,
 * message in amigo_ipm_last_error()).
 */
#ifndef AMIGO_IPM_H
#define AMIGO_IPM_H

#ifdef __cplusplus
extern "C" {
#endif

/* ipm::ProblemInfo<T> (include/interior_point_backend.h:55-63). */
typedef struct {
  int n_primal;
  int n_constraints;
  const int* primal_indices;
  const int* constraint_indices;
  const double* lbx;
  const double* ubx;
  const double* lbh;
} amigo_ipm_problem;

/* ipm::State<T> (include/interior_point_backend.h:67-77). */
typedef struct {
  double* xlam;
  double* zl;
  double* zu;
  double* sl;
  double* su;
} amigo_ipm_state;

/* Reduction kinds for amigo_ipm_reduce (a, b, c are kind-specific). */
enum {
  AMIGO_RED_COMPLEMENTARITY = 0, /* ipm::compute_complementarity (backend.h:396): [sum, count, min] */
  AMIGO_RED_KKT_ERROR = 1,       /* ipm::compute_kkt_error (backend.h:470), a=grad: [dual, primal, comp] */
  AMIGO_RED_BARRIER_LOG = 2,     /* ipm::compute_barrier_log_sum (backend.h:494) */
  AMIGO_RED_DPHI_KKT = 3,        /* ipm::compute_barrier_dphi_from_kkt (backend.h:592), a=res b=px c=diag */
  AMIGO_RED_THETA = 4,           /* ipm::compute_constraint_violation_1norm (backend.h:553), a=grad */
  AMIGO_RED_KKT_SQ = 5,          /* ipm::compute_kkt_error_sq (backend.h:530), a=grad */
  AMIGO_RED_MAX_COMP_DEV = 6,    /* ipm::compute_max_comp_deviation (backend.h:417) */
  AMIGO_RED_COMPL_SQ = 7,        /* ipm::compute_complementarity_sq (backend.h:430) */
  AMIGO_RED_DPHI_DIRECT = 8,     /* ipm::compute_barrier_dphi (backend.h:513), a=grad b=px */
  AMIGO_RED_ABS_SUMS = 9         /* optimality scaling sums (optimality_scaling.py:25-60) */
};

const char* amigo_ipm_last_error(void);
int amigo_ipm_workspace_doubles(void);

/* ipm::compute_residual_and_infeasibility (backend.h:227-266) + the L2 norm
 * of InteriorPointOptimizer::compute_residual (optimizer.h:214-225).
 * out3 = [||res||^2, dual_sq, primal_sq]. */
int amigo_ipm_residual(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                       double mu, const double* grad, double* res, double* ws,
                       double* out3, void* stream);

/* ipm::compute_diagonal (backend.h:277-288). */
int amigo_ipm_diagonal(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                       double* diag, void* stream);

/* ipm::compute_bound_dual_step (backend.h:302-321). */
int amigo_ipm_bound_dual_step(const amigo_ipm_problem* P,
                              const amigo_ipm_state* S, double mu,
                              const double* px, double* dzl, double* dzu,
                              void* stream);

/* ipm::compute_max_step (backend.h:328-369) as bound in amigo.cpp:768-775:
 * out4 = [alpha_x, x_index, alpha_z, z_index] (index -1 when 1 is kept). */
int amigo_ipm_max_step(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                       double tau, const double* px, const double* dzl,
                       const double* dzu, double* ws, double* out4,
                       void* stream);

/* ipm::apply_step (backend.h:375-392) into the trial state T. */
int amigo_ipm_apply_step(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                         double ax, double az, const double* dxlam,
                         const double* dzl, const double* dzu, int n,
                         const amigo_ipm_state* T, void* stream);

/* Scalar reductions (AMIGO_RED_*); out4 receives up to 4 values. */
int amigo_ipm_reduce(int kind, const amigo_ipm_problem* P,
                     const amigo_ipm_state* S, double mu, const double* a,
                     const double* b, const double* c, double* ws, double* out4,
                     void* stream);

/* amigo_ipm_reduce without the ws argument and without any allocation: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The partial sums go to a static device array of
 * 4 x 1024 + 16 doubles (32.9 KB) that each GPU holds until the process
 * ends.  Any n_primal + n_constraints up to 2^31 - 1 fits it (the grid is
 * capped at 1024 blocks); larger or negative sizes, an unknown kind or a
 * null pointer return -1 with a message.  All calls on one GPU share its
 * array: calls that overlap on one GPU (other streams, other threads) race
 * and must be serialized by the caller.  Results equal amigo_ipm_reduce's
 * bit for bit. */
int amigo_ipm_reduce_noalloc(int kind, const amigo_ipm_problem* P,
                             const amigo_ipm_state* S, double mu, const double* a,
                             const double* b, const double* c, double* out4,
                             void* stream);

/* ipm::reset_bound_multipliers (backend.h:459-476), in place. */
int amigo_ipm_reset_bound_multipliers(const amigo_ipm_problem* P,
                                      const amigo_ipm_state* S, double mu,
                                      double kappa, void* stream);

/* ipm::compute_affine_start_point (backend.h:480-489): zl_out = max(zl +
 * dzl, beta_min) and zu_out likewise where the bound is finite; the other
 * entries of zl_out / zu_out are left as they are (in place when they are
 * S's duals). */
int amigo_ipm_affine_start_point(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                                 double beta_min, const double* dzl, const double* dzu,
                                 double* zl_out, double* zu_out, void* stream);

/* ipm::compute_dual_residual_vector (backend.h:566-574). */
int amigo_ipm_dual_residual_vector(const amigo_ipm_problem* P,
                                   const amigo_ipm_state* S, const double* grad,
                                   double* out, int n, void* stream);

/* ipm::relax_bounds + project_primals_into_interior + initialize_bound_duals
 * + initialize_slacks (backend.h:79-200; optimizer.h:186-199, 623-634). */
int amigo_ipm_initialize(const double* lbx0, const double* ubx0, int np,
                         const int* pidx, double relax, double cvt,
                         double kappa1, double kappa2, double* lbx, double* ubx,
                         const amigo_ipm_state* S, int do_relax, int do_project,
                         void* stream);

/* Slack flooring of amigo/algorithm/bound_safeguards.py:16-87; out1 =
 * number of adjusted slacks. */
int amigo_ipm_floor_slacks(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                           const double* lbx0, const double* ubx0, double mu,
                           double s_min, double slack_move, double* ws,
                           double* out1, void* stream);

/* CSRMat::mult (include/csr_matrix.h:590-600) on the full pattern. */
int amigo_csr_spmv(int n, const int* rowp, const int* cols, const double* vals,
                   const double* x, double* y, void* stream);
/* amigo_csr_spmv for the rows of the active groups only: row r is computed
 * when grp_active[row_grp[r]] != 0 and otherwise keeps its y (a batch's
 * refinement of the problems still refining; no counterpart in the
 * reference, which solves one problem at a time).  row_grp: n device ints,
 * grp_active: device ints indexed by group; both NULL = every row. */
int amigo_csr_spmv_masked(int n, const int* rowp, const int* cols, const double* vals,
                          const double* x, double* y, const int* row_grp,
                          const int* grp_active, void* stream);

/* InteriorPointOptimizer::apply_hessian_scaling (optimizer.h:573-584). */
int amigo_csr_scale_symmetric(int n, const int* rowp, const int* cols,
                              const double* sv, double* vals, void* stream);
/* amigo_csr_scale_symmetric for the rows of the active groups only (as
 * amigo_csr_spmv_masked; batches, no counterpart in the reference). */
int amigo_csr_scale_symmetric_masked(int n, const int* rowp, const int* cols,
                                     const double* sv, double* vals, const int* row_grp,
                                     const int* grp_active, void* stream);

/* Constraint-row scaling factors of compute_nlp_scaling
 * (optimizer.h:524-540). */
int amigo_ipm_constraint_scaling(int nc, const int* cidx, const int* rowp,
                                 const double* vals, double max_gradient,
                                 double min_value, double* dc, void* stream);

/* v[cidx[j]] *= dc[j] (mode 0), /= dc[j] (mode 1), v[j] *= dc[j] (mode 2):
 * apply_gradient_scaling / unscale_multipliers (optimizer.h:566-598). */
int amigo_ipm_scale_rows(int nc, const int* cidx, const double* dc, double* v,
                         int mode, void* stream);

/* Iterative refinement on the unreduced system
 * (amigo/algorithm/solvers/linear_solver.py:109-264): initial bound-dual
 * steps and the RHS norm (out1); residual blocks, their inf-norm and the
 * condensed correction RHS (out2 = [max|e|, max|(px,dzl,dzu)|]); update. */
int amigo_ipm_ir_init(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                      double mu, const double* grad, const double* rhs_saved,
                      const double* px, double* dzl, double* dzu, double* ws,
                      double* out1, void* stream);
int amigo_ipm_ir_residual(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                          double mu, const double* grad, const double* rhs_saved,
                          const double* Kpx, const double* px, const double* dzl,
                          const double* dzu, double* econd, double* ezl,
                          double* ezu, double* ws, double* out2, void* stream);
int amigo_ipm_ir_update(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                        int n, const double* corr, const double* ezl,
                        const double* ezu, double* px, double* dzl, double* dzu,
                        void* stream);

/* ---------------- batches of independent problems ----------------------
 *
 * K problems laid end to end in one set of vectors (no reference
 * counterpart: the reference solves one problem per InteriorPointOptimizer).
 * Problem k owns primal entries pseg[k]..pseg[k+1], constraint entries
 * cseg[k]..cseg[k+1] and x-length entries xseg[k]..xseg[k+1]; its
 * primal/constraint index lists hold the problem's own variable numbers.
 * The block tables list, for every block of a launch, the problem it
 * serves (blk_*), and per problem its first block and block count
 * (first_*, count_*); count_total[k] must be
 * amigo_ipm_grid_blocks(np_k + nc_k) and count_primal[k]
 * amigo_ipm_grid_blocks(np_k), the grids the problem gets alone, so that
 * each problem's results (reductions included) are bitwise those of the
 * single-problem entry points.  `active` (K ints, NULL = all) selects the
 * problems a call touches; the others' data are left as they are.  Scalars
 * (mu, tau, step lengths) are device arrays of K values.  Reductions write
 * K rows into out_dev (device, K x 4 doubles) and copy them to the host
 * array out* when it is non-NULL.  ws holds 4 x max(nblk_total,
 * nblk_primal) doubles.  All arrays in the struct are device arrays.
 *
 * The launch lists are optional (NULL = launch every block / problem).
 * launch_total lists nlaunch_total block numbers of the total grid (values
 * in 0..nblk_total-1), launch_primal those of the primal grid, and
 * launch_problems nlaunch_problems problem numbers; a call then launches
 * only the listed blocks (for its per-block kernels) and problems (for its
 * per-problem final passes).  They must list every block and every problem
 * whose `active` entry is nonzero, each once; results are bitwise those of
 * the full launch, at a cost proportional to the listed work (the last
 * iterations of a batch, when a few problems of many are still running). */
typedef struct {
  int num_problems;
  const int* pseg;
  const int* cseg;
  const int* xseg;
  int nblk_total;
  const int* blk_total;
  const int* first_total;
  const int* count_total;
  int nblk_primal;
  const int* blk_primal;
  const int* first_primal;
  const int* count_primal;
  const int* active;
  int nlaunch_total;
  const int* launch_total;
  int nlaunch_primal;
  const int* launch_primal;
  int nlaunch_problems;
  const int* launch_problems;
} amigo_ipm_batch;

/* Blocks of the single-problem grid over n entries. */
int amigo_ipm_grid_blocks(int n);

/* Batched counterparts of the entry points above (same formulas and
 * reference functions, per problem). */
int amigo_ipm_batch_residual(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                             const amigo_ipm_state* S, const double* mu, const double* grad,
                             double* res, double* ws, double* out_dev, double* out3,
                             void* stream);
int amigo_ipm_batch_diagonal(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                             const amigo_ipm_state* S, double* diag, void* stream);
int amigo_ipm_batch_bound_dual_step(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                                    const amigo_ipm_state* S, const double* mu,
                                    const double* px, double* dzl, double* dzu, void* stream);
int amigo_ipm_batch_max_step(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                             const amigo_ipm_state* S, const double* tau, const double* px,
                             const double* dzl, const double* dzu, double* ws,
                             double* out_dev, double* out4, void* stream);
int amigo_ipm_batch_apply_step(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                               const amigo_ipm_state* S, const double* ax, const double* az,
                               const double* dxlam, const double* dzl, const double* dzu,
                               const amigo_ipm_state* T, void* stream);
int amigo_ipm_batch_reduce(int kind, const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                           const amigo_ipm_state* S, const double* mu, const double* a,
                           const double* b, const double* c, double* ws, double* out_dev,
                           double* out4, void* stream);
int amigo_ipm_batch_reset_bound_multipliers(const amigo_ipm_batch* B,
                                            const amigo_ipm_problem* P,
                                            const amigo_ipm_state* S, const double* mu,
                                            double kappa, void* stream);
int amigo_ipm_batch_dual_residual_vector(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                                         const amigo_ipm_state* S, const double* grad,
                                         double* out, void* stream);
int amigo_ipm_batch_initialize(const amigo_ipm_batch* B, const double* lbx0,
                               const double* ubx0, const int* pidx, double relax, double cvt,
                               double kappa1, double kappa2, double* lbx, double* ubx,
                               const amigo_ipm_state* S, int do_relax, int do_project,
                               void* stream);
int amigo_ipm_batch_floor_slacks(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                                 const amigo_ipm_state* S, const double* lbx0,
                                 const double* ubx0, const double* mu, const double* s_min,
                                 double slack_move, double* ws, double* out_dev, double* out1,
                                 void* stream);
int amigo_ipm_batch_scale_rows(const amigo_ipm_batch* B, const int* cidx, const double* dc,
                               double* v, int mode, void* stream);
int amigo_ipm_batch_ir_init(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                            const amigo_ipm_state* S, const double* mu, const double* grad,
                            const double* rhs_saved, const double* px, double* dzl, double* dzu,
                            double* ws, double* out_dev, double* out1, void* stream);
int amigo_ipm_batch_ir_residual(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                                const amigo_ipm_state* S, const double* mu, const double* grad,
                                const double* rhs_saved, const double* Kpx, const double* px,
                                const double* dzl, const double* dzu, double* econd,
                                double* ezl, double* ezu, double* ws, double* out_dev,
                                double* out2, void* stream);
int amigo_ipm_batch_ir_update(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                              const amigo_ipm_state* S, const double* corr, const double* ezl,
                              const double* ezu, double* px, double* dzl, double* dzu,
                              void* stream);
/* Largest |x_i| of each of K contiguous segments [seg[k], seg[k+1]) of x
 * (the divergence test, amigo/algorithm/convergence_check.py:82-83, per
 * problem of a batch; NaN propagates as in np.max).  seg: K + 1 device ints;
 * out_dev: K device doubles; out_host (optional): K host doubles, written
 * once the stream has finished. */
int amigo_segment_absmax(const double* x, const int* seg, int K, double* out_dev,
                         double* out_host, void* stream);
/* OptVector::copy (include/interior_point_optimizer.h:21-90) of the active
 * problems' segments: what = 1 x only, 2 zl/zu/sl/su only, 3 all.  A member
 * that is NULL in either state is skipped. */
int amigo_ipm_batch_copy_state(const amigo_ipm_batch* B, const amigo_ipm_state* dst,
                               const amigo_ipm_state* src, int what, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* AMIGO_IPM_H */
