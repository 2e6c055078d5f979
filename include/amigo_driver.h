/*
 * amigo_amd — native interior-point driver (C ABI).
 *
 * Replaces the reference's Python driver am.Optimizer(...).optimize(options)
 * (amigo/algorithm/ipm_driver.py:41-483 with its mixins: Newton direction
 * and inertia correction, newton_direction.py:17-195 and
 * inertia_correction.py:261-393; filter line search with second-order
 * correction and watchdog, filter_line_search.py:16-438; feasibility
 * restoration, feasibility_restoration.py:16-172; the LOQO / monotone
 * barrier updates, barrier_strategy/heuristic.py and base.py;
 * convergence tests,
 * convergence_check.py:25-138; NLP scaling and least-squares multipliers,
 * iterate_initialization.py:16-69, multiplier_initialization.py:19-93;
 * iterative refinement, solvers/linear_solver.py:109-264).
 *
 * One driver solves K >= 1 independent NLPs laid end to end as one
 * block-diagonal problem (K = 1 is the plain single solve): every device
 * call covers the problems that take part in it, and each problem's
 * decisions are its own, so a problem's iterates are the same whether it is
 * solved alone or beside others.  All device pointers are HIP device memory
 * owned by the caller and must outlive the driver; host pointers are read
 * during amigo_ipm_driver_create only.  Every call returns 0 or a negative
 * code (-1 argument, -2 HIP, -3 state, or an amigo_ldl error), with a
 * message from amigo_ipm_driver_last_error().
 */
#ifndef AMIGO_DRIVER_H
#define AMIGO_DRIVER_H

#include "amigo_amd.h"
#include "amigo_ipm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One element group of the union problem (OptimizationProblem's component
 * groups, include/component_group.h:636-764): E elements (each problem's
 * elements padded to whole blocks of 256) of one element type. */
typedef struct {
  int etype;                /* AMIGO_ELEM_* */
  int E;                    /* elements, a multiple of 256 */
  int ndata;                /* data values per element */
  const int* idx;           /* device, E x inputs (union variable numbers, -1 absent) */
  const double* data;       /* device, E x ndata */
  const int* blk_prob;      /* device, E / 256: problem of each block */
  const int* blk_prob_host; /* host, E / 256 */
  double* alpha_blk;        /* device, E / 256: written by the driver */
  long long goff;           /* offset of the group's gradient buffer slots */
  long long hoff;           /* offset of its Hessian buffer slots */
  int boff;                 /* first Lagrangian partial (block) of the group */
} amigo_nlp_group;

/* The union NLP: variables x (primals and multipliers, n), the KKT pattern
 * and assembly lists, the interior-point index maps and block tables
 * (amigo_ipm_batch), and the union SparseLDL handle (amigo_ldl_create_union
 * + amigo_ldl_set_batch with one block per problem). */
typedef struct {
  int K;
  int n, nnz;
  int n_primal, n_constraints;    /* totals over the problems */
  const int* xseg_host;           /* host, K + 1 */
  const int* pseg_host;           /* host, K + 1 */
  const int* cseg_host;           /* host, K + 1 */
  /* variables and bounds */
  double* x;                      /* device, n: initial point in, iterate out */
  const double* x0;               /* device, n: the initial point (kept) */
  const double* lbx0;             /* device, n_primal: bounds as given */
  const double* ubx0;
  double* lbx;                    /* device, n_primal: relaxed bounds (written) */
  double* ubx;
  const double* lbh0;             /* device, n_constraints: constraint targets as given */
  double* lbh;                    /* device, n_constraints: scaled targets (written) */
  const int* pidx;                /* device, n_primal: problem-local primal rows */
  const int* cidx;                /* device, n_constraints: problem-local constraint rows */
  const long long* pidx_global;   /* device, n_primal: union primal rows */
  const long long* cidx_global;   /* device, n_constraints: union constraint rows */
  const int* cidx_global32;       /* device, n_constraints */
  const int* row_prob;            /* device, n: problem of each row */
  const int* is_mult;             /* device, n: 1 on multiplier rows */
  const int* xseg;                /* device, K + 1 */
  const int* pseg;                /* device, K + 1 */
  const int* cseg;                /* device, K + 1 */
  const long long* xseg64;        /* device, K + 1 */
  const long long* hseg64;        /* device, K + 1: CSR offsets of each problem's rows */
  /* interior-point block tables (amigo_ipm_batch without `active`) */
  int nblk_total;
  const int* blk_total;
  const int* first_total;
  const int* count_total;
  int nblk_primal;
  const int* blk_primal;
  const int* first_primal;
  const int* count_primal;
  /* assembly */
  int ngroups;
  const amigo_nlp_group* groups;  /* host array of ngroups */
  double* gbuf;                   /* device element gradient buffer */
  double* hbuf;                   /* device element Hessian buffer */
  double* lpart;                  /* device Lagrangian partials, one per block */
  const int* g_ptr;               /* device, n + 1 */
  const long long* g_src;
  const int* h_ptr;               /* device, nnz + 1 */
  const long long* h_src;
  const int* lptr;                /* device, K + 1: each problem's partial list */
  const int* lidx;
  /* KKT matrix */
  const int* rowp;                /* device, n + 1 */
  const int* cols;                /* device, nnz */
  double* hess;                   /* device, nnz: assembled values (written) */
  amigo_ldl* ldl;                 /* union handle, batch segments set */
  /* the iterate's bound duals and slacks (device, n_primal each, written)
   * and the constraint scales (device, n_constraints, written) */
  double* zl;
  double* zu;
  double* sl;
  double* su;
  double* dc;
  /* optional (may be NULL): the KKT values and diagonal shift of the first
   * Newton-system factorization of a solve are copied here (nnz, n) */
  double* capture_hess;
  double* capture_shift;
} amigo_nlp_desc;

/* The reference's option table (amigo/algorithm/default_options.py:9-122),
 * the entries the driver reads. */
typedef struct {
  int max_iterations;
  int barrier_heuristic;          /* barrier_strategy == "heuristic" (LOQO); else monotone */
  int progress_based_barrier;
  double convergence_tolerance, dual_inf_tol, constr_viol_tol, compl_inf_tol;
  double diverging_iterates_tol;
  double fraction_to_boundary, initial_barrier_param;
  int max_line_search_iterations;
  double backtracking_factor;
  int equal_primal_dual_step;
  int init_least_squares_multipliers, init_affine_step_multipliers;
  double heuristic_barrier_gamma, heuristic_barrier_r;
  int second_order_correction;
  double filter_gamma_theta, filter_gamma_phi, filter_delta, filter_s_theta, filter_s_phi;
  double filter_eta_phi;
  int filter_max_soc;
  double filter_kappa_soc;
  int filter_restoration_max_iter, filter_reset_trigger, max_filter_resets;
  int watchdog_shortened_iter_trigger, watchdog_trial_iter_max;
  double acceptable_tol;
  int acceptable_iter;
  double acceptable_dual_inf_tol, acceptable_constr_viol_tol, acceptable_compl_inf_tol;
  int adaptive_tau;
  double tau_min;
  double barrier_progress_tol;
  int max_consecutive_rejections;
  double barrier_increase_factor;
  int max_inertia_corrections, inertia_tolerance;
  double mu_linear_decrease_factor, mu_superlinear_decrease_power, barrier_tol_factor;
  double nlp_scaling_max_gradient;
  int record_objective;
  int precision;                  /* AMIGO_PREC_FP64 or AMIGO_PREC_MIXED */
  int profile;                    /* 1: synchronize and time each phase */
} amigo_ipm_options;

/* One record per iteration (amigo/algorithm/iteration_logger.py:15-75). */
typedef struct {
  int iteration, line_iters, x_index, z_index, filter_size;
  double time, residual, barrier_param, alpha_x, alpha_z, inertia_delta;
  double inf_du, inf_pr, compl_inf, nlp_error, objective;
} amigo_ipm_iteration;

/* Termination of convergence_check.py:25-138. */
enum {
  AMIGO_IPM_RUNNING = 0,          /* iteration limit reached */
  AMIGO_IPM_CONVERGED = 1,
  AMIGO_IPM_ACCEPTABLE = 2,
  AMIGO_IPM_DIVERGED = 3,
  AMIGO_IPM_PRECISION_FLOOR = 4
};

typedef struct {
  int status;                     /* AMIGO_IPM_* */
  int num_records;                /* iteration records (iterations + 1) */
  int num_inertia;                /* inertia-correction factorizations with an inertia */
  int num_solver_inertia;         /* every factorization's inertia, initialization included */
  int fp64_fallbacks;             /* mixed precision: refinement that fell back to FP64 */
  int ir_steps;                   /* refinement steps taken */
  int restorations;               /* feasibility restoration phases entered */
  double obj_scale;               /* NLP objective scale */
  double time;                    /* seconds from the start to this problem's end */
} amigo_ipm_result;

typedef struct amigo_ipm_driver amigo_ipm_driver;

/* Called from the solving thread when problem k has finished (its iterate
 * no longer changes). */
typedef void (*amigo_ipm_done_fn)(void* ctx, int k);

const char* amigo_ipm_driver_last_error(void);
int amigo_ipm_driver_create(const amigo_nlp_desc* desc, amigo_ipm_driver** out);
int amigo_ipm_driver_destroy(amigo_ipm_driver* d);
/* optimize(options) of every problem, from x0 (ipm_driver.py:106-483). */
int amigo_ipm_driver_solve(amigo_ipm_driver* d, const amigo_ipm_options* opt,
                           amigo_ipm_done_fn on_done, void* ctx, void* stream);
int amigo_ipm_driver_result(const amigo_ipm_driver* d, int k, amigo_ipm_result* out);
/* Copies up to cap iteration records of problem k; returns the count. */
int amigo_ipm_driver_iterations(const amigo_ipm_driver* d, int k, amigo_ipm_iteration* out,
                                int cap);
/* (n_pos, n_neg) pairs of problem k: which = 0 the inertia-correction
 * factorizations, 1 every factorization.  Returns the count copied. */
int amigo_ipm_driver_inertia(const amigo_ipm_driver* d, int k, int which, int* out, int cap);
/* Factorization and solve calls of the last solve (each covers the problems
 * that take part in it). */
int amigo_ipm_driver_counts(const amigo_ipm_driver* d, long long* num_factor,
                            long long* num_solve);
/* Copies a working vector of the union (n doubles, device) into dst:
 * which = 0 gradient, 1 condensed residual, 2 the diagonal shift of the
 * last factorization, 3 the last Newton direction px. */
int amigo_ipm_driver_copy_vector(const amigo_ipm_driver* d, int which, double* dst, void* stream);
/* Phase times in seconds (with options.profile): [initialize, convergence
 * and barrier, factorization with inertia correction, solve and
 * refinement, line search, restoration, step finish], and host seconds. */
int amigo_ipm_driver_phase_times(const amigo_ipm_driver* d, double* out8);

/* Decision trace of the inertia-correction rules alone (host only; no
 * reference counterpart: a test hook).  outcomes[i] are the results of
 * successive factorizations: 0 inertia correct, 1 singular, 2 too few
 * negative pivots, 3 wrong inertia; -1 starts a new KKT matrix, whose
 * barrier parameter is the next entry of mu.  For every factorization the
 * rules ask for, the perturbation (delta_w, delta_c) it uses goes to trace
 * (2 doubles per entry, at most cap entries); a matrix ends at outcome 0 or
 * when the rules give up, and its further outcomes up to the next -1 are
 * skipped.  Returns the number of entries written. */
int amigo_ipm_regularization_trace(const int* outcomes, int n, const double* mu,
                                   int has_constraints, double* trace, int cap);

#ifdef __cplusplus
}
#endif

#endif /* AMIGO_DRIVER_H */
