/*
 * amigo_amd — KKT gradient / Hessian assembly of component groups (C ABI).
 *
 * Replaces OptimizationProblem::gradient / hessian / lagrangian
 * (include/optimization_problem.h:718-818) for the element types below;
 * device pointers throughout, results stream-ordered.
 */
#ifndef AMIGO_ASSEMBLY_H
#define AMIGO_ASSEMBLY_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  AMIGO_ELEM_LINK = 0,        /* (x, lambda): lambda * x            */
  AMIGO_ELEM_DI_INTERVAL = 1, /* double-integrator collocation interval */
  AMIGO_ELEM_ELLIPTIC2D = 2,  /* 5-point semilinear elliptic control node */
  AMIGO_ELEM_ELLIPTIC3D = 3,  /* 7-point semilinear elliptic control node */
  AMIGO_ELEM_CARTPOLE = 4     /* cart-pole collocation interval */
};

/* First input of a padding element (a batch pads each problem's elements
 * to whole blocks of 256). */
#define AMIGO_ELEM_PAD (-2)

const char* amigo_assembly_last_error(void);

/* Number of inputs per element of a type (-1 if unknown). */
int amigo_elem_num_inputs(int type);

/* Element evaluator of a generated element type: the arguments of
 * amigo_elements_eval_masked without the type. */
typedef int (*amigo_elem_eval_fn)(int E, const int* idx, const double* data, int ndata,
                                  const double* x, double alpha, const double* alpha_blk,
                                  const int* blk_grp, const int* grp_active, double* gbuf,
                                  double* hbuf, double* lpart, void* stream);

/* First type number of generated element types; AMIGO_ELEM_MAX_GENERATED
 * of them can be registered. */
#define AMIGO_ELEM_GENERATED 100
#define AMIGO_ELEM_MAX_GENERATED 256

/* Register (or replace) the evaluator of generated element type `type`
 * (AMIGO_ELEM_GENERATED <= type < AMIGO_ELEM_GENERATED +
 * AMIGO_ELEM_MAX_GENERATED) with nc inputs per element: every evaluation
 * entry point below then dispatches that type to fn.  The kernels come from
 * component expressions (amigo_amd.codegen), the counterpart of the
 * reference's generated component groups (amigo/component.py:835-1167,
 * amigo/model.py:1409-1622). */
int amigo_elem_register(int type, int nc, amigo_elem_eval_fn fn);

/* Per-element gradient (E x NC) and Hessian (E x NC x NC, row-major) of
 * alpha * f_e + lambda^T c_e at x; idx is E x NC (-1 = absent input), data
 * is E x ndata.  gbuf / hbuf / lpart may be NULL; lpart receives one partial
 * Lagrangian sum per block of 256 elements. */
int amigo_elements_eval(int type, int E, const int* idx, const double* data,
                        int ndata, const double* x, double alpha, double* gbuf,
                        double* hbuf, double* lpart, void* stream);

/* Same, for a batch of problems laid end to end with each problem's
 * elements padded to whole blocks of 256: block b scales the objective by
 * alpha_blk[b] (device, one value per block; NULL = alpha for all), so each
 * problem keeps its own objective scale (the obj_scale of
 * InteriorPointOptimizer::compute_nlp_scaling, optimizer.h:502-563). */
int amigo_elements_eval_scaled(int type, int E, const int* idx, const double* data,
                               int ndata, const double* x, double alpha,
                               const double* alpha_blk, double* gbuf, double* hbuf,
                               double* lpart, void* stream);

/* out[i] (+)= sum over p in [ptr[i], ptr[i+1]) of buf[src[p]], summed in
 * list order (deterministic). */
int amigo_gather_sum(int n, const int* ptr, const long long* src,
                     const double* buf, double* out, int accumulate,
                     void* stream);

/* amigo_elements_eval_scaled for the blocks of the active groups only:
 * block b (256 elements) is evaluated when grp_active[blk_grp[b]] != 0 and
 * otherwise writes nothing (a batch's problems still iterating; no
 * counterpart in the reference, which evaluates one problem at a time).
 * blk_grp: one device int per block; grp_active: device ints by group;
 * both NULL = every block. */
int amigo_elements_eval_masked(int type, int E, const int* idx, const double* data,
                               int ndata, const double* x, double alpha,
                               const double* alpha_blk, const int* blk_grp,
                               const int* grp_active, double* gbuf, double* hbuf,
                               double* lpart, void* stream);

/* amigo_gather_sum for the outputs of the active groups only: the outputs
 * are grouped by seg (nseg + 1 increasing device offsets, seg[0] = 0,
 * seg[nseg] = n); outputs of groups with seg_active[g] == 0 keep their
 * value (batches, as above).  seg = seg_active = NULL: every output. */
int amigo_gather_sum_masked(int n, const int* ptr, const long long* src,
                            const double* buf, double* out, int accumulate,
                            const long long* seg, int nseg, const int* seg_active,
                            void* stream);

/* amigo_gather_sum_masked for a few groups of many: the outputs of the
 * nlist groups list[0..nlist) only (device ints, each a group number g
 * with outputs seg[g] .. seg[g+1]; max_len >= the longest of them); the
 * other outputs keep their value and the launch's size follows the listed
 * outputs, not n (no counterpart in the reference, which sums one
 * problem's outputs at a time). */
int amigo_gather_sum_listed(const int* ptr, const long long* src, const double* buf,
                            double* out, int accumulate, const long long* seg,
                            const int* list, int nlist, long long max_len, void* stream);

/* out_dev[0] = sum of part[0..nb) in order; copied to out_host if given. */
int amigo_sum_partials(const double* part, int nb, double* out_dev,
                       double* out_host, void* stream);

/* Per-problem Lagrangians of a batch (OptimizationProblem::lagrangian,
 * optimization_problem.h:718-733, per problem): out_dev[k] = sum of
 * part[lidx[p]] for p in lptr[k] .. lptr[k+1] (device arrays), in list
 * order; copied to out_host (K doubles) if given. */
int amigo_sum_partials_batch(const double* part, int K, const int* lptr, const int* lidx,
                             double* out_dev, double* out_host, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* AMIGO_ASSEMBLY_H */
