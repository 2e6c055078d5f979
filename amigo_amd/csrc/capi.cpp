// C ABI of the amigo_amd library (declarations in include/amigo_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/amigo_amd.h"
#include "ldl_numeric.h"
#include "symbolic.h"

using namespace amigo_amd;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

bool check_csr(int n, const int* rowp, const int* cols) {
  if (n < 0 || !rowp || (n > 0 && !cols)) return false;
  if (rowp[0] != 0) return false;
  for (int i = 0; i < n; i++) {
    if (rowp[i + 1] < rowp[i]) return false;
    for (int p = rowp[i]; p < rowp[i + 1]; p++) {
      if (cols[p] < 0 || cols[p] >= n) return false;
      if (p > rowp[i] && cols[p] <= cols[p - 1]) return false;
    }
  }
  return true;
}
}  // namespace

struct amigo_ldl {
  int n = 0, nnz = 0;
  int solver_type = AMIGO_SOLVER_LDL;
  double ustab = 0.01, pivot_tol = 1e-14, delay_growth = 2.0;
  int order = AMIGO_ORDER_NATURAL;
  SymbolicPlan plan;
  LdlNumeric num;
  bool device_ready = false;
  int scaling = 0;
  int sqdef = -1;  // quasi-definite split of a SparseCholesky handle (-1: none)
  std::vector<int> rowp, cols;
};

namespace {
int ensure_device(amigo_ldl* h) {
  if (h->device_ready) return AMIGO_OK;
  std::string err;
  int rc = h->num.setup(h->plan, h->rowp.data(), h->cols.data(),
                        h->delay_growth, err);
  if (rc != 0) return fail(rc, "device setup: " + err);
  h->device_ready = true;
  return AMIGO_OK;
}
}  // namespace

extern "C" {

const char* amigo_last_error(void) { return g_err.c_str(); }

int amigo_order(int order, int n, const int* rowp, const int* cols, int* perm,
                int* iperm) {
  if (!check_csr(n, rowp, cols) || !perm || !iperm)
    return fail(AMIGO_ERR_ARG, "amigo_order: invalid CSR or output");
  try {
    std::vector<int> p, ip;
    compute_ordering(order, n, rowp, cols, p, ip);
    std::memcpy(perm, p.data(), sizeof(int) * n);
    std::memcpy(iperm, ip.data(), sizeof(int) * n);
  } catch (const std::exception& e) {
    return fail(AMIGO_ERR_NOMEM, e.what());
  }
  return AMIGO_OK;
}

int amigo_block_amd(int n, const int* rowp, const int* cols, int nmult,
                    const int* mult, int* perm) {
  if (!check_csr(n, rowp, cols) || !perm || nmult < 0 || (nmult > 0 && !mult))
    return fail(AMIGO_ERR_ARG, "amigo_block_amd: invalid arguments");
  try {
    std::vector<int> rp(rowp, rowp + n + 1), cl(cols, cols + rowp[n]);
    std::vector<int> mm(mult, mult + nmult);
    for (int v : mm)
      if (v < 0 || v >= n) return fail(AMIGO_ERR_ARG, "multiplier index out of range");
    if (block_amd(n, rp.data(), cl.data(), nmult, mm.data(), perm) != 0)
      return fail(AMIGO_ERR_ARG, "quotient graph rows are not sorted");
  } catch (const std::exception& e) {
    return fail(AMIGO_ERR_NOMEM, e.what());
  }
  return AMIGO_OK;
}

int amigo_etree_post_colcount(int n, const int* rowp, const int* cols,
                              const int* perm, const int* iperm, int* parent,
                              int* post, int* colcount) {
  if (!check_csr(n, rowp, cols) || !parent || !post || !colcount)
    return fail(AMIGO_ERR_ARG, "amigo_etree_post_colcount: invalid arguments");
  etree(n, rowp, cols, perm, iperm, parent);
  postorder(n, parent, post);
  column_counts(n, rowp, cols, post, perm, iperm, parent, colcount);
  return AMIGO_OK;
}

int amigo_ldl_create(int n, const int* rowp, const int* cols, int solver_type,
                     double ustab, double pivot_tol, double delay_growth,
                     int order, amigo_ldl** out) {
  return amigo_ldl_create_grouped(n, rowp, cols, solver_type, ustab, pivot_tol,
                                  delay_growth, order, nullptr, out);
}

namespace {
int create_handle(int n, const int* rowp, const int* cols, int solver_type,
                  double ustab, double pivot_tol, double delay_growth, int order,
                  const int* var_group, const int* user_perm, amigo_ldl** out) {
  if (!out) return fail(AMIGO_ERR_ARG, "amigo_ldl_create: null output");
  *out = nullptr;
  if (!check_csr(n, rowp, cols))
    return fail(AMIGO_ERR_ARG,
                "amigo_ldl_create: CSR must have sorted, in-range column indices");
  if (order < 0 || order > 3) return fail(AMIGO_ERR_ARG, "unknown ordering");
  if (solver_type != AMIGO_SOLVER_LDL && solver_type != AMIGO_SOLVER_CHOLESKY)
    return fail(AMIGO_ERR_ARG, "unknown solver type");
  if (user_perm) {
    std::vector<char> seen(n, 0);
    for (int i = 0; i < n; i++) {
      const int v = user_perm[i];
      if (v < 0 || v >= n || seen[v])
        return fail(AMIGO_ERR_ARG, "amigo_ldl_create_perm: perm is not a permutation");
      seen[v] = 1;
    }
  }
  std::unique_ptr<amigo_ldl> h;
  try {
    h.reset(new amigo_ldl);
  } catch (...) {
    return fail(AMIGO_ERR_NOMEM, "out of host memory");
  }
  // Parameter clamps of SparseLDL's constructor (sparse_ldl.h:76-96).
  h->n = n;
  h->nnz = rowp[n];
  h->solver_type = solver_type;
  h->ustab = ustab > 0.5 ? 0.5 : (ustab < 0.0 ? 0.0 : ustab);
  h->pivot_tol = pivot_tol;
  if (solver_type == AMIGO_SOLVER_CHOLESKY) {
    delay_growth = 1.0;
  } else {
    delay_growth = delay_growth < 1.0 ? 1.0 : (delay_growth > 10.0 ? 10.0 : delay_growth);
  }
  h->delay_growth = delay_growth;
  h->order = order;
  try {
    if (var_group)
      for (int i = 0; i < n; i++)
        if (var_group[i] < 0 || var_group[i] >= n)
          return fail(AMIGO_ERR_ARG, "variable group id out of range");
    symbolic_analysis(order, n, rowp, cols, h->plan,
                      order == AMIGO_ORDER_NESTED_DISSECTION ? var_group : nullptr,
                      user_perm);
  } catch (const std::exception& e) {
    return fail(AMIGO_ERR_NOMEM, std::string("symbolic analysis: ") + e.what());
  }
  // The device plan is uploaded at the first factorization; keep the
  // pattern for it.
  h->rowp.assign(rowp, rowp + n + 1);
  h->cols.assign(cols, cols + rowp[n]);
  *out = h.release();
  return AMIGO_OK;
}
}  // namespace

int amigo_ldl_create_grouped(int n, const int* rowp, const int* cols,
                             int solver_type, double ustab, double pivot_tol,
                             double delay_growth, int order,
                             const int* var_group, amigo_ldl** out) {
  return create_handle(n, rowp, cols, solver_type, ustab, pivot_tol, delay_growth,
                       order, var_group, nullptr, out);
}

int amigo_ldl_create_perm(int n, const int* rowp, const int* cols, int solver_type,
                          double ustab, double pivot_tol, double delay_growth,
                          const int* perm, amigo_ldl** out) {
  if (!perm) return fail(AMIGO_ERR_ARG, "amigo_ldl_create_perm: null perm");
  return create_handle(n, rowp, cols, solver_type, ustab, pivot_tol, delay_growth,
                       AMIGO_ORDER_AMD, nullptr, perm, out);
}

int amigo_ldl_create_union(int n, const int* rowp, const int* cols, int num_blocks,
                           const amigo_ldl* const* blocks, amigo_ldl** out) {
  if (!out) return fail(AMIGO_ERR_ARG, "amigo_ldl_create_union: null output");
  *out = nullptr;
  if (num_blocks < 1 || !blocks)
    return fail(AMIGO_ERR_ARG, "amigo_ldl_create_union: need at least one block");
  if (!check_csr(n, rowp, cols))
    return fail(AMIGO_ERR_ARG,
                "amigo_ldl_create_union: CSR must have sorted, in-range column indices");
  std::vector<const SymbolicPlan*> plans(num_blocks);
  std::vector<int> voff(num_blocks);
  int v = 0;
  long long nz = 0;
  for (int b = 0; b < num_blocks; b++) {
    const amigo_ldl* B = blocks[b];
    if (!B) return fail(AMIGO_ERR_ARG, "amigo_ldl_create_union: null block handle");
    if (B->solver_type != blocks[0]->solver_type || B->ustab != blocks[0]->ustab ||
        B->pivot_tol != blocks[0]->pivot_tol || B->delay_growth != blocks[0]->delay_growth)
      return fail(AMIGO_ERR_ARG, "amigo_ldl_create_union: blocks differ in solver parameters");
    // Row lengths of the union must be the blocks' own (block-diagonal).
    for (int i = 0; i < B->n; i++)
      if (v + i >= n || rowp[v + i + 1] - rowp[v + i] != B->rowp[i + 1] - B->rowp[i])
        return fail(AMIGO_ERR_ARG, "amigo_ldl_create_union: pattern is not the blocks' diagonal union");
    plans[b] = &B->plan;
    voff[b] = v;
    v += B->n;
    nz += B->nnz;
  }
  if (v != n || nz != rowp[n])
    return fail(AMIGO_ERR_ARG, "amigo_ldl_create_union: sizes do not add up");
  std::unique_ptr<amigo_ldl> h;
  try {
    h.reset(new amigo_ldl);
    const amigo_ldl* B0 = blocks[0];
    h->n = n;
    h->nnz = rowp[n];
    h->solver_type = B0->solver_type;
    h->ustab = B0->ustab;
    h->pivot_tol = B0->pivot_tol;
    h->delay_growth = B0->delay_growth;
    h->order = B0->order;
    concat_plans(plans, voff, h->plan);
    h->rowp.assign(rowp, rowp + n + 1);
    h->cols.assign(cols, cols + rowp[n]);
  } catch (const std::exception& e) {
    return fail(AMIGO_ERR_NOMEM, std::string("amigo_ldl_create_union: ") + e.what());
  }
  *out = h.release();
  return AMIGO_OK;
}

int amigo_ldl_destroy(amigo_ldl* h) {
  delete h;
  return AMIGO_OK;
}

int amigo_ldl_sizes(const amigo_ldl* h, int* num_nodes, int* contrib_total,
                    int* num_levels, int* max_front) {
  if (!h) return fail(AMIGO_ERR_ARG, "null handle");
  if (num_nodes) *num_nodes = h->plan.num_nodes;
  if (contrib_total) *contrib_total = int(h->plan.contrib_rows.size());
  if (num_levels || max_front) {
    // Height of the node tree and largest static front (host-side).
    const SymbolicPlan& P = h->plan;
    std::vector<int> lev(P.num_nodes, 0);
    int nl = 0, mf = 0;
    for (int k = 0; k < P.num_nodes; k++) {
      for (int q = P.children_ptr[k]; q < P.children_ptr[k + 1]; q++)
        lev[k] = std::max(lev[k], lev[P.children[q]] + 1);
      nl = std::max(nl, lev[k] + 1);
      mf = std::max(mf, P.node_ptr[k + 1] - P.node_ptr[k] + P.contrib_ptr[k + 1] -
                            P.contrib_ptr[k]);
    }
    if (num_levels) *num_levels = nl;
    if (max_front) *max_front = mf;
  }
  return AMIGO_OK;
}

int amigo_ldl_get_symbolic(const amigo_ldl* h, int* node_ptr, int* node_to_var,
                           int* children_ptr, int* children, int* contrib_ptr,
                           int* contrib_rows, int* invperm) {
  if (!h) return fail(AMIGO_ERR_ARG, "null handle");
  const SymbolicPlan& P = h->plan;
  auto cp = [](int* dst, const std::vector<int>& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), sizeof(int) * v.size());
  };
  cp(node_ptr, P.node_ptr);
  cp(node_to_var, P.node_to_var);
  cp(children_ptr, P.children_ptr);
  cp(children, P.children);
  cp(contrib_ptr, P.contrib_ptr);
  cp(contrib_rows, P.contrib_rows);
  if (invperm) {
    if (P.iperm.empty()) {
      for (int i = 0; i < h->n; i++) invperm[i] = i;
    } else {
      cp(invperm, P.iperm);
    }
  }
  return AMIGO_OK;
}

int amigo_ldl_factor_shifted(amigo_ldl* h, const double* values_dev,
                             const double* shift_dev, int precision,
                             void* stream) {
  if (!h || !values_dev) return fail(AMIGO_ERR_ARG, "amigo_ldl_factor: null argument");
  if (precision != AMIGO_PREC_FP64 && precision != AMIGO_PREC_MIXED)
    return fail(AMIGO_ERR_ARG, "unknown precision");
  int rd = ensure_device(h);
  if (rd) return rd;
  std::string err;
  int rc = h->num.factor(values_dev, shift_dev, h->ustab, h->pivot_tol,
                         h->solver_type == AMIGO_SOLVER_CHOLESKY,
                         precision == AMIGO_PREC_MIXED, static_cast<hipStream_t>(stream), err);
  if (rc != 0) return fail(rc, "amigo_ldl_factor: " + err);
  return AMIGO_OK;
}

int amigo_ldl_factor(amigo_ldl* h, const double* values_dev, int precision,
                     void* stream) {
  return amigo_ldl_factor_shifted(h, values_dev, nullptr, precision, stream);
}

int amigo_chol_create(int n, const int* rowp, const int* cols, int sqdef_index,
                      amigo_ldl** out) {
  const int rc = create_handle(n, rowp, cols, AMIGO_SOLVER_CHOLESKY, 0.0, 0.0, 1.0,
                               AMIGO_ORDER_NATURAL, nullptr, nullptr, out);
  if (rc) return rc;
  // SparseCholesky's constructor: a split outside [0, n] means the whole
  // matrix is the definite block (sparse_cholesky.h:73-76).
  (*out)->sqdef = (sqdef_index < 0 || sqdef_index > n) ? n : sqdef_index;
  return AMIGO_OK;
}

int amigo_chol_factor(amigo_ldl* h, const double* values_dev, void* stream) {
  if (!h) return fail(AMIGO_ERR_ARG, "amigo_chol_factor: null handle");
  if (h->solver_type != AMIGO_SOLVER_CHOLESKY || h->sqdef < 0)
    return fail(AMIGO_ERR_ARG, "amigo_chol_factor: handle not made by amigo_chol_create");
  int rc = amigo_ldl_factor_shifted(h, values_dev, nullptr, AMIGO_PREC_FP64, stream);
  if (rc) return rc;
  std::string err;
  int bad = 0;
  rc = h->num.sqd_check(h->sqdef, &bad, static_cast<hipStream_t>(stream), err);
  if (rc) return fail(rc, "amigo_chol_factor: " + err);
  return bad;
}

int amigo_ldl_set_batch(amigo_ldl* h, int num_blocks, const int* node_seg) {
  if (!h) return fail(AMIGO_ERR_ARG, "null handle");
  if (num_blocks < 1 || !node_seg)
    return fail(AMIGO_ERR_ARG, "amigo_ldl_set_batch: need num_blocks >= 1 and node_seg");
  if (node_seg[0] != 0 || node_seg[num_blocks] != h->plan.num_nodes)
    return fail(AMIGO_ERR_ARG, "amigo_ldl_set_batch: node_seg must run from 0 to num_nodes");
  for (int b = 0; b < num_blocks; b++)
    if (node_seg[b + 1] < node_seg[b])
      return fail(AMIGO_ERR_ARG, "amigo_ldl_set_batch: node_seg must be nondecreasing");
  int rc = ensure_device(h);
  if (rc) return rc;
  std::string err;
  rc = h->num.set_batch(num_blocks, node_seg, err);
  if (rc) return fail(rc, "set_batch: " + err);
  return AMIGO_OK;
}

int amigo_ldl_set_batch_mask(amigo_ldl* h, const int* mask, void* stream) {
  if (!h) return fail(AMIGO_ERR_ARG, "null handle");
  if (h->num.batch_blocks() == 0) return fail(AMIGO_ERR_STATE, "no batch segments set");
  std::string err;
  int rc = h->num.set_batch_mask(mask, (hipStream_t)stream, err);
  if (rc) return fail(rc, "set_batch_mask: " + err);
  return AMIGO_OK;
}

int amigo_ldl_get_batch_inertia(const amigo_ldl* h, int* npos, int* nneg, int* nzero,
                                int* nnan) {
  if (!h) return fail(AMIGO_ERR_ARG, "null handle");
  if (!h->num.factored()) return fail(AMIGO_ERR_STATE, "no factorization");
  if (h->num.batch_blocks() == 0) return fail(AMIGO_ERR_STATE, "no batch segments set");
  h->num.batch_inertia(npos, nneg, nzero, nnan);
  if (h->solver_type == AMIGO_SOLVER_CHOLESKY) {
    const int B = h->num.batch_blocks();
    for (int b = 0; b < B; b++) {
      if (npos) npos[b] = 0;
      if (nneg) nneg[b] = 0;
      if (nzero) nzero[b] = 0;
    }
  }
  return AMIGO_OK;
}

int amigo_ldl_get_inertia(const amigo_ldl* h, int* npos, int* nneg,
                          int* nzero) {
  if (!h) return fail(AMIGO_ERR_ARG, "null handle");
  if (!h->num.factored()) return fail(AMIGO_ERR_STATE, "no factorization");
  int p, q, z;
  h->num.inertia(&p, &q, &z);
  if (h->solver_type == AMIGO_SOLVER_CHOLESKY) p = q = z = 0;
  if (npos) *npos = p;
  if (nneg) *nneg = q;
  if (nzero) *nzero = z;
  return AMIGO_OK;
}

int amigo_ldl_pivot_counts(const amigo_ldl* h, int* num_pivots,
                           int* num_delayed) {
  if (!h || !num_pivots || !num_delayed) return fail(AMIGO_ERR_ARG, "null argument");
  if (!h->num.factored()) return fail(AMIGO_ERR_STATE, "no factorization");
  std::string err;
  int rc = h->num.pivot_counts(num_pivots, num_delayed, err);
  if (rc) return fail(rc, err);
  return AMIGO_OK;
}

int amigo_ldl_solve(amigo_ldl* h, int nrhs, double* x_dev, int ldx,
                    void* stream) {
  if (!h || !x_dev || nrhs < 0 || (nrhs > 0 && ldx < h->n))
    return fail(AMIGO_ERR_ARG, "amigo_ldl_solve: invalid arguments");
  std::string err;
  int rc = h->num.solve(nrhs, x_dev, ldx, static_cast<hipStream_t>(stream), err);
  if (rc) return fail(rc, "amigo_ldl_solve: " + err);
  return AMIGO_OK;
}

int amigo_ldl_set_scaling(amigo_ldl* h, int mode) {
  if (!h || mode < 0 || mode > 1) return fail(AMIGO_ERR_ARG, "amigo_ldl_set_scaling: bad argument");
  h->scaling = mode;
  h->num.set_scaling(mode);
  return AMIGO_OK;
}

int amigo_ldl_set_timing(amigo_ldl* h, int on) {
  if (!h) return fail(AMIGO_ERR_ARG, "null handle");
  h->num.set_timing(on != 0);
  if (on == 2) h->num.reset_timing();
  return AMIGO_OK;
}

int amigo_ldl_get_timing(const amigo_ldl* h, double* ms6, long long* count6,
                         double* flops_small, double* flops_big) {
  if (!h || !ms6 || !count6) return fail(AMIGO_ERR_ARG, "null argument");
  h->num.get_timing(ms6, count6);
  if (flops_small) *flops_small = h->num.flops_small();
  if (flops_big) *flops_big = h->num.flops_big();
  return AMIGO_OK;
}

int amigo_ldl_get_stats(const amigo_ldl* h, int* num_delayed, int* max_front) {
  if (!h) return fail(AMIGO_ERR_ARG, "null handle");
  if (!h->num.factored()) return fail(AMIGO_ERR_STATE, "no factorization");
  int d = 0, m = 0;
  h->num.stats(&d, &m);
  if (num_delayed) *num_delayed = d;
  if (max_front) *max_front = m;
  return AMIGO_OK;
}

int amigo_ldl_work(const amigo_ldl* h, double* factor_flops,
                   double* solve_bytes) {
  if (!h) return fail(AMIGO_ERR_ARG, "null handle");
  // Algorithmic counts from the symbolic, no-delay structure.
  const SymbolicPlan& P = h->plan;
  double f = 0.0, b = 0.0;
  for (int k = 0; k < P.num_nodes; k++) {
    const double ns = P.node_ptr[k + 1] - P.node_ptr[k];
    const double m = ns + (P.contrib_ptr[k + 1] - P.contrib_ptr[k]);
    for (int q = 0; q < (int)ns; q++) {
      const double r = m - q - 1;
      f += r + r * (m - q);
    }
    b += 8.0 * 2.0 * m * ns + 8.0 * 2.0 * m + 4.0 * m;
  }
  if (factor_flops) *factor_flops = f;
  if (solve_bytes) *solve_bytes = b;
  return AMIGO_OK;
}

int amigo_ldl_set_graphs(amigo_ldl* h, int on) {
  if (!h) return fail(AMIGO_ERR_ARG, "null handle");
  h->num.set_graphs(on != 0);
  return AMIGO_OK;
}

int amigo_ldl_memory(const amigo_ldl* h, long long* bytes) {
  if (!h || !bytes) return fail(AMIGO_ERR_ARG, "null argument");
  // Device bytes the numeric phase allocates for this plan: front pools
  // (delay_growth x static sizes, the sizing rule of LdlNumeric::setup), the
  // index / position pools, the forward-solve update vectors (one RHS), the
  // panel workspace of the widest level, and the static plan arrays.
  const SymbolicPlan& P = h->plan;
  const double g = std::max(1.0, h->delay_growth);
  double r = 0.0, iv = 0.0;
  for (int k = 0; k < P.num_nodes; k++) {
    const double m = (P.node_ptr[k + 1] - P.node_ptr[k]) + (P.contrib_ptr[k + 1] - P.contrib_ptr[k]);
    r += m * m;
    iv += m;
  }
  const double rcap = g * r + 4096.0, icap = g * iv + 4096.0;
  double total = 8.0 * rcap + 2.0 * 4.0 * icap + 8.0 * icap;
  total += 8.0 * 68.0 * (iv + 64.0 * P.num_nodes);          // panel workspace bound
  total += 4.0 * (6.0 * P.num_nodes + 3.0 * P.contrib_rows.size() + 4.0 * h->nnz + 2.0 * h->n);
  *bytes = (long long)total;
  return AMIGO_OK;
}

}  // extern "C"
