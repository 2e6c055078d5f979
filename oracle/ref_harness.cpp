// Test-only harness: exposes the reference's own SparseLDL / SparseCholesky /
// BlockAMD (compiled from the reference headers where they lie, never copied)
// through a C ABI so the tests can pin the host symbolic analysis and the
// device numeric factorization against it.
//
// Reference interfaces wrapped here:
//   include/sparse_ldl.h:65-206        SparseLDL ctor/factor/solve/get_inertia
//   include/sparse_ldl.h:2204-2358     symbolic_analysis (etree/post/colcount)
//   include/ordering_utils.h:38-56     OrderingUtils::reorder
//   include/ordering_utils.h:313-360   OrderingUtils::copy_for_reorder
//   include/sparse_cholesky.h:64-376   SparseCholesky
//
// This file is ORACLE infrastructure: only tests/, smoke() and bench.py's
// cpu_baseline leg load the library it builds.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <deque>
#include <iomanip>
#include <iostream>
#include <memory>
#include <mutex>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include <mpi.h>
#include <omp.h>

// The harness reads the symbolic arrays that SparseLDL keeps as private
// members; the standard headers are included above so that only the
// reference classes are affected.
#define private public
#define protected public
#include "sparse_cholesky.h"
#include "sparse_ldl.h"
#undef private
#undef protected

using namespace amigo;

namespace {
struct RefLDL {
  std::shared_ptr<CSRMat<double>> mat;
  std::unique_ptr<SparseLDL<double>> ldl;
  int n = 0;
  int nnz = 0;
};

struct RefChol {
  std::shared_ptr<CSRMat<double>> mat;
  std::unique_ptr<SparseCholesky<double>> chol;
  int n = 0;
  int nnz = 0;
};

std::shared_ptr<CSRMat<double>> make_mat(int n, const int* rowp,
                                         const int* cols, int sqdef_index) {
  int nnz = rowp[n];
  int* rp = new int[n + 1];
  int* cl = new int[nnz];
  std::copy(rowp, rowp + n + 1, rp);
  std::copy(cols, cols + nnz, cl);
  return std::make_shared<CSRMat<double>>(n, n, nnz, rp, cl,
                                          MemoryLocation::HOST_ONLY, nullptr,
                                          nullptr, sqdef_index);
}

OrderingType to_order(int order) {
  switch (order) {
    case 0:
      return OrderingType::NESTED_DISSECTION;
    case 1:
      return OrderingType::AMD;
    case 2:
      return OrderingType::MULTI_COLOR;
    default:
      return OrderingType::NATURAL;
  }
}
}  // namespace

extern "C" {

// perm/iperm as SparseLDL's symbolic analysis computes them
// (copy_for_reorder + reorder).  order: 0=ND, 1=AMD, 3=NATURAL.
int ref_order(int order, int n, const int* rowp, const int* cols, int* perm,
              int* iperm) {
  OrderingType ot = to_order(order);
  int *rp = nullptr, *cl = nullptr, *p = nullptr, *ip = nullptr;
  OrderingUtils::copy_for_reorder(ot, n, rowp, cols, &rp, &cl);
  OrderingUtils::reorder(ot, n, rp, cl, &p, &ip);
  std::copy(p, p + n, perm);
  std::copy(ip, ip + n, iperm);
  delete[] rp;
  delete[] cl;
  delete[] p;
  delete[] ip;
  return 0;
}

// Raw BlockAMD::amd on an already-prepared pattern (diagonal included,
// rows sorted) with an optional multiplier list.
int ref_block_amd(int n, const int* rowp, const int* cols, int nmult,
                  const int* mult, int* perm) {
  int* rp = new int[n + 1];
  int* cl = new int[rowp[n]];
  std::copy(rowp, rowp + n + 1, rp);
  std::copy(cols, cols + rowp[n], cl);
  std::vector<int> m(mult, mult + nmult);
  auto st = BlockAMD::amd(n, rp, cl, nmult, m.data(), perm, 0);
  delete[] rp;
  delete[] cl;
  return st == BlockAMD::AMDStatus::SUCCESS ? 0 : 1;
}

// OrderingUtils::create_csr_from_element_conn (ordering_utils.h:374-470) as
// reorder_model calls it (amigo/amigo.cpp:228-264: no diagonal, sorted rows,
// no linear elements): element groups one after another in conn, group g
// with ne[g] elements of npe[g] nodes whose last ncon[g] are constraint
// nodes.  Returns nnz; fills rowp (nrows + 1) and, when nnz <= cap, cols.
int ref_element_csr(int nrows, int ngroups, const int* ne, const int* npe, const int* ncon,
                    const int* conn, int* rowp_out, int* cols_out, int cap) {
  std::vector<int> start(ngroups + 1, 0);
  std::vector<size_t> off(ngroups + 1, 0);
  for (int g = 0; g < ngroups; g++) {
    start[g + 1] = start[g] + ne[g];
    off[g + 1] = off[g] + (size_t)ne[g] * npe[g];
  }
  auto nodes = [&](int e, const int** ptr, int* nc = nullptr, bool* lin = nullptr) {
    const int g = int(std::upper_bound(start.begin(), start.end(), e) - start.begin()) - 1;
    *ptr = conn + off[g] + (size_t)(e - start[g]) * npe[g];
    if (nc) *nc = ncon[g];
    if (lin) *lin = false;
    return npe[g];
  };
  int *rp = nullptr, *cl = nullptr;
  OrderingUtils::create_csr_from_element_conn(nrows, nrows, start[ngroups], nodes, false, true,
                                              &rp, &cl);
  const int nnz = rp[nrows];
  std::copy(rp, rp + nrows + 1, rowp_out);
  if (nnz <= cap) std::copy(cl, cl + nnz, cols_out);
  delete[] rp;
  delete[] cl;
  return nnz;
}

void* ref_ldl_create(int n, const int* rowp, const int* cols, int cholesky,
                     double ustab, double pivot_tol, double delay_growth,
                     int order) {
  auto* h = new RefLDL;
  h->n = n;
  h->nnz = rowp[n];
  h->mat = make_mat(n, rowp, cols, -1);
  auto st = cholesky ? SparseLDL<double>::SolverType::CHOLESKY
                     : SparseLDL<double>::SolverType::LDL;
  h->ldl.reset(new SparseLDL<double>(h->mat, st, ustab, pivot_tol,
                                     delay_growth, to_order(order)));
  return h;
}

void ref_ldl_destroy(void* p) { delete static_cast<RefLDL*>(p); }

int ref_ldl_num_nodes(void* p) {
  return static_cast<RefLDL*>(p)->ldl->num_nodes;
}

int ref_ldl_contrib_total(void* p) {
  auto* h = static_cast<RefLDL*>(p);
  return h->ldl->contrib_ptr[h->ldl->num_nodes];
}

// node_ptr[K+1], node_to_var[n], children_ptr[K+2], children[K],
// contrib_ptr[K+1], contrib_rows[contrib_total], invperm[n] (identity when
// the ordering is NATURAL).
void ref_ldl_get_symbolic(void* p, int* node_ptr, int* node_to_var,
                          int* children_ptr, int* children, int* contrib_ptr,
                          int* contrib_rows, int* invperm) {
  auto* h = static_cast<RefLDL*>(p);
  auto& L = *h->ldl;
  int K = L.num_nodes;
  std::copy(L.node_ptr, L.node_ptr + K + 1, node_ptr);
  std::copy(L.node_to_var, L.node_to_var + h->n, node_to_var);
  std::copy(L.node_children_ptr, L.node_children_ptr + K + 2, children_ptr);
  std::copy(L.node_children, L.node_children + K, children);
  std::copy(L.contrib_ptr, L.contrib_ptr + K + 1, contrib_ptr);
  std::copy(L.contrib_rows, L.contrib_rows + L.contrib_ptr[K], contrib_rows);
  if (L.invperm) {
    std::copy(L.invperm, L.invperm + h->n, invperm);
  } else {
    for (int i = 0; i < h->n; i++) invperm[i] = i;
  }
}

// Etree, postorder and column counts in the permuted ordering, computed by
// the reference's own member functions for the permutation given.
void ref_etree_post_colcount(void* p, const int* perm, const int* iperm,
                             int* parent, int* post, int* colcount) {
  auto* h = static_cast<RefLDL*>(p);
  auto& L = *h->ldl;
  int n = h->n;
  const int *rowp, *cols;
  h->mat->get_data(nullptr, nullptr, nullptr, &rowp, &cols,
                   (const double**)nullptr);
  std::vector<int> work(4 * n);
  L.compute_etree(n, rowp, cols, perm, iperm, parent, work.data());
  L.post_order_etree(n, parent, post, work.data());
  L.count_column_nonzeros(n, rowp, cols, post, perm, iperm, parent, colcount,
                          work.data());
}

int ref_ldl_factor(void* p, const double* data) {
  auto* h = static_cast<RefLDL*>(p);
  double* d = h->mat->get_data_ptr();
  std::copy(data, data + h->nnz, d);
  return h->ldl->factor();
}

void ref_ldl_inertia(void* p, int* npos, int* nneg) {
  static_cast<RefLDL*>(p)->ldl->get_inertia(npos, nneg);
}

void ref_ldl_solve(void* p, int nrhs, double* x, int ldx) {
  static_cast<RefLDL*>(p)->ldl->solve(nrhs, x, ldx);
}

// Per node: number of pivots and delayed pivots of the last factor().
void ref_ldl_pivot_counts(void* p, int* num_pivots, int* num_delayed) {
  auto* h = static_cast<RefLDL*>(p);
  auto& L = *h->ldl;
  for (int k = 0; k < L.num_nodes; k++) {
    int np = 0, nd = 0;
    L.fact.get_factor(k, &np, nullptr, &nd, nullptr, nullptr);
    num_pivots[k] = np;
    num_delayed[k] = nd;
  }
}

void ref_set_num_threads(int nt) { omp_set_num_threads(nt); }

void* ref_chol_create(int n, const int* rowp, const int* cols,
                      int sqdef_index) {
  auto* h = new RefChol;
  h->n = n;
  h->nnz = rowp[n];
  h->mat = make_mat(n, rowp, cols, sqdef_index);
  h->chol.reset(new SparseCholesky<double>(h->mat));
  return h;
}

void ref_chol_destroy(void* p) { delete static_cast<RefChol*>(p); }

int ref_chol_factor(void* p, const double* data) {
  auto* h = static_cast<RefChol*>(p);
  double* d = h->mat->get_data_ptr();
  std::copy(data, data + h->nnz, d);
  return h->chol->factor();
}

void ref_chol_solve(void* p, double* x) {
  auto* h = static_cast<RefChol*>(p);
  Vector<double> v(h->n, 0, MemoryLocation::HOST_ONLY);
  std::copy(x, x + h->n, v.get_array());
  h->chol->solve(&v);
  std::copy(v.get_array(), v.get_array() + h->n, x);
}

}  // extern "C"
