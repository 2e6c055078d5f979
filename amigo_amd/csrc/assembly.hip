// KKT assembly for component groups on gfx950.
//
// Each group is E elements of one type with NC inputs (variables and
// multipliers, gathered from x through an E x NC index array; -1 marks an
// absent input, e.g. a Dirichlet neighbour).  A thread evaluates its
// element's Lagrangian contribution alpha*f_e + lambda^T c_e, its gradient
// and its dense NC x NC Hessian into per-element buffers; deterministic
// gather kernels then sum the buffers into the gradient vector and the CSR
// values through precomputed contribution lists (no atomics, no per-entry
// binary search).  This replaces the reference's per-element Hessian-vector
// loop and add_row search (include/component_group.h:133-163,
// include/csr_matrix.h:539-552) and the CUDA atomic scatter
// (include/cuda/component_group_backend.cuh:37-120).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <cmath>
#include <string>

#include "../../include/amigo_assembly.h"

namespace {

constexpr int NT = 256;
thread_local std::string g_asm_err;

// ---- hyper-dual scalar for second derivatives of small nonlinear terms ----
// The hyper-dual arithmetic (cart-pole dynamics) is compiled without
// contraction of products into fused multiply-adds: the first-derivative
// part of each operation is then the same IEEE sequence whichever direction
// carries it and whichever other parts are known zero, so a gradient-only
// evaluation reproduces the gradient of the Hessian evaluation bit for bit.
#pragma clang fp contract(off)
struct HD {
  double v, a, b, ab;  // value, d/da, d/db, d2/dadb
};
__device__ __forceinline__ HD hd(double v) { return {v, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ HD operator+(HD x, HD y) { return {x.v + y.v, x.a + y.a, x.b + y.b, x.ab + y.ab}; }
__device__ __forceinline__ HD operator-(HD x, HD y) { return {x.v - y.v, x.a - y.a, x.b - y.b, x.ab - y.ab}; }
__device__ __forceinline__ HD operator*(HD x, HD y) {
  return {x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b,
          x.ab * y.v + x.a * y.b + x.b * y.a + x.v * y.ab};
}
__device__ __forceinline__ HD operator*(double s, HD x) { return {s * x.v, s * x.a, s * x.b, s * x.ab}; }
__device__ __forceinline__ HD inv(HD x) {
  const double f = 1.0 / x.v, f2 = -f * f, f3 = 2.0 * f * f * f;
  return {f, f2 * x.a, f2 * x.b, f2 * x.ab + f3 * x.a * x.b};
}
__device__ __forceinline__ HD operator/(HD x, HD y) { return x * inv(y); }
__device__ __forceinline__ HD hsin(HD x) {
  const double s = sin(x.v), c = cos(x.v);
  return {s, c * x.a, c * x.b, c * x.ab - s * x.a * x.b};
}
__device__ __forceinline__ HD hcos(HD x) {
  const double s = sin(x.v), c = cos(x.v);
  return {c, -s * x.a, -s * x.b, -s * x.ab - c * x.a * x.b};
}
#pragma clang fp contract(fast)

// ---- element types --------------------------------------------------------
// AMIGO_ELEM_LINK: inputs (x, lambda); c = x.  L = lambda x.
__device__ void eval_link(const double* z, const double* d, double alpha, double& L,
                          double* g, double* H) {
  (void)d;
  (void)alpha;
  L = z[1] * z[0];
  g[0] = z[1];
  g[1] = z[0];
  H[0] = 0.0;
  H[1] = 1.0;
  H[2] = 1.0;
  H[3] = 0.0;
}

// AMIGO_ELEM_DI_INTERVAL: double integrator, trapezoidal collocation.
// inputs (q0, v0, u0, q1, v1, u1, lq, lv); data (dt).
// f_e = (u0^2 + u1^2) dt / 4
// c_q = q1 - q0 - dt/2 (v0 + v1),  c_v = v1 - v0 - dt/2 (u0 + u1)
__device__ void eval_di(const double* z, const double* d, double alpha, double& L,
                        double* g, double* H) {
  const double dt = d[0], h = 0.5 * dt;
  const double q0 = z[0], v0 = z[1], u0 = z[2], q1 = z[3], v1 = z[4], u1 = z[5];
  const double lq = z[6], lv = z[7];
  const double cq = q1 - q0 - h * (v0 + v1);
  const double cv = v1 - v0 - h * (u0 + u1);
  L = alpha * (u0 * u0 + u1 * u1) * dt * 0.25 + lq * cq + lv * cv;
  g[0] = -lq;
  g[1] = -h * lq - lv;
  g[2] = alpha * u0 * h - h * lv;
  g[3] = lq;
  g[4] = -h * lq + lv;
  g[5] = alpha * u1 * h - h * lv;
  g[6] = cq;
  g[7] = cv;
  for (int q = 0; q < 64; q++) H[q] = 0.0;
  H[2 * 8 + 2] = alpha * h;
  H[5 * 8 + 5] = alpha * h;
  const double jq[6] = {-1.0, -h, 0.0, 1.0, -h, 0.0};
  const double jv[6] = {0.0, -1.0, -h, 0.0, 1.0, -h};
  for (int q = 0; q < 6; q++) {
    H[6 * 8 + q] = H[q * 8 + 6] = jq[q];
    H[7 * 8 + q] = H[q * 8 + 7] = jv[q];
  }
}

// AMIGO_ELEM_ELLIPTIC2D / 3D: one grid node of -Lap(y) + y^3 - u = 0 with the
// tracking objective, the constraint multiplied by h^2 (the usual
// finite-difference scaling: Jacobian entries O(1)).
// inputs (y, y_nb[2*dim], u, lambda); data (h, yd, a).
// f_e = h^dim/2 (y - yd)^2 + a h^dim/2 u^2
// c   = (2 dim y - sum y_nb) + h^2 (y^3 - u)
template <int DIM>
__device__ void eval_elliptic(const double* z, const double* d, double alpha, double& L,
                              double* g, double* H) {
  constexpr int NB = 2 * DIM, NC = NB + 3;
  const double h = d[0], yd = d[1], a = d[2];
  const double hd_ = DIM == 2 ? h * h : h * h * h;
  const double h2 = h * h;
  const double y = z[0], u = z[NB + 1], lam = z[NB + 2];
  double s = 0.0;
  for (int q = 0; q < NB; q++) s += z[1 + q];
  const double c = (2.0 * DIM * y - s) + h2 * (y * y * y - u);
  L = alpha * (0.5 * hd_ * (y - yd) * (y - yd) + 0.5 * a * hd_ * u * u) + lam * c;
  const double dcy = 2.0 * DIM + 3.0 * h2 * y * y;
  g[0] = alpha * hd_ * (y - yd) + lam * dcy;
  for (int q = 0; q < NB; q++) g[1 + q] = -lam;
  g[NB + 1] = alpha * a * hd_ * u - lam * h2;
  g[NB + 2] = c;
  for (int q = 0; q < NC * NC; q++) H[q] = 0.0;
  H[0] = alpha * hd_ + 6.0 * lam * h2 * y;
  H[(NB + 1) * NC + NB + 1] = alpha * a * hd_;
  H[(NB + 2) * NC + 0] = H[0 * NC + NB + 2] = dcy;
  for (int q = 0; q < NB; q++) H[(NB + 2) * NC + 1 + q] = H[(1 + q) * NC + NB + 2] = -1.0;
  H[(NB + 2) * NC + NB + 1] = H[(NB + 1) * NC + NB + 2] = -h2;
}

// Cart-pole accelerations, point-mass pole of length l on a cart, angle th
// measured from hanging (th = 0 down, th = pi upright), the convention of the
// reference's examples/trajectory/cart/cart_pole.py:76-86:
//   D = mc + mp sin^2(th)
//   xdd = (u + mp sin(th) (l w^2 + g cos(th))) / D
//   thdd = (-u cos(th) - mp l w^2 sin(th) cos(th) - (mc + mp) g sin(th)) / (l D)
#pragma clang fp contract(off)
template <class T>
__device__ void cp_accel(T th, T w, T u, double mc, double mp, double l, T& xdd, T& thdd) {
  const double gr = 9.81;
  T s = hsin(th), c = hcos(th);
  T D = hd(mc) + mp * (s * s);
  xdd = (u + mp * (s * (l * (w * w) + gr * c))) / D;
  thdd = ((-1.0) * (u * c) - (mp * l) * ((w * w) * (s * c)) - ((mc + mp) * gr) * s) / (l * D);
}

// AMIGO_ELEM_CARTPOLE: trapezoidal collocation interval.
// inputs (x0, th0, xd0, w0, u0, x1, th1, xd1, w1, u1, l1, l2, l3, l4);
// data (dt, mc, mp, l).
// f_e = (u0^2 + u1^2) dt / 2
// c = s1 - s0 - dt/2 (f(s0, u0) + f(s1, u1)),  f = (xd, w, xdd, thdd)
// WANT_H = false (gradient only): one hyper-dual direction per variable
// instead of one pair per Hessian entry.  The first-derivative part of every
// hyper-dual operation reads only values and first-direction parts, so the
// gradient is the same as the one the Hessian path computes.
template <bool WANT_H>
__device__ void eval_cartpole(const double* z, const double* d, double alpha, double& L,
                              double* g, double* H) {
  constexpr int NC = 14;
  const double dt = d[0], mc = d[1], mp = d[2], l = d[3], h = 0.5 * dt;
  if constexpr (WANT_H)
    for (int q = 0; q < NC * NC; q++) H[q] = 0.0;
  for (int q = 0; q < NC; q++) g[q] = 0.0;
  const double* lam = z + 10;
  L = alpha * (z[4] * z[4] + z[9] * z[9]) * dt * 0.5;
  g[4] += alpha * z[4] * dt;
  g[9] += alpha * z[9] * dt;
  if constexpr (WANT_H) {
    H[4 * NC + 4] += alpha * dt;
    H[9 * NC + 9] += alpha * dt;
  }
  // Linear part: lam . (s1 - s0) - h lam . ((xd0, w0) + (xd1, w1)) in rows 0,1.
  for (int r = 0; r < 4; r++) {
    const double cr_lin = z[5 + r] - z[r];
    L += lam[r] * cr_lin;
    g[5 + r] += lam[r];
    g[r] -= lam[r];
    g[10 + r] += cr_lin;
    if constexpr (WANT_H) {
      H[(10 + r) * NC + 5 + r] = H[(5 + r) * NC + 10 + r] = 1.0;
      H[(10 + r) * NC + r] = H[r * NC + 10 + r] = -1.0;
    }
  }
  for (int e = 0; e < 2; e++) {
    const int o = 5 * e;  // offset of (x, th, xd, w, u) for this end
    // rows 0,1 of f: xd and w (linear)
    L -= h * (lam[0] * z[o + 2] + lam[1] * z[o + 3]);
    g[o + 2] -= h * lam[0];
    g[o + 3] -= h * lam[1];
    g[10] -= h * z[o + 2];
    g[11] -= h * z[o + 3];
    if constexpr (WANT_H) {
      H[10 * NC + o + 2] -= h;
      H[(o + 2) * NC + 10] -= h;
      H[11 * NC + o + 3] -= h;
      H[(o + 3) * NC + 11] -= h;
    }
    // rows 2,3: nonlinear in (th, w, u) = inputs (o+1, o+3, o+4).
    const int vi[3] = {o + 1, o + 3, o + 4};
    double val2 = 0.0, val3 = 0.0, g2[3], g3[3], h2[3][3], h3[3][3];
    if constexpr (!WANT_H) {
      for (int p = 0; p < 3; p++) {
        HD th = hd(z[o + 1]), w = hd(z[o + 3]), u = hd(z[o + 4]);
        HD* vars[3] = {&th, &w, &u};
        vars[p]->a = 1.0;
        HD xdd, thdd;
        cp_accel(th, w, u, mc, mp, l, xdd, thdd);
        val2 = xdd.v;
        val3 = thdd.v;
        g2[p] = xdd.a;
        g3[p] = thdd.a;
      }
    } else
    for (int p = 0; p < 3; p++) {
      for (int q = p; q < 3; q++) {
        HD th = hd(z[o + 1]), w = hd(z[o + 3]), u = hd(z[o + 4]);
        HD* vars[3] = {&th, &w, &u};
        vars[p]->a = 1.0;
        vars[q]->b = 1.0;
        HD xdd, thdd;
        cp_accel(th, w, u, mc, mp, l, xdd, thdd);
        val2 = xdd.v;
        val3 = thdd.v;
        g2[p] = xdd.a;
        g3[p] = thdd.a;
        g2[q] = xdd.b;
        g3[q] = thdd.b;
        h2[p][q] = h2[q][p] = xdd.ab;
        h3[p][q] = h3[q][p] = thdd.ab;
      }
    }
    L -= h * (lam[2] * val2 + lam[3] * val3);
    g[12] -= h * val2;
    g[13] -= h * val3;
    for (int p = 0; p < 3; p++) {
      g[vi[p]] -= h * (lam[2] * g2[p] + lam[3] * g3[p]);
      if constexpr (!WANT_H) continue;
      H[12 * NC + vi[p]] -= h * g2[p];
      H[vi[p] * NC + 12] -= h * g2[p];
      H[13 * NC + vi[p]] -= h * g3[p];
      H[vi[p] * NC + 13] -= h * g3[p];
      for (int q = 0; q < 3; q++)
        H[vi[p] * NC + vi[q]] -= h * (lam[2] * h2[p][q] + lam[3] * h3[p][q]);
    }
  }
}

#pragma clang fp contract(fast)

template <int NC, bool WANT_H>
__device__ void eval_elem(int type, const double* z, const double* d, double alpha,
                          double& L, double* g, double* H) {
  if constexpr (NC == 2) {
    eval_link(z, d, alpha, L, g, H);
  } else if constexpr (NC == 8) {
    eval_di(z, d, alpha, L, g, H);
  } else if constexpr (NC == 7) {
    eval_elliptic<2>(z, d, alpha, L, g, H);
  } else if constexpr (NC == 9) {
    eval_elliptic<3>(z, d, alpha, L, g, H);
  } else {
    eval_cartpole<WANT_H>(z, d, alpha, L, g, H);
  }
}

// One thread per element: gather inputs, evaluate, store g and H (and the
// Lagrangian partial sum per block into lpart when non-null).  With
// alpha_blk, block b scales the objective by alpha_blk[b] (a batch of
// problems, each padded to whole blocks); an element whose first input is
// AMIGO_ELEM_PAD is padding and contributes nothing.
template <int NC, bool WANT_H>
__global__ __launch_bounds__(NT) void element_kernel(int type, int E, const int* __restrict__ idx,
                                                     const double* __restrict__ data, int ndata,
                                                     const double* __restrict__ x, double alpha,
                                                     const double* __restrict__ alpha_blk,
                                                     double* __restrict__ gbuf,
                                                     double* __restrict__ hbuf,
                                                     double* __restrict__ lpart,
                                                     const int* __restrict__ blk_grp,
                                                     const int* __restrict__ grp_active) {
  __shared__ double red[NT / 64];
  // A block of a problem left out (batches) writes nothing.
  if (blk_grp && !grp_active[blk_grp[blockIdx.x]]) return;
  const int e = blockIdx.x * NT + threadIdx.x;
  if (alpha_blk) alpha = alpha_blk[blockIdx.x];
  double Lval = 0.0;
  if (e < E && idx[(size_t)e * NC] != AMIGO_ELEM_PAD) {
    double z[NC], g[NC], H[NC * NC];
#pragma unroll
    for (int q = 0; q < NC; q++) {
      const int k = idx[(size_t)e * NC + q];
      z[q] = k >= 0 ? x[k] : 0.0;
    }
    eval_elem<NC, WANT_H>(type, z, data + (size_t)e * ndata, alpha, Lval, g, H);
    if (gbuf) {
#pragma unroll
      for (int q = 0; q < NC; q++) gbuf[(size_t)e * NC + q] = g[q];
    }
    if constexpr (WANT_H) {
#pragma unroll
      for (int q = 0; q < NC * NC; q++) hbuf[(size_t)e * NC * NC + q] = H[q];
    }
  }
  if (lpart) {
    double v = Lval;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int w = 0; w < NT / 64; w++) s += red[w];
      lpart[blockIdx.x] = s;
    }
  }
}

// out[i] = (accumulate ? out[i] : 0) + sum_{p in [ptr[i], ptr[i+1])} buf[src[p]]
// With seg (nseg + 1 increasing offsets of the outputs' groups), outputs of
// groups with seg_active[g] == 0 are skipped and keep their value.
__global__ void gather_sum_kernel(int n, const int* __restrict__ ptr,
                                  const long long* __restrict__ src,
                                  const double* __restrict__ buf, double* __restrict__ out,
                                  int accumulate, const long long* __restrict__ seg, int nseg,
                                  const int* __restrict__ seg_active) {
  for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
    if (seg) {
      int lo = 0, hi = nseg;  // seg[lo] <= i < seg[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (seg[mid] <= i) lo = mid; else hi = mid;
      }
      if (!seg_active[lo]) continue;
    }
    double s = accumulate ? out[i] : 0.0;
    for (int p = ptr[i]; p < ptr[i + 1]; p++) s += buf[src[p]];
    out[i] = s;
  }
}

// The listed groups' outputs only: block b serves group list[b / chunks]
// and strides over its outputs seg[g] .. seg[g+1] with the group's other
// chunks; each output is the same ordered sum as in gather_sum_kernel.
__global__ void gather_sum_listed_kernel(const int* __restrict__ ptr,
                                         const long long* __restrict__ src,
                                         const double* __restrict__ buf, double* __restrict__ out,
                                         int accumulate, const long long* __restrict__ seg,
                                         const int* __restrict__ list, int chunks) {
  const int e = blockIdx.x / chunks, c = blockIdx.x - e * chunks;
  const int g = list[e];
  const long long hi = seg[g + 1];
  for (long long i = seg[g] + (long long)c * NT + threadIdx.x; i < hi; i += (long long)chunks * NT) {
    double s = accumulate ? out[i] : 0.0;
    for (int p = ptr[i]; p < ptr[i + 1]; p++) s += buf[src[p]];
    out[i] = s;
  }
}

// Batch: problem k's sum runs over part[lidx[lptr[k] .. lptr[k+1]]] in list
// order (its blocks of every element group, groups in order), the same
// sequence sum_partials_kernel adds for the problem alone.
__global__ void sum_partials_batch_kernel(const double* part, int K, const int* lptr,
                                          const int* lidx, double* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  double s = 0.0;
  for (int p = lptr[k]; p < lptr[k + 1]; p++) s += part[lidx[p]];
  out[k] = s;
}

__global__ void sum_partials_kernel(const double* part, int nb, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[b];
    out[0] = s;
  }
}

// Generated element types: evaluator and input count by type.
struct GenType {
  amigo_elem_eval_fn fn = nullptr;
  int nc = -1;
};
GenType g_gen[AMIGO_ELEM_MAX_GENERATED];

const GenType* gen_type(int type) {
  const int t = type - AMIGO_ELEM_GENERATED;
  if (t < 0 || t >= AMIGO_ELEM_MAX_GENERATED || !g_gen[t].fn) return nullptr;
  return &g_gen[t];
}

int set_err(const char* w, hipError_t e) {
  g_asm_err = std::string(w) + ": " + hipGetErrorString(e);
  return -2;
}

}  // namespace

extern "C" {

const char* amigo_assembly_last_error(void) { return g_asm_err.c_str(); }

int amigo_elem_register(int type, int nc, amigo_elem_eval_fn fn) {
  const int t = type - AMIGO_ELEM_GENERATED;
  if (t < 0 || t >= AMIGO_ELEM_MAX_GENERATED || nc <= 0 || fn == nullptr) {
    g_asm_err = "generated element type out of range or no evaluator";
    return -1;
  }
  g_gen[t].fn = fn;
  g_gen[t].nc = nc;
  return 0;
}

int amigo_elem_num_inputs(int type) {
  if (const GenType* gt = gen_type(type)) return gt->nc;
  switch (type) {
    case AMIGO_ELEM_LINK:
      return 2;
    case AMIGO_ELEM_DI_INTERVAL:
      return 8;
    case AMIGO_ELEM_ELLIPTIC2D:
      return 7;
    case AMIGO_ELEM_ELLIPTIC3D:
      return 9;
    case AMIGO_ELEM_CARTPOLE:
      return 14;
    default:
      return -1;
  }
}

int amigo_elements_eval(int type, int E, const int* idx, const double* data, int ndata,
                        const double* x, double alpha, double* gbuf, double* hbuf,
                        double* lpart, void* stream) {
  return amigo_elements_eval_scaled(type, E, idx, data, ndata, x, alpha, nullptr, gbuf, hbuf,
                                    lpart, stream);
}

int amigo_elements_eval_scaled(int type, int E, const int* idx, const double* data, int ndata,
                               const double* x, double alpha, const double* alpha_blk,
                               double* gbuf, double* hbuf, double* lpart, void* stream) {
  return amigo_elements_eval_masked(type, E, idx, data, ndata, x, alpha, alpha_blk, nullptr,
                                    nullptr, gbuf, hbuf, lpart, stream);
}

int amigo_elements_eval_masked(int type, int E, const int* idx, const double* data, int ndata,
                               const double* x, double alpha, const double* alpha_blk,
                               const int* blk_grp, const int* grp_active, double* gbuf,
                               double* hbuf, double* lpart, void* stream) {
  if (E <= 0) return 0;
  if ((blk_grp == nullptr) != (grp_active == nullptr)) {
    g_asm_err = "blk_grp and grp_active go together";
    return -1;
  }
  if (const GenType* gt = gen_type(type)) {
    const int rc = gt->fn(E, idx, data, ndata, x, alpha, alpha_blk, blk_grp, grp_active, gbuf,
                          hbuf, lpart, stream);
    if (rc != 0) g_asm_err = "generated element evaluator failed";
    return rc;
  }
  const int nb = (E + NT - 1) / NT;
  hipStream_t s = (hipStream_t)stream;
  const bool want_h = hbuf != nullptr;
  auto launch = [&](auto nc) {
    constexpr int NC = decltype(nc)::value;
    if (want_h)
      hipLaunchKernelGGL((element_kernel<NC, true>), dim3(nb), dim3(NT), 0, s, type, E, idx,
                         data, ndata, x, alpha, alpha_blk, gbuf, hbuf, lpart, blk_grp,
                         grp_active);
    else
      hipLaunchKernelGGL((element_kernel<NC, false>), dim3(nb), dim3(NT), 0, s, type, E, idx,
                         data, ndata, x, alpha, alpha_blk, gbuf, hbuf, lpart, blk_grp,
                         grp_active);
  };
  switch (type) {
    case AMIGO_ELEM_LINK:
      launch(std::integral_constant<int, 2>{});
      break;
    case AMIGO_ELEM_DI_INTERVAL:
      launch(std::integral_constant<int, 8>{});
      break;
    case AMIGO_ELEM_ELLIPTIC2D:
      launch(std::integral_constant<int, 7>{});
      break;
    case AMIGO_ELEM_ELLIPTIC3D:
      launch(std::integral_constant<int, 9>{});
      break;
    case AMIGO_ELEM_CARTPOLE:
      launch(std::integral_constant<int, 14>{});
      break;
    default:
      g_asm_err = "unknown element type";
      return -1;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err("element_kernel", e);
  return 0;
}

int amigo_gather_sum(int n, const int* ptr, const long long* src, const double* buf,
                     double* out, int accumulate, void* stream) {
  return amigo_gather_sum_masked(n, ptr, src, buf, out, accumulate, nullptr, 0, nullptr, stream);
}

int amigo_gather_sum_masked(int n, const int* ptr, const long long* src, const double* buf,
                            double* out, int accumulate, const long long* seg, int nseg,
                            const int* seg_active, void* stream) {
  if (n <= 0) return 0;
  if ((seg == nullptr) != (seg_active == nullptr) || (seg && nseg <= 0)) {
    g_asm_err = "seg, nseg and seg_active go together";
    return -1;
  }
  int nb = (n + NT - 1) / NT;
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(gather_sum_kernel, dim3(nb), dim3(NT), 0, (hipStream_t)stream, n, ptr, src,
                     buf, out, accumulate, seg, nseg, seg_active);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err("gather_sum_kernel", e);
  return 0;
}

int amigo_gather_sum_listed(const int* ptr, const long long* src, const double* buf, double* out,
                            int accumulate, const long long* seg, const int* list, int nlist,
                            long long max_len, void* stream) {
  if (nlist <= 0 || max_len <= 0) return 0;
  if (!ptr || !src || !buf || !out || !seg || !list) {
    g_asm_err = "gather_sum_listed: null array";
    return -1;
  }
  long long chunks = (max_len + NT - 1) / NT;
  if (chunks > 64) chunks = 64;
  if ((long long)nlist * chunks > 0x7fffffffLL) {
    g_asm_err = "gather_sum_listed: list too long";
    return -1;
  }
  hipLaunchKernelGGL(gather_sum_listed_kernel, dim3(unsigned(nlist * chunks)), dim3(NT), 0,
                     (hipStream_t)stream, ptr, src, buf, out, accumulate, seg, list, int(chunks));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err("gather_sum_listed_kernel", e);
  return 0;
}

int amigo_sum_partials(const double* part, int nb, double* out_dev, double* out_host,
                       void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, s, part, nb, out_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err("sum_partials_kernel", e);
  if (out_host) {
    e = hipMemcpyAsync(out_host, out_dev, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return set_err("hipMemcpyAsync", e);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_err("hipStreamSynchronize", e);
  }
  return 0;
}

int amigo_sum_partials_batch(const double* part, int K, const int* lptr, const int* lidx,
                             double* out_dev, double* out_host, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (K <= 0) return 0;
  hipLaunchKernelGGL(sum_partials_batch_kernel, dim3((K + 63) / 64), dim3(64), 0, s, part, K,
                     lptr, lidx, out_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err("sum_partials_batch_kernel", e);
  if (out_host) {
    e = hipMemcpyAsync(out_host, out_dev, sizeof(double) * K, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return set_err("hipMemcpyAsync", e);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_err("hipStreamSynchronize", e);
  }
  return 0;
}

}  // extern "C"
