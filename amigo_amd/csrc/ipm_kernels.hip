// Interior-point vector kernels for gfx950 (the ipm:: functions of
// include/interior_point_backend.h and the wrappers of
// include/interior_point_optimizer.h:214-634).
//
// Element-wise kernels stream the vectors once.  Reductions are two-pass and
// deterministic: a grid of wave64 reductions writes per-block partials to the
// caller's workspace, a single workgroup combines them in a fixed order, and
// the results are copied to the host.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cfloat>
#include <climits>
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/amigo_ipm.h"

namespace {

constexpr int NT = 256;
constexpr int MAXB = 1024;  // partial blocks per reduction

thread_local std::string g_ipm_err;

int set_err(const char* where, hipError_t e) {
  g_ipm_err = std::string(where) + ": " + hipGetErrorString(e);
  return -2;
}

__device__ __forceinline__ bool finite_bound(double b) { return !isinf(b); }

// ---- batches of independent problems ---------------------------------------
// A batch lays K problems end to end: problem k owns primal entries
// pseg[k]..pseg[k+1], constraint entries cseg[k]..cseg[k+1] and x-length
// entries xseg[k]..xseg[k+1]; its primal / constraint index lists hold the
// problem's own (local) variable numbers.  Block b of a batched launch
// belongs to problem prob[b]; that problem's blocks are first[k] ..
// first[k] + count[k] - 1, where count[k] is the grid the problem gets when
// it runs alone.  Each block then runs exactly the loop it runs alone (same
// elements per thread, same partial slot order), so every per-problem
// result, reductions included, is bitwise the single-problem result.  All
// pointers are null for a single problem.
struct BT {
  const int* prob;
  const int* first;
  const int* count;
  const int* pseg;
  const int* cseg;
  const int* xseg;
  const int* active;    // [K] nonzero = the problem takes part (null = all)
  const int* launch;    // block numbers of a compacted launch (null = block b is blockIdx.x)
  const int* problems;  // problem numbers of a compacted final pass (null = blockIdx.x)
};

struct View {
  int k, bx, gx, slot, xo, p0, c0, n;
  bool skip;
};

__device__ __forceinline__ View view(const BT& B, amigo_ipm_problem& P, amigo_ipm_state& S) {
  View v;
  if (!B.prob) {
    v.k = 0;
    v.bx = blockIdx.x;
    v.gx = gridDim.x;
    v.slot = blockIdx.x;
    v.xo = v.p0 = v.c0 = 0;
    v.n = P.n_primal + P.n_constraints;
    v.skip = false;
    return v;
  }
  const int b = B.launch ? B.launch[blockIdx.x] : (int)blockIdx.x;
  const int k = B.prob[b];
  v.k = k;
  v.bx = b - B.first[k];
  v.gx = B.count[k];
  v.slot = b;
  v.skip = B.active && B.active[k] == 0;
  v.p0 = B.pseg[k];
  v.c0 = B.cseg[k];
  v.xo = B.xseg[k];
  v.n = B.xseg[k + 1] - v.xo;
  P.n_primal = B.pseg[k + 1] - v.p0;
  P.n_constraints = B.cseg[k + 1] - v.c0;
  P.primal_indices += v.p0;
  P.constraint_indices += v.c0;
  P.lbx += v.p0;
  P.ubx += v.p0;
  P.lbh += v.c0;
  S.xlam += v.xo;
  S.zl += v.p0;
  S.zu += v.p0;
  S.sl += v.p0;
  S.su += v.p0;
  return v;
}

__device__ __forceinline__ void shift_state(amigo_ipm_state& T, const View& v) {
  T.xlam += v.xo;
  T.zl += v.p0;
  T.zu += v.p0;
  T.sl += v.p0;
  T.su += v.p0;
}

// Scalar of the block's problem: sb[k] in a batch, s otherwise.
__device__ __forceinline__ double pick(double s, const double* sb, const View& v) {
  return sb ? sb[v.k] : s;
}

// ---- reduction helpers ------------------------------------------------------
enum Op { SUM = 0, MAX = 1, MIN = 2 };

__device__ __forceinline__ double combine(double a, double b, int op) {
  return op == SUM ? a + b : (op == MAX ? fmax(a, b) : fmin(a, b));
}
__device__ __forceinline__ double ident(int op) {
  return op == SUM ? 0.0 : (op == MAX ? -DBL_MAX : DBL_MAX);
}

template <int NV>
__device__ void block_reduce(double (&v)[NV], const int (&ops)[NV], double* red) {
#pragma unroll
  for (int q = 0; q < NV; q++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[q] = combine(v[q], __shfl_xor(v[q], off, 64), ops[q]);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int q = 0; q < NV; q++) red[wid * NV + q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < NV; q++) {
      double r = red[q];
      for (int w = 1; w < NT / 64; w++) r = combine(r, red[w * NV + q], ops[q]);
      v[q] = r;
    }
  }
}

// Final pass: ws holds nb partial rows of NV values; writes NV results to out.
template <int NV>
__global__ void finish_kernel(const double* ws, int nb, int o0, int o1, int o2,
                              int o3, double* out) {
  __shared__ double red[(NT / 64) * 4];
  const int ops[4] = {o0, o1, o2, o3};
  double v[NV];
  int opv[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    opv[q] = ops[q];
    v[q] = ident(ops[q]);
  }
  // Fixed assignment of partials to threads and a fixed tree: deterministic.
  for (int b = threadIdx.x; b < nb; b += NT)
    for (int q = 0; q < NV; q++) v[q] = combine(v[q], ws[b * NV + q], opv[q]);
  block_reduce<NV>(v, opv, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; q++) out[q] = v[q];
}

// Batched final pass: block k combines problem k's partials (rows
// first[k] .. first[k]+count[k]-1 of ws) exactly as finish_kernel combines a
// single problem's, and writes out[k*NV + q].
template <int NV>
__global__ void finish_batch_kernel(BT B, const double* ws, int o0, int o1, int o2, int o3,
                                    double* out) {
  __shared__ double red[(NT / 64) * 4];
  const int k = B.problems ? B.problems[blockIdx.x] : (int)blockIdx.x;
  if (B.active && B.active[k] == 0) return;
  const int ops[4] = {o0, o1, o2, o3};
  const int nb = B.count[k];
  ws += (size_t)B.first[k] * NV;
  double v[NV];
  int opv[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    opv[q] = ops[q];
    v[q] = ident(ops[q]);
  }
  for (int b = threadIdx.x; b < nb; b += NT)
    for (int q = 0; q < NV; q++) v[q] = combine(v[q], ws[b * NV + q], opv[q]);
  block_reduce<NV>(v, opv, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; q++) out[(size_t)k * NV + q] = v[q];
}

int nblocks(int n) {
  int b = (n + NT - 1) / NT;
  if (b < 1) b = 1;
  return b > MAXB ? MAXB : b;
}

int finish_and_copy(int NV, const int* ops, double* ws, int nb, double* host,
                    hipStream_t s) {
  double* out = ws + (size_t)MAXB * 4;
  switch (NV) {
    case 1:
      hipLaunchKernelGGL(finish_kernel<1>, dim3(1), dim3(NT), 0, s, ws, nb, ops[0], 0, 0, 0, out);
      break;
    case 2:
      hipLaunchKernelGGL(finish_kernel<2>, dim3(1), dim3(NT), 0, s, ws, nb, ops[0], ops[1], 0, 0, out);
      break;
    case 3:
      hipLaunchKernelGGL(finish_kernel<3>, dim3(1), dim3(NT), 0, s, ws, nb, ops[0], ops[1], ops[2], 0, out);
      break;
    default:
      hipLaunchKernelGGL(finish_kernel<4>, dim3(1), dim3(NT), 0, s, ws, nb, ops[0], ops[1], ops[2], ops[3], out);
      break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err("finish_kernel", e);
  if (host) {
    e = hipMemcpyAsync(host, out, sizeof(double) * NV, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return set_err("hipMemcpyAsync", e);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_err("hipStreamSynchronize", e);
  }
  return 0;
}

// ---- element-wise kernels --------------------------------------------------

// Condensed RHS (interior_point_backend.h:202-222 / :227-266).  Writes res
// for every primal and constraint row (res is zeroed for the others by the
// caller's layout: every row is one or the other).  Partials:
// [||res||^2, dual_sq, primal_sq].
__global__ void residual_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S, double mu,
                                const double* __restrict__ mub,
                                const double* __restrict__ grad,
                                double* __restrict__ res, double* ws) {
  const View V = view(B, P, S);
  if (V.skip) return;
  mu = pick(mu, mub, V);
  grad += V.xo;
  res += V.xo;
  __shared__ double red[(NT / 64) * 3];
  double v[3] = {0.0, 0.0, 0.0};
  const int ops[3] = {SUM, SUM, SUM};
  const int total = P.n_primal + P.n_constraints;
  for (int t = V.bx * NT + threadIdx.x; t < total; t += V.gx * NT) {
    if (t < P.n_primal) {
      const int i = t;
      const int idx = P.primal_indices[i];
      const double rhs_stat = grad[idx] - S.zl[i] + S.zu[i];
      v[1] += rhs_stat * rhs_stat;
      double aug = rhs_stat;
      if (finite_bound(P.lbx[i])) {
        const double gap = S.sl[i];
        aug += (gap * S.zl[i] - mu) / gap;
      }
      if (finite_bound(P.ubx[i])) {
        const double gap = S.su[i];
        aug -= (gap * S.zu[i] - mu) / gap;
      }
      res[idx] = -aug;
      v[0] += aug * aug;
    } else {
      const int j = t - P.n_primal;
      const int idx = P.constraint_indices[j];
      const double rp = grad[idx] - P.lbh[j];
      v[2] += rp * rp;
      res[idx] = -rp;
      v[0] += rp * rp;
    }
  }
  block_reduce<3>(v, ops, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; q++) ws[V.slot * 3 + q] = v[q];
}

// Sigma (interior_point_backend.h:277-288).
__global__ void diagonal_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S,
                                double* __restrict__ diag) {
  const View V = view(B, P, S);
  if (V.skip) return;
  diag += V.xo;
  const int total = P.n_primal + P.n_constraints;
  for (int t = V.bx * NT + threadIdx.x; t < total; t += V.gx * NT) {
    if (t < P.n_primal) {
      double s = 0.0;
      if (finite_bound(P.lbx[t])) s += S.zl[t] / S.sl[t];
      if (finite_bound(P.ubx[t])) s += S.zu[t] / S.su[t];
      diag[P.primal_indices[t]] = s;
    } else {
      diag[P.constraint_indices[t - P.n_primal]] = 0.0;
    }
  }
}

// Bound-dual back-substitution (interior_point_backend.h:302-321).
__global__ void bound_dual_step_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S,
                                       double mu, const double* __restrict__ mub,
                                       const double* __restrict__ px,
                                       double* __restrict__ dzl,
                                       double* __restrict__ dzu) {
  const View V = view(B, P, S);
  if (V.skip) return;
  mu = pick(mu, mub, V);
  px += V.xo;
  dzl += V.p0;
  dzu += V.p0;
  for (int i = V.bx * NT + threadIdx.x; i < P.n_primal; i += V.gx * NT) {
    const double dx = px[P.primal_indices[i]];
    double a = 0.0, b = 0.0;
    if (finite_bound(P.lbx[i])) {
      const double gap = S.sl[i];
      a = -((gap * S.zl[i] - mu) + S.zl[i] * dx) / gap;
    }
    if (finite_bound(P.ubx[i])) {
      const double gap = S.su[i];
      b = -((gap * S.zu[i] - mu) - S.zu[i] * dx) / gap;
    }
    dzl[i] = a;
    dzu[i] = b;
  }
}

// Fraction to the boundary (interior_point_backend.h:328-369): minimum step
// and the first index (loop order, lower bound before upper) attaining it;
// only candidates below 1 count, as the reference starts from 1.
__global__ void max_step_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S, double tau,
                                const double* __restrict__ taub,
                                const double* __restrict__ px,
                                const double* __restrict__ dzl,
                                const double* __restrict__ dzu, double* ws) {
  const View V = view(B, P, S);
  if (V.skip) return;
  tau = pick(tau, taub, V);
  px += V.xo;
  dzl += V.p0;
  dzu += V.p0;
  __shared__ double rv[2 * (NT / 64)];
  __shared__ long long rk[2 * (NT / 64)];
  double ax = 1.0, az = 1.0;
  long long kx = LLONG_MAX, kz = LLONG_MAX;
  for (int i = V.bx * NT + threadIdx.x; i < P.n_primal; i += V.gx * NT) {
    const double dx = px[P.primal_indices[i]];
    if (finite_bound(P.lbx[i])) {
      if (dx < 0.0) {
        const double a = -tau * S.sl[i] / dx;
        if (a < 1.0 && (a < ax || (a == ax && 2LL * i < kx))) {
          ax = a;
          kx = 2LL * i;
        }
      }
      if (dzl[i] < 0.0) {
        const double a = -tau * S.zl[i] / dzl[i];
        if (a < 1.0 && (a < az || (a == az && 2LL * i < kz))) {
          az = a;
          kz = 2LL * i;
        }
      }
    }
    if (finite_bound(P.ubx[i])) {
      if (dx > 0.0) {
        const double a = tau * S.su[i] / dx;
        if (a < 1.0 && (a < ax || (a == ax && 2LL * i + 1 < kx))) {
          ax = a;
          kx = 2LL * i + 1;
        }
      }
      if (dzu[i] < 0.0) {
        const double a = -tau * S.zu[i] / dzu[i];
        if (a < 1.0 && (a < az || (a == az && 2LL * i + 1 < kz))) {
          az = a;
          kz = 2LL * i + 1;
        }
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double oa = __shfl_xor(ax, off, 64);
    long long ok = __shfl_xor(kx, off, 64);
    if (oa < ax || (oa == ax && ok < kx)) {
      ax = oa;
      kx = ok;
    }
    double ob = __shfl_xor(az, off, 64);
    long long oz = __shfl_xor(kz, off, 64);
    if (ob < az || (ob == az && oz < kz)) {
      az = ob;
      kz = oz;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    rv[2 * wid] = ax;
    rk[2 * wid] = kx;
    rv[2 * wid + 1] = az;
    rk[2 * wid + 1] = kz;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NT / 64; w++) {
      if (rv[2 * w] < ax || (rv[2 * w] == ax && rk[2 * w] < kx)) {
        ax = rv[2 * w];
        kx = rk[2 * w];
      }
      if (rv[2 * w + 1] < az || (rv[2 * w + 1] == az && rk[2 * w + 1] < kz)) {
        az = rv[2 * w + 1];
        kz = rk[2 * w + 1];
      }
    }
    ws[V.slot * 4 + 0] = ax;
    ws[V.slot * 4 + 1] = (double)kx;
    ws[V.slot * 4 + 2] = az;
    ws[V.slot * 4 + 3] = (double)kz;
  }
}

// The step bounds of each problem from its partial rows: the smallest
// (value, index) pair in lexicographic order, which no reduction order can
// change.  One 256-thread workgroup per problem (block k for problem k).
__device__ __forceinline__ void step_min(double& a, double& k, double b, double l) {
  if (b < a || (b == a && l < k)) {
    a = b;
    k = l;
  }
}

__global__ __launch_bounds__(256) void max_step_finish(BT B, int K, amigo_ipm_problem P,
                                                       const double* ws, int nb, double* out) {
  __shared__ double sv[4][256];
  const int t = threadIdx.x;
  if (B.prob) {
    const int k = B.problems ? B.problems[blockIdx.x] : (int)blockIdx.x;
    if (k >= K || (B.active && B.active[k] == 0)) return;
    nb = B.count[k];
    ws += (size_t)B.first[k] * 4;
    P.primal_indices += B.pseg[k];
    out += (size_t)k * 4;
  }
  double ax = 1.0, az = 1.0;
  double kx = 9.0e18, kz = 9.0e18;
  for (int b = t; b < nb; b += 256) {
    step_min(ax, kx, ws[4 * b], ws[4 * b + 1]);
    step_min(az, kz, ws[4 * b + 2], ws[4 * b + 3]);
  }
  sv[0][t] = ax;
  sv[1][t] = kx;
  sv[2][t] = az;
  sv[3][t] = kz;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      double a = sv[0][t], ka = sv[1][t], c = sv[2][t], kc = sv[3][t];
      step_min(a, ka, sv[0][t + w], sv[1][t + w]);
      step_min(c, kc, sv[2][t + w], sv[3][t + w]);
      sv[0][t] = a;
      sv[1][t] = ka;
      sv[2][t] = c;
      sv[3][t] = kc;
    }
    __syncthreads();
  }
  if (t == 0) {
    ax = sv[0][0];
    kx = sv[1][0];
    az = sv[2][0];
    kz = sv[3][0];
    out[0] = ax;
    out[1] = kx < 9.0e18 ? (double)P.primal_indices[(long long)kx / 2] : -1.0;
    out[2] = az;
    out[3] = kz < 9.0e18 ? (double)P.primal_indices[(long long)kz / 2] : -1.0;
  }
}

// Trial step (interior_point_backend.h:375-392).
__global__ void apply_step_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S, double ax,
                                  double az, const double* __restrict__ axb,
                                  const double* __restrict__ azb,
                                  const double* __restrict__ dxlam,
                                  const double* __restrict__ dzl,
                                  const double* __restrict__ dzu, int n,
                                  amigo_ipm_state T) {
  const View V = view(B, P, S);
  if (V.skip) return;
  ax = pick(ax, axb, V);
  az = pick(az, azb, V);
  if (B.prob) n = V.n;
  dxlam += V.xo;
  dzl += V.p0;
  dzu += V.p0;
  shift_state(T, V);
  for (int t = V.bx * NT + threadIdx.x; t < n; t += V.gx * NT) {
    T.xlam[t] = S.xlam[t] + ax * dxlam[t];
    if (t < P.n_primal) {
      const int i = t;
      const double dx = ax * dxlam[P.primal_indices[i]];
      if (finite_bound(P.lbx[i])) {
        T.zl[i] = S.zl[i] + az * dzl[i];
        T.sl[i] = S.sl[i] + dx;
      } else {
        T.zl[i] = S.zl[i];
        T.sl[i] = S.sl[i];
      }
      if (finite_bound(P.ubx[i])) {
        T.zu[i] = S.zu[i] + az * dzu[i];
        T.su[i] = S.su[i] - dx;
      } else {
        T.zu[i] = S.zu[i];
        T.su[i] = S.su[i];
      }
    }
  }
}

// Generic scalar reductions over the primals / constraints.
enum RedKind {
  RK_COMPL = 0,        // [sum gap*z, count, min gap*z]  (backend.h:396-414)
  RK_KKT_ERR = 1,      // [max dual, max primal, max comp] (backend.h:470-489)
  RK_LOGSUM = 2,       // [barrier log-sum]                 (backend.h:494-508)
  RK_DPHI_KKT = 3,     // [dphi]                            (backend.h:592-606)
  RK_THETA = 4,        // [sum |c - lbh|]                   (backend.h:553-562)
  RK_KKT_SQ = 5,       // [dual_sq, primal_sq, comp_sq]     (backend.h:530-549)
  RK_MAX_COMP_DEV = 6, // [max |gap*z - mu|]                (backend.h:417-427)
  RK_COMPL_SQ = 7,     // [sum (gap*z - mu)^2]              (backend.h:430-444)
  RK_DPHI_DIRECT = 8,  // [dphi]                            (backend.h:513-524)
  RK_ABS_SUMS = 9,     // [sum|zl|+|zu|, sum|y|, n_bounds, n_constr]
};

__global__ void reduce_kernel(BT B, int kind, amigo_ipm_problem P, amigo_ipm_state S,
                              double mu, const double* __restrict__ mub,
                              const double* __restrict__ a,
                              const double* __restrict__ b,
                              const double* __restrict__ c, double* ws) {
  const View V = view(B, P, S);
  if (V.skip) return;
  mu = pick(mu, mub, V);
  if (a) a += V.xo;
  if (b) b += V.xo;
  if (c) c += V.xo;
  __shared__ double red[(NT / 64) * 4];
  int ops[4] = {SUM, SUM, SUM, SUM};
  if (kind == RK_COMPL) ops[2] = MIN;
  if (kind == RK_KKT_ERR) ops[0] = ops[1] = ops[2] = MAX;
  if (kind == RK_MAX_COMP_DEV) ops[0] = MAX;
  double v[4];
  for (int q = 0; q < 4; q++) v[q] = ident(ops[q]);
  if (kind == RK_KKT_ERR || kind == RK_MAX_COMP_DEV)
    for (int q = 0; q < 4; q++) v[q] = 0.0;  // the reference starts its maxima at 0
  const int total = P.n_primal + P.n_constraints;
  for (int t = V.bx * NT + threadIdx.x; t < total; t += V.gx * NT) {
    const bool prim = t < P.n_primal;
    const int i = t, j = t - P.n_primal;
    const bool hl = prim && finite_bound(P.lbx[i]);
    const bool hu = prim && finite_bound(P.ubx[i]);
    switch (kind) {
      case RK_COMPL:
        if (hl) {
          const double cc = S.sl[i] * S.zl[i];
          v[0] += cc;
          v[1] += 1.0;
          v[2] = fmin(v[2], cc);
        }
        if (hu) {
          const double cc = S.su[i] * S.zu[i];
          v[0] += cc;
          v[1] += 1.0;
          v[2] = fmin(v[2], cc);
        }
        break;
      case RK_KKT_ERR:
        if (prim) {
          const int idx = P.primal_indices[i];
          v[0] = fmax(v[0], fabs(a[idx] - S.zl[i] + S.zu[i]));
          if (hl) v[2] = fmax(v[2], fabs(S.sl[i] * S.zl[i] - mu));
          if (hu) v[2] = fmax(v[2], fabs(S.su[i] * S.zu[i] - mu));
        } else {
          v[1] = fmax(v[1], fabs(a[P.constraint_indices[j]] - P.lbh[j]));
        }
        break;
      case RK_LOGSUM:
        if (hl && S.sl[i] > 0) v[0] -= mu * log(S.sl[i]);
        if (hu && S.su[i] > 0) v[0] -= mu * log(S.su[i]);
        break;
      case RK_DPHI_KKT:  // a = res, b = px, c = diag
        if (prim) {
          const int idx = P.primal_indices[i];
          v[0] -= a[idx] * b[idx];
        } else {
          const int idx = P.constraint_indices[j];
          const double jdx = a[idx] - c[idx] * b[idx];
          v[0] -= S.xlam[idx] * jdx;
        }
        break;
      case RK_THETA:
        if (!prim) v[0] += fabs(a[P.constraint_indices[j]] - P.lbh[j]);
        break;
      case RK_KKT_SQ:
        if (prim) {
          const int idx = P.primal_indices[i];
          const double rd = a[idx] - S.zl[i] + S.zu[i];
          v[0] += rd * rd;
          if (hl) {
            const double cc = S.sl[i] * S.zl[i];
            v[2] += cc * cc;
          }
          if (hu) {
            const double cc = S.su[i] * S.zu[i];
            v[2] += cc * cc;
          }
        } else {
          const double rp = a[P.constraint_indices[j]] - P.lbh[j];
          v[1] += rp * rp;
        }
        break;
      case RK_MAX_COMP_DEV:
        if (hl) v[0] = fmax(v[0], fabs(S.sl[i] * S.zl[i] - mu));
        if (hu) v[0] = fmax(v[0], fabs(S.su[i] * S.zu[i] - mu));
        break;
      case RK_COMPL_SQ:
        if (hl) {
          const double r = S.sl[i] * S.zl[i] - mu;
          v[0] += r * r;
        }
        if (hu) {
          const double r = S.su[i] * S.zu[i] - mu;
          v[0] += r * r;
        }
        break;
      case RK_DPHI_DIRECT:  // a = grad, b = px
        if (prim) {
          const int idx = P.primal_indices[i];
          const double dx = b[idx];
          v[0] += a[idx] * dx;
          if (hl) v[0] -= mu * dx / S.sl[i];
          if (hu) v[0] += mu * dx / S.su[i];
        }
        break;
      case RK_ABS_SUMS:
        if (prim) {
          v[0] += fabs(S.zl[i]) + fabs(S.zu[i]);
          v[2] += (hl ? 1.0 : 0.0) + (hu ? 1.0 : 0.0);
        } else {
          v[1] += fabs(S.xlam[P.constraint_indices[j]]);
          v[3] += 1.0;
        }
        break;
    }
  }
  block_reduce<4>(v, ops, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; q++) ws[V.slot * 4 + q] = v[q];
}

// z clamp into [mu/(kappa gap), kappa mu/gap] (backend.h:459-476).
__global__ void reset_multipliers_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S,
                                         double mu, const double* __restrict__ mub,
                                         double kappa) {
  const View V = view(B, P, S);
  if (V.skip) return;
  mu = pick(mu, mub, V);
  for (int i = V.bx * NT + threadIdx.x; i < P.n_primal; i += V.gx * NT) {
    if (finite_bound(P.lbx[i])) {
      const double g = S.sl[i];
      S.zl[i] = fmax(fmin(S.zl[i], kappa * mu / g), mu / (kappa * g));
    }
    if (finite_bound(P.ubx[i])) {
      const double g = S.su[i];
      S.zu[i] = fmax(fmin(S.zu[i], kappa * mu / g), mu / (kappa * g));
    }
  }
}

// z + dz floored at beta_min where the bound is finite, written to the
// output duals (backend.h:480-489); other entries are left as they are.
__global__ void affine_start_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S,
                                    double beta_min, const double* __restrict__ dzl,
                                    const double* __restrict__ dzu, double* __restrict__ zl_out,
                                    double* __restrict__ zu_out) {
  const View V = view(B, P, S);
  if (V.skip) return;
  dzl += V.p0;
  dzu += V.p0;
  zl_out += V.p0;
  zu_out += V.p0;
  for (int i = V.bx * NT + threadIdx.x; i < P.n_primal; i += V.gx * NT) {
    if (finite_bound(P.lbx[i])) zl_out[i] = fmax(S.zl[i] + dzl[i], beta_min);
    if (finite_bound(P.ubx[i])) zu_out[i] = fmax(S.zu[i] + dzu[i], beta_min);
  }
}

// r_d = grad - zl + zu on primals, 0 elsewhere (backend.h:566-574).
__global__ void dual_residual_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S,
                                     const double* __restrict__ grad,
                                     double* __restrict__ out, int n) {
  const View V = view(B, P, S);
  if (V.skip) return;
  if (B.prob) n = V.n;
  out += V.xo;
  for (int t = V.bx * NT + threadIdx.x; t < n; t += V.gx * NT) out[t] = 0.0;
}
__global__ void dual_residual_fill(BT B, amigo_ipm_problem P, amigo_ipm_state S,
                                   const double* __restrict__ grad,
                                   double* __restrict__ out) {
  const View V = view(B, P, S);
  if (V.skip) return;
  grad += V.xo;
  out += V.xo;
  for (int i = V.bx * NT + threadIdx.x; i < P.n_primal; i += V.gx * NT) {
    const int idx = P.primal_indices[i];
    out[idx] = grad[idx] - S.zl[i] + S.zu[i];
  }
}

// Starting point (backend.h:79-200): relaxed bounds, projection of the
// primals into the interior, unit bound duals, slacks from the primals.
__global__ void init_kernel(BT B, const double* __restrict__ lbx0,
                            const double* __restrict__ ubx0, int np,
                            const int* __restrict__ pidx, double relax,
                            double cvt, double kappa1, double kappa2,
                            double* __restrict__ lbx, double* __restrict__ ubx,
                            amigo_ipm_state S, int do_relax, int do_project) {
  amigo_ipm_problem P{np, 0, pidx, nullptr, lbx0, ubx0, nullptr};
  const View V = view(B, P, S);
  if (V.skip) return;
  np = P.n_primal;
  lbx0 += V.p0;
  ubx0 += V.p0;
  pidx += V.p0;
  lbx += V.p0;
  ubx += V.p0;
  for (int i = V.bx * NT + threadIdx.x; i < np; i += V.gx * NT) {
    double lb = lbx0[i], ub = ubx0[i];
    if (do_relax && relax > 0) {
      if (finite_bound(lb)) lb = lb - fmin(cvt, relax * fmax(1.0, fabs(lb)));
      if (finite_bound(ub)) ub = ub + fmin(cvt, relax * fmax(1.0, fabs(ub)));
      lbx[i] = lb;
      ubx[i] = ub;
    } else {
      lb = lbx[i];
      ub = ubx[i];
    }
    if (do_project) {
      const int idx = pidx[i];
      double x = S.xlam[idx];
      const bool hl = finite_bound(lb), hu = finite_bound(ub);
      if (hl && hu) {
        const double range = ub - lb;
        const double pl = fmin(kappa1 * fmax(1.0, fabs(lb)), kappa2 * range);
        const double pu = fmin(kappa1 * fmax(1.0, fabs(ub)), kappa2 * range);
        x = fmax(fmin(x, ub - pu), lb + pl);
      } else if (hl) {
        x = fmax(x, lb + kappa1 * fmax(1.0, fabs(lb)));
      } else if (hu) {
        x = fmin(x, ub - kappa1 * fmax(1.0, fabs(ub)));
      }
      S.xlam[idx] = x;
      S.zl[i] = hl ? 1.0 : 0.0;
      S.zu[i] = hu ? 1.0 : 0.0;
      S.sl[i] = hl ? (x - lb) : 1.0;
      S.su[i] = hu ? (ub - x) : 1.0;
    }
  }
}

// Slack flooring (amigo/algorithm/bound_safeguards.py:16-87): where a stored
// slack fell below s_min, reset it to min(max(mu/z, s_min), cap) and move the
// primal to stay consistent.  ws[0] receives the count (per block).
__global__ void floor_slacks_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S,
                                    const double* __restrict__ lbx0,
                                    const double* __restrict__ ubx0, double mu,
                                    const double* __restrict__ mub, double s_min,
                                    const double* __restrict__ sminb, double slack_move,
                                    double* ws) {
  const View V = view(B, P, S);
  if (V.skip) return;
  mu = pick(mu, mub, V);
  s_min = pick(s_min, sminb, V);
  lbx0 += V.p0;
  ubx0 += V.p0;
  __shared__ double red[(NT / 64) * 1];
  double v[1] = {0.0};
  const int ops[1] = {SUM};
  for (int i = V.bx * NT + threadIdx.x; i < P.n_primal; i += V.gx * NT) {
    const int idx = P.primal_indices[i];
    if (finite_bound(lbx0[i]) && S.sl[i] < s_min) {
      const double g = fmax(S.sl[i], 0.0);
      const double z = fmax(S.zl[i], DBL_MIN);
      const double target = fmax(mu / z, s_min);
      const double cap = slack_move * fmax(1.0, fabs(lbx0[i])) + g;
      S.sl[i] = fmin(target, cap);
      S.xlam[idx] = lbx0[i] + S.sl[i];
      v[0] += 1.0;
    }
    if (finite_bound(ubx0[i]) && S.su[i] < s_min) {
      const double g = fmax(S.su[i], 0.0);
      const double z = fmax(S.zu[i], DBL_MIN);
      const double target = fmax(mu / z, s_min);
      const double cap = slack_move * fmax(1.0, fabs(ubx0[i])) + g;
      S.su[i] = fmin(target, cap);
      S.xlam[idx] = ubx0[i] - S.su[i];
      v[0] += 1.0;
    }
  }
  block_reduce<1>(v, ops, red);
  if (threadIdx.x == 0) ws[V.slot] = v[0];
}

// After slack flooring moved any slack of a problem, every bounded primal of
// that problem is set from its slack: x = lb + sl, then x = ub - su where
// the upper bound is finite (bound_safeguards.py:68-80).  count[k] is the
// problem's number of adjusted slacks (device).
__global__ void snap_primals_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S,
                                    const double* __restrict__ lbx0,
                                    const double* __restrict__ ubx0,
                                    const double* __restrict__ count) {
  const View V = view(B, P, S);
  if (V.skip || !(count[V.k] > 0.0)) return;
  lbx0 += V.p0;
  ubx0 += V.p0;
  for (int i = V.bx * NT + threadIdx.x; i < P.n_primal; i += V.gx * NT) {
    const int idx = P.primal_indices[i];
    if (finite_bound(lbx0[i])) S.xlam[idx] = lbx0[i] + S.sl[i];
    if (finite_bound(ubx0[i])) S.xlam[idx] = ubx0[i] - S.su[i];
  }
}

// ---- sparse matrix helpers -------------------------------------------------
// y = A x for the full symmetric CSR (one wave per row).
// With row_grp, the rows whose group is not active (grp_active[row_grp[r]]
// == 0: problems of a batch left out) are skipped and keep their y.
__global__ void spmv_kernel(int n, const int* __restrict__ rowp,
                            const int* __restrict__ cols,
                            const double* __restrict__ vals,
                            const double* __restrict__ x, double* __restrict__ y,
                            const int* __restrict__ row_grp,
                            const int* __restrict__ grp_active) {
  const int lane = threadIdx.x & 63;
  const int row = (blockIdx.x * NT + threadIdx.x) >> 6;
  const int nw = (gridDim.x * NT) >> 6;
  for (int r = row; r < n; r += nw) {
    if (row_grp && !grp_active[row_grp[r]]) continue;
    double s = 0.0;
    for (int p = rowp[r] + lane; p < rowp[r + 1]; p += 64) s += vals[p] * x[cols[p]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[r] = s;
  }
}

// vals[p] *= sv[row] * sv[col] (interior_point_optimizer.h:573-584).
__global__ void scale_csr_kernel(int n, const int* __restrict__ rowp,
                                 const int* __restrict__ cols,
                                 const double* __restrict__ sv, double* __restrict__ vals,
                                 const int* __restrict__ row_grp,
                                 const int* __restrict__ grp_active) {
  // One thread per row (KKT rows hold a few entries; each entry's product
  // is the same whichever thread forms it).
  for (int r = blockIdx.x * NT + threadIdx.x; r < n; r += gridDim.x * NT) {
    if (row_grp && !grp_active[row_grp[r]]) continue;
    const double d = sv[r];
    for (int p = rowp[r]; p < rowp[r + 1]; p++) vals[p] *= d * sv[cols[p]];
  }
}

// Per-constraint-row inf-norm -> dc = min(1, max_gradient/|row|_inf), floored
// at min_value (interior_point_optimizer.h:524-540).
__global__ void constraint_scale_kernel(int nc, const int* __restrict__ cidx,
                                        const int* __restrict__ rowp,
                                        const double* __restrict__ vals,
                                        double max_gradient, double min_value,
                                        double* __restrict__ dc) {
  const int lane = threadIdx.x & 63;
  const int w = (blockIdx.x * NT + threadIdx.x) >> 6;
  const int nw = (gridDim.x * NT) >> 6;
  for (int j = w; j < nc; j += nw) {
    const int r = cidx[j];
    double m = 0.0;
    for (int p = rowp[r] + lane; p < rowp[r + 1]; p += 64) m = fmax(m, fabs(vals[p]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (lane == 0) {
      double d = m > max_gradient ? max_gradient / m : 1.0;
      dc[j] = d < min_value ? min_value : d;
    }
  }
}

// Iterative refinement on the full 8-block system (linear_solver.py:109-264):
// residual blocks, their inf-norms, and the condensed correction RHS.
// ws partial layout: [max|e|, max|(px,dzl,dzu)|].
__global__ void ir_residual_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S,
                                   double mu, const double* __restrict__ mub,
                                   const double* __restrict__ grad,
                                   const double* __restrict__ rhs_saved,
                                   const double* __restrict__ Kpx,
                                   const double* __restrict__ px,
                                   const double* __restrict__ dzl,
                                   const double* __restrict__ dzu,
                                   double* __restrict__ econd,
                                   double* __restrict__ ezl,
                                   double* __restrict__ ezu, double* ws) {
  const View V = view(B, P, S);
  if (V.skip) return;
  mu = pick(mu, mub, V);
  grad += V.xo;
  rhs_saved += V.xo;
  Kpx += V.xo;
  px += V.xo;
  econd += V.xo;
  dzl += V.p0;
  dzu += V.p0;
  ezl += V.p0;
  ezu += V.p0;
  __shared__ double red[(NT / 64) * 2];
  double v[2] = {0.0, 0.0};
  const int ops[2] = {MAX, MAX};
  const int total = P.n_primal + P.n_constraints;
  for (int t = V.bx * NT + threadIdx.x; t < total; t += V.gx * NT) {
    if (t < P.n_primal) {
      const int i = t, idx = P.primal_indices[i];
      const bool fl = finite_bound(P.lbx[i]), fu = finite_bound(P.ubx[i]);
      const double gl = fl ? S.sl[i] : 1.0, gu = fu ? S.su[i] : 1.0;
      const double zl = S.zl[i], zu = S.zu[i];
      const double sigma = (fl ? zl / gl : 0.0) + (fu ? zu / gu : 0.0);
      const double dx = px[idx];
      const double rhs_x = -(grad[idx] - zl + zu);
      const double e_x = rhs_x - Kpx[idx] + sigma * dx + dzl[i] - dzu[i];
      const double e_zl = fl ? (mu - gl * zl) - zl * dx - gl * dzl[i] : 0.0;
      const double e_zu = fu ? (mu - gu * zu) + zu * dx - gu * dzu[i] : 0.0;
      v[0] = fmax(v[0], fmax(fabs(e_x), fmax(fabs(e_zl), fabs(e_zu))));
      v[1] = fmax(v[1], fmax(fabs(dx), fmax(fabs(dzl[i]), fabs(dzu[i]))));
      econd[idx] = e_x + (fl ? e_zl / gl : 0.0) - (fu ? e_zu / gu : 0.0);
      ezl[i] = e_zl;
      ezu[i] = e_zu;
    } else {
      const int idx = P.constraint_indices[t - P.n_primal];
      const double e_c = rhs_saved[idx] - Kpx[idx];
      v[0] = fmax(v[0], fabs(e_c));
      v[1] = fmax(v[1], fabs(px[idx]));
      econd[idx] = e_c;
    }
  }
  block_reduce<2>(v, ops, red);
  if (threadIdx.x == 0) {
    ws[V.slot * 2 + 0] = v[0];
    ws[V.slot * 2 + 1] = v[1];
  }
}

// Norm of the 8-block RHS: max over |rhs_x|, |rhs_c|, |rhs_zl|, |rhs_zu|.
__global__ void ir_rhs_norm_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S, double mu,
                                   const double* __restrict__ mub,
                                   const double* __restrict__ grad,
                                   const double* __restrict__ rhs_saved, double* ws) {
  const View V = view(B, P, S);
  if (V.skip) return;
  mu = pick(mu, mub, V);
  grad += V.xo;
  rhs_saved += V.xo;
  __shared__ double red[(NT / 64) * 1];
  double v[1] = {0.0};
  const int ops[1] = {MAX};
  const int total = P.n_primal + P.n_constraints;
  for (int t = V.bx * NT + threadIdx.x; t < total; t += V.gx * NT) {
    if (t < P.n_primal) {
      const int i = t, idx = P.primal_indices[i];
      const bool fl = finite_bound(P.lbx[i]), fu = finite_bound(P.ubx[i]);
      const double gl = fl ? S.sl[i] : 1.0, gu = fu ? S.su[i] : 1.0;
      v[0] = fmax(v[0], fabs(grad[idx] - S.zl[i] + S.zu[i]));
      if (fl) v[0] = fmax(v[0], fabs(mu - gl * S.zl[i]));
      if (fu) v[0] = fmax(v[0], fabs(mu - gu * S.zu[i]));
    } else {
      v[0] = fmax(v[0], fabs(rhs_saved[P.constraint_indices[t - P.n_primal]]));
    }
  }
  block_reduce<1>(v, ops, red);
  if (threadIdx.x == 0) ws[V.slot] = v[0];
}

// dzl += (e_zl - zl dc)/gl, dzu += (e_zu + zu dc)/gu, px += corr.
__global__ void ir_update_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S, int n,
                                 const double* __restrict__ corr,
                                 const double* __restrict__ ezl,
                                 const double* __restrict__ ezu,
                                 double* __restrict__ px, double* __restrict__ dzl,
                                 double* __restrict__ dzu) {
  const View V = view(B, P, S);
  if (V.skip) return;
  if (B.prob) n = V.n;
  corr += V.xo;
  px += V.xo;
  ezl += V.p0;
  ezu += V.p0;
  dzl += V.p0;
  dzu += V.p0;
  for (int t = V.bx * NT + threadIdx.x; t < n; t += V.gx * NT) {
    if (t < P.n_primal) {
      const int i = t;
      const double dc = corr[P.primal_indices[i]];
      if (finite_bound(P.lbx[i])) dzl[i] += (ezl[i] - S.zl[i] * dc) / S.sl[i];
      if (finite_bound(P.ubx[i])) dzu[i] += (ezu[i] + S.zu[i] * dc) / S.su[i];
    }
    px[t] += corr[t];
  }
}

// Initial bound-dual steps from px (linear_solver.py:177-179).
__global__ void ir_init_dz_kernel(BT B, amigo_ipm_problem P, amigo_ipm_state S, double mu,
                                  const double* __restrict__ mub,
                                  const double* __restrict__ px,
                                  double* __restrict__ dzl, double* __restrict__ dzu) {
  const View V = view(B, P, S);
  if (V.skip) return;
  mu = pick(mu, mub, V);
  px += V.xo;
  dzl += V.p0;
  dzu += V.p0;
  for (int i = V.bx * NT + threadIdx.x; i < P.n_primal; i += V.gx * NT) {
    const double dx = px[P.primal_indices[i]];
    const bool fl = finite_bound(P.lbx[i]), fu = finite_bound(P.ubx[i]);
    const double gl = fl ? S.sl[i] : 1.0, gu = fu ? S.su[i] : 1.0;
    dzl[i] = fl ? (mu - gl * S.zl[i] - S.zl[i] * dx) / gl : 0.0;
    dzu[i] = fu ? (mu - gu * S.zu[i] + S.zu[i] * dx) / gu : 0.0;
  }
}

__global__ void scatter_scale_kernel(BT B, int nc, const int* __restrict__ cidx,
                                     const double* __restrict__ dc,
                                     double* __restrict__ v, int mode) {
  amigo_ipm_problem P{0, nc, nullptr, cidx, nullptr, nullptr, dc};
  amigo_ipm_state S0{nullptr, nullptr, nullptr, nullptr, nullptr};
  const View V = view(B, P, S0);
  if (V.skip) return;
  nc = P.n_constraints;
  cidx += V.c0;
  dc += V.c0;
  v += mode == 2 ? V.c0 : V.xo;
  for (int j = V.bx * NT + threadIdx.x; j < nc; j += V.gx * NT) {
    if (mode == 0)
      v[cidx[j]] *= dc[j];
    else if (mode == 1)
      v[cidx[j]] /= dc[j];
    else
      v[j] *= dc[j];
  }
}

// Copy of the active problems' segments: what = 1 -> x only, 2 -> zl, zu,
// sl, su only, 3 -> all five vectors (OptVector::copy of the problems that
// take part).  Null destination members are skipped.
__global__ void copy_state_kernel(BT B, amigo_ipm_state D, amigo_ipm_state S, int what) {
  // A member is copied only when both states have it: the flags are taken
  // before the views shift the pointers by the problem's offsets.
  const bool cx = (what & 1) && D.xlam && S.xlam;
  const bool czl = (what & 2) && D.zl && S.zl, czu = (what & 2) && D.zu && S.zu;
  const bool csl = (what & 2) && D.sl && S.sl, csu = (what & 2) && D.su && S.su;
  amigo_ipm_problem P{0, 0, nullptr, nullptr, nullptr, nullptr, nullptr};
  const View V = view(B, P, S);
  if (V.skip) return;
  shift_state(D, V);
  const int np = B.pseg[V.k + 1] - V.p0;
  for (int t = V.bx * NT + threadIdx.x; t < V.n; t += V.gx * NT) {
    if (cx) D.xlam[t] = S.xlam[t];
    if (t < np) {
      if (czl) D.zl[t] = S.zl[t];
      if (czu) D.zu[t] = S.zu[t];
      if (csl) D.sl[t] = S.sl[t];
      if (csu) D.su[t] = S.su[t];
    }
  }
}

int grid_for(int n) { return nblocks(n); }

#define LAUNCH_CHECK(name)                                  \
  do {                                                      \
    hipError_t e__ = hipGetLastError();                     \
    if (e__ != hipSuccess) return set_err(name, e__);       \
  } while (0)

}  // namespace

// Largest |x_i| of each problem's segment [seg[k], seg[k+1]) -- the
// divergence test of convergence_check.py:82-83.  One workgroup per problem;
// a maximum is exact, so the result does not depend on the reduction order,
// and a NaN propagates as in np.max.
__device__ __forceinline__ double nan_max(double a, double b) {
  return (b > a || isnan(b)) && !isnan(a) ? b : a;
}

// |x| maximum of each segment (NaN if any entry is NaN): blockIdx.x is the
// segment, blockIdx.y strides over it, and the blocks combine with an
// atomic maximum of the bit patterns (non-negative doubles order like their
// bits; a NaN's bits exceed every number's), so the result does not depend
// on the order.  out starts at zero (segment_zero_kernel).
__global__ void segment_zero_kernel(int K, unsigned long long* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < K) out[k] = 0ull;
}

__global__ __launch_bounds__(NT) void segment_absmax_kernel(const double* __restrict__ x,
                                                            const int* __restrict__ seg,
                                                            unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[NT / 64];
  const int k = blockIdx.x;
  unsigned long long m = 0ull;
  for (int i = seg[k] + blockIdx.y * NT + threadIdx.x; i < seg[k + 1]; i += gridDim.y * NT) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(x[i]));
    m = b > m ? b : m;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long r = red[0];
    for (int w = 1; w < NT / 64; w++) r = red[w] > r ? red[w] : r;
    if (r) atomicMax(out + k, r);
  }
}

// ---- scalar reductions without allocation -----------------------------------
// The workspace of amigo_ipm_reduce_noalloc: the per-block partial rows of one
// reduction (at most MAXB blocks x 4 values) followed by the 4 combined
// results, as in a caller's ws.  One copy per GPU, held until the process
// ends; every call on that GPU reuses it.
constexpr int kScratchDoubles = MAXB * 4 + 16;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current.  In host code the array's name is only a
// stand-in for the symbol; kernels get this address.
int scratch_on_current_device(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_err("hipGetDevice", e);
  if (dev < 0 || dev >= kMaxDevices) {
    g_ipm_err = "amigo_ipm_reduce_noalloc: device index " + std::to_string(dev) +
                " beyond the " + std::to_string(kMaxDevices) + " supported";
    return -1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return set_err("hipGetSymbolAddress(scratch)", e);
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" {

const char* amigo_ipm_last_error(void) { return g_ipm_err.c_str(); }

int amigo_ipm_workspace_doubles(void) { return MAXB * 4 + 16; }

int amigo_ipm_grid_blocks(int n) { return nblocks(n); }

int amigo_ipm_residual(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                       double mu, const double* grad, double* res, double* ws,
                       double* out3, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const int total = P->n_primal + P->n_constraints;
  const int nb = nblocks(total);
  hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(NT), 0, s, BT{}, *P, *S, mu, nullptr,
                     grad, res, ws);
  LAUNCH_CHECK("residual_kernel");
  const int ops[3] = {SUM, SUM, SUM};
  return finish_and_copy(3, ops, ws, nb, out3, s);
}

int amigo_ipm_diagonal(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                       double* diag, void* stream) {
  const int total = P->n_primal + P->n_constraints;
  hipLaunchKernelGGL(diagonal_kernel, dim3(grid_for(total)), dim3(NT), 0, (hipStream_t)stream,
                     BT{}, *P, *S, diag);
  LAUNCH_CHECK("diagonal_kernel");
  return 0;
}

int amigo_ipm_bound_dual_step(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                              double mu, const double* px, double* dzl, double* dzu,
                              void* stream) {
  hipLaunchKernelGGL(bound_dual_step_kernel, dim3(grid_for(P->n_primal)), dim3(NT), 0,
                     (hipStream_t)stream, BT{}, *P, *S, mu, nullptr, px, dzl, dzu);
  LAUNCH_CHECK("bound_dual_step_kernel");
  return 0;
}

int amigo_ipm_max_step(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                       double tau, const double* px, const double* dzl,
                       const double* dzu, double* ws, double* out4, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const int nb = nblocks(P->n_primal);
  hipLaunchKernelGGL(max_step_kernel, dim3(nb), dim3(NT), 0, s, BT{}, *P, *S, tau, nullptr,
                     px, dzl, dzu, ws);
  double* out = ws + (size_t)MAXB * 4;
  hipLaunchKernelGGL(max_step_finish, dim3(1), dim3(256), 0, s, BT{}, 1, *P, ws, nb, out);
  LAUNCH_CHECK("max_step_kernel");
  hipError_t e = hipMemcpyAsync(out4, out, 4 * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return set_err("hipMemcpyAsync", e);
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return set_err("hipStreamSynchronize", e);
  return 0;
}

int amigo_ipm_apply_step(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                         double ax, double az, const double* dxlam, const double* dzl,
                         const double* dzu, int n, const amigo_ipm_state* T,
                         void* stream) {
  hipLaunchKernelGGL(apply_step_kernel, dim3(grid_for(n)), dim3(NT), 0, (hipStream_t)stream,
                     BT{}, *P, *S, ax, az, nullptr, nullptr, dxlam, dzl, dzu, n, *T);
  LAUNCH_CHECK("apply_step_kernel");
  return 0;
}

static void reduce_ops(int kind, int* ops) {
  ops[0] = ops[1] = ops[2] = ops[3] = SUM;
  if (kind == RK_COMPL) ops[2] = MIN;
  if (kind == RK_KKT_ERR) ops[0] = ops[1] = ops[2] = MAX;
  if (kind == RK_MAX_COMP_DEV) ops[0] = MAX;
}

int amigo_ipm_reduce(int kind, const amigo_ipm_problem* P, const amigo_ipm_state* S,
                     double mu, const double* a, const double* b, const double* c,
                     double* ws, double* out4, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const int total = P->n_primal + P->n_constraints;
  const int nb = nblocks(total);
  hipLaunchKernelGGL(reduce_kernel, dim3(nb), dim3(NT), 0, s, BT{}, kind, *P, *S, mu, nullptr,
                     a, b, c, ws);
  LAUNCH_CHECK("reduce_kernel");
  int ops[4];
  reduce_ops(kind, ops);
  return finish_and_copy(4, ops, ws, nb, out4, s);
}

int amigo_ipm_reduce_noalloc(int kind, const amigo_ipm_problem* P, const amigo_ipm_state* S,
                             double mu, const double* a, const double* b, const double* c,
                             double* out4, void* stream) {
  if (!P || !S || !out4) {
    g_ipm_err = "amigo_ipm_reduce_noalloc: null problem, state or output";
    return -1;
  }
  if (kind < RK_COMPL || kind > RK_ABS_SUMS) {
    g_ipm_err = "amigo_ipm_reduce_noalloc: unknown reduction kind " + std::to_string(kind);
    return -1;
  }
  const long long total = (long long)P->n_primal + P->n_constraints;
  if (P->n_primal < 0 || P->n_constraints < 0 || total > INT_MAX) {
    g_ipm_err = "amigo_ipm_reduce_noalloc: n_primal + n_constraints must lie in [0, 2^31 - 1]";
    return -1;
  }
  const int nb = nblocks((int)total);
  if ((long long)nb * 4 + 4 > kScratchDoubles) {
    g_ipm_err = "amigo_ipm_reduce_noalloc: " + std::to_string(nb) +
                " partial blocks exceed the static workspace";
    return -1;
  }
  double* ws = nullptr;
  const int rc = scratch_on_current_device(&ws);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(reduce_kernel, dim3(nb), dim3(NT), 0, s, BT{}, kind, *P, *S, mu, nullptr,
                     a, b, c, ws);
  LAUNCH_CHECK("reduce_kernel");
  int ops[4];
  reduce_ops(kind, ops);
  return finish_and_copy(4, ops, ws, nb, out4, s);
}

int amigo_ipm_reset_bound_multipliers(const amigo_ipm_problem* P,
                                      const amigo_ipm_state* S, double mu,
                                      double kappa, void* stream) {
  hipLaunchKernelGGL(reset_multipliers_kernel, dim3(grid_for(P->n_primal)), dim3(NT), 0,
                     (hipStream_t)stream, BT{}, *P, *S, mu, nullptr, kappa);
  LAUNCH_CHECK("reset_multipliers_kernel");
  return 0;
}

int amigo_ipm_affine_start_point(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                                 double beta_min, const double* dzl, const double* dzu,
                                 double* zl_out, double* zu_out, void* stream) {
  hipLaunchKernelGGL(affine_start_kernel, dim3(grid_for(P->n_primal)), dim3(NT), 0,
                     (hipStream_t)stream, BT{}, *P, *S, beta_min, dzl, dzu, zl_out, zu_out);
  LAUNCH_CHECK("affine_start_kernel");
  return 0;
}

int amigo_ipm_dual_residual_vector(const amigo_ipm_problem* P,
                                   const amigo_ipm_state* S, const double* grad,
                                   double* out, int n, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(dual_residual_kernel, dim3(grid_for(n)), dim3(NT), 0, s, BT{}, *P, *S,
                     grad, out, n);
  hipLaunchKernelGGL(dual_residual_fill, dim3(grid_for(P->n_primal)), dim3(NT), 0, s, BT{}, *P,
                     *S, grad, out);
  LAUNCH_CHECK("dual_residual_kernel");
  return 0;
}

int amigo_ipm_initialize(const double* lbx0, const double* ubx0, int np,
                         const int* pidx, double relax, double cvt, double kappa1,
                         double kappa2, double* lbx, double* ubx,
                         const amigo_ipm_state* S, int do_relax, int do_project,
                         void* stream) {
  hipLaunchKernelGGL(init_kernel, dim3(grid_for(np)), dim3(NT), 0, (hipStream_t)stream, BT{},
                     lbx0, ubx0, np, pidx, relax, cvt, kappa1, kappa2, lbx, ubx, *S, do_relax,
                     do_project);
  LAUNCH_CHECK("init_kernel");
  return 0;
}

int amigo_ipm_floor_slacks(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                           const double* lbx0, const double* ubx0, double mu,
                           double s_min, double slack_move, double* ws,
                           double* out1, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const int nb = nblocks(P->n_primal);
  hipLaunchKernelGGL(floor_slacks_kernel, dim3(nb), dim3(NT), 0, s, BT{}, *P, *S, lbx0, ubx0,
                     mu, nullptr, s_min, nullptr, slack_move, ws);
  LAUNCH_CHECK("floor_slacks_kernel");
  const int ops[1] = {SUM};
  const int rc = finish_and_copy(1, ops, ws, nb, out1, s);
  if (rc) return rc;
  // finish_and_copy leaves the count at ws[MAXB * 4] on the device.
  hipLaunchKernelGGL(snap_primals_kernel, dim3(nb), dim3(NT), 0, s, BT{}, *P, *S, lbx0, ubx0,
                     ws + (size_t)MAXB * 4);
  LAUNCH_CHECK("snap_primals_kernel");
  return 0;
}

int amigo_csr_spmv(int n, const int* rowp, const int* cols, const double* vals,
                   const double* x, double* y, void* stream) {
  return amigo_csr_spmv_masked(n, rowp, cols, vals, x, y, nullptr, nullptr, stream);
}

int amigo_csr_spmv_masked(int n, const int* rowp, const int* cols, const double* vals,
                          const double* x, double* y, const int* row_grp,
                          const int* grp_active, void* stream) {
  if ((row_grp == nullptr) != (grp_active == nullptr)) return -1;
  const int waves = n;
  int nb = (waves * 64 + NT - 1) / NT;
  if (nb > 8192) nb = 8192;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(spmv_kernel, dim3(nb), dim3(NT), 0, (hipStream_t)stream, n, rowp, cols,
                     vals, x, y, row_grp, grp_active);
  LAUNCH_CHECK("spmv_kernel");
  return 0;
}

int amigo_csr_scale_symmetric(int n, const int* rowp, const int* cols,
                              const double* sv, double* vals, void* stream) {
  return amigo_csr_scale_symmetric_masked(n, rowp, cols, sv, vals, nullptr, nullptr, stream);
}

int amigo_csr_scale_symmetric_masked(int n, const int* rowp, const int* cols,
                                     const double* sv, double* vals, const int* row_grp,
                                     const int* grp_active, void* stream) {
  if ((row_grp == nullptr) != (grp_active == nullptr)) return -1;
  int nb = (n + NT - 1) / NT;
  if (nb > 65535) nb = 65535;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(scale_csr_kernel, dim3(nb), dim3(NT), 0, (hipStream_t)stream, n, rowp,
                     cols, sv, vals, row_grp, grp_active);
  LAUNCH_CHECK("scale_csr_kernel");
  return 0;
}

int amigo_ipm_constraint_scaling(int nc, const int* cidx, const int* rowp,
                                 const double* vals, double max_gradient,
                                 double min_value, double* dc, void* stream) {
  int nb = (nc * 64 + NT - 1) / NT;
  if (nb > 8192) nb = 8192;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(constraint_scale_kernel, dim3(nb), dim3(NT), 0, (hipStream_t)stream, nc,
                     cidx, rowp, vals, max_gradient, min_value, dc);
  LAUNCH_CHECK("constraint_scale_kernel");
  return 0;
}

int amigo_ipm_scale_rows(int nc, const int* cidx, const double* dc, double* v,
                         int mode, void* stream) {
  hipLaunchKernelGGL(scatter_scale_kernel, dim3(grid_for(nc)), dim3(NT), 0, (hipStream_t)stream,
                     BT{}, nc, cidx, dc, v, mode);
  LAUNCH_CHECK("scatter_scale_kernel");
  return 0;
}

int amigo_ipm_ir_init(const amigo_ipm_problem* P, const amigo_ipm_state* S, double mu,
                      const double* grad, const double* rhs_saved, const double* px,
                      double* dzl, double* dzu, double* ws, double* out1,
                      void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(ir_init_dz_kernel, dim3(grid_for(P->n_primal)), dim3(NT), 0, s, BT{}, *P,
                     *S, mu, nullptr, px, dzl, dzu);
  const int total = P->n_primal + P->n_constraints;
  const int nb = nblocks(total);
  hipLaunchKernelGGL(ir_rhs_norm_kernel, dim3(nb), dim3(NT), 0, s, BT{}, *P, *S, mu, nullptr,
                     grad, rhs_saved, ws);
  LAUNCH_CHECK("ir_init");
  const int ops[1] = {MAX};
  return finish_and_copy(1, ops, ws, nb, out1, s);
}

int amigo_ipm_ir_residual(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                          double mu, const double* grad, const double* rhs_saved,
                          const double* Kpx, const double* px, const double* dzl,
                          const double* dzu, double* econd, double* ezl, double* ezu,
                          double* ws, double* out2, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const int total = P->n_primal + P->n_constraints;
  const int nb = nblocks(total);
  hipLaunchKernelGGL(ir_residual_kernel, dim3(nb), dim3(NT), 0, s, BT{}, *P, *S, mu, nullptr,
                     grad, rhs_saved, Kpx, px, dzl, dzu, econd, ezl, ezu, ws);
  LAUNCH_CHECK("ir_residual_kernel");
  const int ops[2] = {MAX, MAX};
  return finish_and_copy(2, ops, ws, nb, out2, s);
}

int amigo_ipm_ir_update(const amigo_ipm_problem* P, const amigo_ipm_state* S, int n,
                        const double* corr, const double* ezl, const double* ezu,
                        double* px, double* dzl, double* dzu, void* stream) {
  hipLaunchKernelGGL(ir_update_kernel, dim3(grid_for(n)), dim3(NT), 0, (hipStream_t)stream,
                     BT{}, *P, *S, n, corr, ezl, ezu, px, dzl, dzu);
  LAUNCH_CHECK("ir_update_kernel");
  return 0;
}

// ---- batches -----------------------------------------------------------------

static BT bt_total(const amigo_ipm_batch* B) {
  return BT{B->blk_total, B->first_total, B->count_total, B->pseg,          B->cseg,
            B->xseg,      B->active,      B->launch_total, B->launch_problems};
}
static BT bt_primal(const amigo_ipm_batch* B) {
  return BT{B->blk_primal, B->first_primal, B->count_primal,  B->pseg,          B->cseg,
            B->xseg,       B->active,       B->launch_primal, B->launch_problems};
}
// Grids of a batched launch: the listed blocks / problems, or all of them.
static dim3 grid_total(const amigo_ipm_batch* B) {
  return dim3(B->launch_total ? B->nlaunch_total : B->nblk_total);
}
static dim3 grid_primal(const amigo_ipm_batch* B) {
  return dim3(B->launch_primal ? B->nlaunch_primal : B->nblk_primal);
}
static dim3 grid_problems(const amigo_ipm_batch* B) {
  return dim3(B->launch_problems ? B->nlaunch_problems : B->num_problems);
}

static int batch_ok(const amigo_ipm_batch* B) {
  if (!B || B->num_problems < 1 || B->nblk_total < 1 || B->nblk_primal < 1 || !B->pseg ||
      !B->cseg || !B->xseg || !B->blk_total || !B->first_total || !B->count_total ||
      !B->blk_primal || !B->first_primal || !B->count_primal ||
      (B->launch_total && (B->nlaunch_total < 1 || B->nlaunch_total > B->nblk_total)) ||
      (B->launch_primal && (B->nlaunch_primal < 1 || B->nlaunch_primal > B->nblk_primal)) ||
      (B->launch_problems &&
       (B->nlaunch_problems < 1 || B->nlaunch_problems > B->num_problems))) {
    g_ipm_err = "invalid batch description";
    return -1;
  }
  return 0;
}

// Combine per-problem partials into out_dev[K*NV] and copy to the host.
static int batch_finish(const amigo_ipm_batch* B, BT bt, int NV, const int* ops,
                        const double* ws, double* out_dev, double* out_host, hipStream_t s) {
  const int K = B->num_problems;
  switch (NV) {
    case 1:
      hipLaunchKernelGGL(finish_batch_kernel<1>, grid_problems(B), dim3(NT), 0, s, bt, ws, ops[0], 0, 0,
                         0, out_dev);
      break;
    case 2:
      hipLaunchKernelGGL(finish_batch_kernel<2>, grid_problems(B), dim3(NT), 0, s, bt, ws, ops[0], ops[1],
                         0, 0, out_dev);
      break;
    case 3:
      hipLaunchKernelGGL(finish_batch_kernel<3>, grid_problems(B), dim3(NT), 0, s, bt, ws, ops[0], ops[1],
                         ops[2], 0, out_dev);
      break;
    default:
      hipLaunchKernelGGL(finish_batch_kernel<4>, grid_problems(B), dim3(NT), 0, s, bt, ws, ops[0], ops[1],
                         ops[2], ops[3], out_dev);
      break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err("finish_batch_kernel", e);
  if (out_host) {
    e = hipMemcpyAsync(out_host, out_dev, sizeof(double) * NV * (size_t)K,
                       hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return set_err("hipMemcpyAsync", e);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_err("hipStreamSynchronize", e);
  }
  return 0;
}

int amigo_ipm_batch_residual(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                             const amigo_ipm_state* S, const double* mu, const double* grad,
                             double* res, double* ws, double* out_dev, double* out3,
                             void* stream) {
  if (batch_ok(B)) return -1;
  hipStream_t s = (hipStream_t)stream;
  const BT bt = bt_total(B);
  hipLaunchKernelGGL(residual_kernel, grid_total(B), dim3(NT), 0, s, bt, *P, *S, 0.0, mu,
                     grad, res, ws);
  LAUNCH_CHECK("residual_kernel");
  const int ops[3] = {SUM, SUM, SUM};
  return batch_finish(B, bt, 3, ops, ws, out_dev, out3, s);
}

int amigo_ipm_batch_diagonal(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                             const amigo_ipm_state* S, double* diag, void* stream) {
  if (batch_ok(B)) return -1;
  hipLaunchKernelGGL(diagonal_kernel, grid_total(B), dim3(NT), 0, (hipStream_t)stream,
                     bt_total(B), *P, *S, diag);
  LAUNCH_CHECK("diagonal_kernel");
  return 0;
}

int amigo_ipm_batch_bound_dual_step(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                                    const amigo_ipm_state* S, const double* mu,
                                    const double* px, double* dzl, double* dzu, void* stream) {
  if (batch_ok(B)) return -1;
  hipLaunchKernelGGL(bound_dual_step_kernel, grid_primal(B), dim3(NT), 0,
                     (hipStream_t)stream, bt_primal(B), *P, *S, 0.0, mu, px, dzl, dzu);
  LAUNCH_CHECK("bound_dual_step_kernel");
  return 0;
}

int amigo_ipm_batch_max_step(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                             const amigo_ipm_state* S, const double* tau, const double* px,
                             const double* dzl, const double* dzu, double* ws,
                             double* out_dev, double* out4, void* stream) {
  if (batch_ok(B)) return -1;
  hipStream_t s = (hipStream_t)stream;
  const BT bt = bt_primal(B);
  const int K = B->num_problems;
  hipLaunchKernelGGL(max_step_kernel, grid_primal(B), dim3(NT), 0, s, bt, *P, *S, 0.0,
                     tau, px, dzl, dzu, ws);
  hipLaunchKernelGGL(max_step_finish, grid_problems(B), dim3(256), 0, s, bt, K, *P, ws, 0,
                     out_dev);
  LAUNCH_CHECK("max_step_kernel");
  if (out4) {
    hipError_t e = hipMemcpyAsync(out4, out_dev, 4 * sizeof(double) * (size_t)K,
                                  hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return set_err("hipMemcpyAsync", e);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_err("hipStreamSynchronize", e);
  }
  return 0;
}

int amigo_ipm_batch_apply_step(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                               const amigo_ipm_state* S, const double* ax, const double* az,
                               const double* dxlam, const double* dzl, const double* dzu,
                               const amigo_ipm_state* T, void* stream) {
  if (batch_ok(B)) return -1;
  hipLaunchKernelGGL(apply_step_kernel, grid_total(B), dim3(NT), 0, (hipStream_t)stream,
                     bt_total(B), *P, *S, 0.0, 0.0, ax, az, dxlam, dzl, dzu, 0, *T);
  LAUNCH_CHECK("apply_step_kernel");
  return 0;
}

int amigo_ipm_batch_reduce(int kind, const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                           const amigo_ipm_state* S, const double* mu, const double* a,
                           const double* b, const double* c, double* ws, double* out_dev,
                           double* out4, void* stream) {
  if (batch_ok(B)) return -1;
  hipStream_t s = (hipStream_t)stream;
  const BT bt = bt_total(B);
  hipLaunchKernelGGL(reduce_kernel, grid_total(B), dim3(NT), 0, s, bt, kind, *P, *S, 0.0,
                     mu, a, b, c, ws);
  LAUNCH_CHECK("reduce_kernel");
  int ops[4];
  reduce_ops(kind, ops);
  return batch_finish(B, bt, 4, ops, ws, out_dev, out4, s);
}

int amigo_ipm_batch_reset_bound_multipliers(const amigo_ipm_batch* B,
                                            const amigo_ipm_problem* P,
                                            const amigo_ipm_state* S, const double* mu,
                                            double kappa, void* stream) {
  if (batch_ok(B)) return -1;
  hipLaunchKernelGGL(reset_multipliers_kernel, grid_primal(B), dim3(NT), 0,
                     (hipStream_t)stream, bt_primal(B), *P, *S, 0.0, mu, kappa);
  LAUNCH_CHECK("reset_multipliers_kernel");
  return 0;
}

int amigo_ipm_batch_dual_residual_vector(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                                         const amigo_ipm_state* S, const double* grad,
                                         double* out, void* stream) {
  if (batch_ok(B)) return -1;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(dual_residual_kernel, grid_total(B), dim3(NT), 0, s, bt_total(B), *P,
                     *S, grad, out, 0);
  hipLaunchKernelGGL(dual_residual_fill, grid_primal(B), dim3(NT), 0, s, bt_primal(B), *P,
                     *S, grad, out);
  LAUNCH_CHECK("dual_residual_kernel");
  return 0;
}

int amigo_ipm_batch_initialize(const amigo_ipm_batch* B, const double* lbx0,
                               const double* ubx0, const int* pidx, double relax, double cvt,
                               double kappa1, double kappa2, double* lbx, double* ubx,
                               const amigo_ipm_state* S, int do_relax, int do_project,
                               void* stream) {
  if (batch_ok(B)) return -1;
  hipLaunchKernelGGL(init_kernel, grid_primal(B), dim3(NT), 0, (hipStream_t)stream,
                     bt_primal(B), lbx0, ubx0, 0, pidx, relax, cvt, kappa1, kappa2, lbx, ubx, *S,
                     do_relax, do_project);
  LAUNCH_CHECK("init_kernel");
  return 0;
}

int amigo_ipm_batch_floor_slacks(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                                 const amigo_ipm_state* S, const double* lbx0,
                                 const double* ubx0, const double* mu, const double* s_min,
                                 double slack_move, double* ws, double* out_dev, double* out1,
                                 void* stream) {
  if (batch_ok(B)) return -1;
  hipStream_t s = (hipStream_t)stream;
  const BT bt = bt_primal(B);
  hipLaunchKernelGGL(floor_slacks_kernel, grid_primal(B), dim3(NT), 0, s, bt, *P, *S, lbx0,
                     ubx0, 0.0, mu, 0.0, s_min, slack_move, ws);
  LAUNCH_CHECK("floor_slacks_kernel");
  const int ops[1] = {SUM};
  const int rc = batch_finish(B, bt, 1, ops, ws, out_dev, out1, s);
  if (rc) return rc;
  hipLaunchKernelGGL(snap_primals_kernel, grid_primal(B), dim3(NT), 0, s, bt, *P, *S, lbx0,
                     ubx0, out_dev);
  LAUNCH_CHECK("snap_primals_kernel");
  return 0;
}

int amigo_ipm_batch_scale_rows(const amigo_ipm_batch* B, const int* cidx, const double* dc,
                               double* v, int mode, void* stream) {
  if (batch_ok(B)) return -1;
  hipLaunchKernelGGL(scatter_scale_kernel, grid_total(B), dim3(NT), 0, (hipStream_t)stream,
                     bt_total(B), 0, cidx, dc, v, mode);
  LAUNCH_CHECK("scatter_scale_kernel");
  return 0;
}

int amigo_ipm_batch_ir_init(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                            const amigo_ipm_state* S, const double* mu, const double* grad,
                            const double* rhs_saved, const double* px, double* dzl, double* dzu,
                            double* ws, double* out_dev, double* out1, void* stream) {
  if (batch_ok(B)) return -1;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(ir_init_dz_kernel, grid_primal(B), dim3(NT), 0, s, bt_primal(B), *P,
                     *S, 0.0, mu, px, dzl, dzu);
  const BT bt = bt_total(B);
  hipLaunchKernelGGL(ir_rhs_norm_kernel, grid_total(B), dim3(NT), 0, s, bt, *P, *S, 0.0,
                     mu, grad, rhs_saved, ws);
  LAUNCH_CHECK("ir_init");
  const int ops[1] = {MAX};
  return batch_finish(B, bt, 1, ops, ws, out_dev, out1, s);
}

int amigo_ipm_batch_ir_residual(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                                const amigo_ipm_state* S, const double* mu, const double* grad,
                                const double* rhs_saved, const double* Kpx, const double* px,
                                const double* dzl, const double* dzu, double* econd,
                                double* ezl, double* ezu, double* ws, double* out_dev,
                                double* out2, void* stream) {
  if (batch_ok(B)) return -1;
  hipStream_t s = (hipStream_t)stream;
  const BT bt = bt_total(B);
  hipLaunchKernelGGL(ir_residual_kernel, grid_total(B), dim3(NT), 0, s, bt, *P, *S, 0.0,
                     mu, grad, rhs_saved, Kpx, px, dzl, dzu, econd, ezl, ezu, ws);
  LAUNCH_CHECK("ir_residual_kernel");
  const int ops[2] = {MAX, MAX};
  return batch_finish(B, bt, 2, ops, ws, out_dev, out2, s);
}

int amigo_ipm_batch_ir_update(const amigo_ipm_batch* B, const amigo_ipm_problem* P,
                              const amigo_ipm_state* S, const double* corr, const double* ezl,
                              const double* ezu, double* px, double* dzl, double* dzu,
                              void* stream) {
  if (batch_ok(B)) return -1;
  hipLaunchKernelGGL(ir_update_kernel, grid_total(B), dim3(NT), 0, (hipStream_t)stream,
                     bt_total(B), *P, *S, 0, corr, ezl, ezu, px, dzl, dzu);
  LAUNCH_CHECK("ir_update_kernel");
  return 0;
}

int amigo_segment_absmax(const double* x, const int* seg, int K, double* out_dev,
                         double* out_host, void* stream) {
  if (K <= 0) return 0;
  if (x == nullptr || seg == nullptr || out_dev == nullptr) return -1;
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* ob = reinterpret_cast<unsigned long long*>(out_dev);
  hipLaunchKernelGGL(segment_zero_kernel, dim3((K + 255) / 256), dim3(256), 0, s, K, ob);
  LAUNCH_CHECK("segment_zero_kernel");
  // Blocks per segment: 64 for one segment, fewer as the segments fill the
  // device themselves.
  const int per = std::max(1, std::min(64, 4096 / K));
  hipLaunchKernelGGL(segment_absmax_kernel, dim3(K, per), dim3(NT), 0, s, x, seg, ob);
  LAUNCH_CHECK("segment_absmax_kernel");
  if (out_host) {
    hipError_t e = hipMemcpyAsync(out_host, out_dev, sizeof(double) * (size_t)K,
                                  hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return set_err("hipMemcpyAsync", e);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_err("hipStreamSynchronize", e);
  }
  return 0;
}

int amigo_ipm_batch_copy_state(const amigo_ipm_batch* B, const amigo_ipm_state* dst,
                               const amigo_ipm_state* src, int what, void* stream) {
  if (batch_ok(B)) return -1;
  hipLaunchKernelGGL(copy_state_kernel, grid_total(B), dim3(NT), 0, (hipStream_t)stream,
                     bt_total(B), *dst, *src, what);
  LAUNCH_CHECK("copy_state_kernel");
  return 0;
}

}  // extern "C"
