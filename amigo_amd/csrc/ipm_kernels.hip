// Interior-point vector kernels for gfx950 (the ipm:: functions of
// include/interior_point_backend.h and the wrappers of
// include/interior_point_optimizer.h:214-634).
//
// Element-wise kernels stream the vectors once.  Reductions are two-pass and
// deterministic: a grid of wave64 reductions writes per-block partials to the
// caller's workspace, a single workgroup combines them in a fixed order, and
// the results are copied to the host.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>
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
__global__ void residual_kernel(amigo_ipm_problem P, amigo_ipm_state S, double mu,
                                const double* __restrict__ grad,
                                double* __restrict__ res, double* ws) {
  __shared__ double red[(NT / 64) * 3];
  double v[3] = {0.0, 0.0, 0.0};
  const int ops[3] = {SUM, SUM, SUM};
  const int total = P.n_primal + P.n_constraints;
  for (int t = blockIdx.x * NT + threadIdx.x; t < total; t += gridDim.x * NT) {
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
    for (int q = 0; q < 3; q++) ws[blockIdx.x * 3 + q] = v[q];
}

// Sigma (interior_point_backend.h:277-288).
__global__ void diagonal_kernel(amigo_ipm_problem P, amigo_ipm_state S,
                                double* __restrict__ diag) {
  const int total = P.n_primal + P.n_constraints;
  for (int t = blockIdx.x * NT + threadIdx.x; t < total; t += gridDim.x * NT) {
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
__global__ void bound_dual_step_kernel(amigo_ipm_problem P, amigo_ipm_state S,
                                       double mu, const double* __restrict__ px,
                                       double* __restrict__ dzl,
                                       double* __restrict__ dzu) {
  for (int i = blockIdx.x * NT + threadIdx.x; i < P.n_primal; i += gridDim.x * NT) {
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
__global__ void max_step_kernel(amigo_ipm_problem P, amigo_ipm_state S, double tau,
                                const double* __restrict__ px,
                                const double* __restrict__ dzl,
                                const double* __restrict__ dzu, double* ws) {
  __shared__ double rv[2 * (NT / 64)];
  __shared__ long long rk[2 * (NT / 64)];
  double ax = 1.0, az = 1.0;
  long long kx = LLONG_MAX, kz = LLONG_MAX;
  for (int i = blockIdx.x * NT + threadIdx.x; i < P.n_primal; i += gridDim.x * NT) {
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
    ws[blockIdx.x * 4 + 0] = ax;
    ws[blockIdx.x * 4 + 1] = (double)kx;
    ws[blockIdx.x * 4 + 2] = az;
    ws[blockIdx.x * 4 + 3] = (double)kz;
  }
}

__global__ void max_step_finish(amigo_ipm_problem P, const double* ws, int nb,
                                double* out) {
  if (threadIdx.x != 0) return;
  double ax = 1.0, az = 1.0;
  double kx = 9.0e18, kz = 9.0e18;
  for (int b = 0; b < nb; b++) {
    const double a = ws[4 * b], k = ws[4 * b + 1];
    if (a < ax || (a == ax && k < kx)) {
      ax = a;
      kx = k;
    }
    const double c = ws[4 * b + 2], l = ws[4 * b + 3];
    if (c < az || (c == az && l < kz)) {
      az = c;
      kz = l;
    }
  }
  out[0] = ax;
  out[1] = kx < 9.0e18 ? (double)P.primal_indices[(long long)kx / 2] : -1.0;
  out[2] = az;
  out[3] = kz < 9.0e18 ? (double)P.primal_indices[(long long)kz / 2] : -1.0;
}

// Trial step (interior_point_backend.h:375-392).
__global__ void apply_step_kernel(amigo_ipm_problem P, amigo_ipm_state S, double ax,
                                  double az, const double* __restrict__ dxlam,
                                  const double* __restrict__ dzl,
                                  const double* __restrict__ dzu, int n,
                                  amigo_ipm_state T) {
  for (int t = blockIdx.x * NT + threadIdx.x; t < n; t += gridDim.x * NT) {
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

__global__ void reduce_kernel(int kind, amigo_ipm_problem P, amigo_ipm_state S,
                              double mu, const double* __restrict__ a,
                              const double* __restrict__ b,
                              const double* __restrict__ c, double* ws) {
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
  for (int t = blockIdx.x * NT + threadIdx.x; t < total; t += gridDim.x * NT) {
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
    for (int q = 0; q < 4; q++) ws[blockIdx.x * 4 + q] = v[q];
}

// z clamp into [mu/(kappa gap), kappa mu/gap] (backend.h:459-476).
__global__ void reset_multipliers_kernel(amigo_ipm_problem P, amigo_ipm_state S,
                                         double mu, double kappa) {
  for (int i = blockIdx.x * NT + threadIdx.x; i < P.n_primal; i += gridDim.x * NT) {
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

// r_d = grad - zl + zu on primals, 0 elsewhere (backend.h:566-574).
__global__ void dual_residual_kernel(amigo_ipm_problem P, amigo_ipm_state S,
                                     const double* __restrict__ grad,
                                     double* __restrict__ out, int n) {
  for (int t = blockIdx.x * NT + threadIdx.x; t < n; t += gridDim.x * NT) out[t] = 0.0;
}
__global__ void dual_residual_fill(amigo_ipm_problem P, amigo_ipm_state S,
                                   const double* __restrict__ grad,
                                   double* __restrict__ out) {
  for (int i = blockIdx.x * NT + threadIdx.x; i < P.n_primal; i += gridDim.x * NT) {
    const int idx = P.primal_indices[i];
    out[idx] = grad[idx] - S.zl[i] + S.zu[i];
  }
}

// Starting point (backend.h:79-200): relaxed bounds, projection of the
// primals into the interior, unit bound duals, slacks from the primals.
__global__ void init_kernel(const double* __restrict__ lbx0,
                            const double* __restrict__ ubx0, int np,
                            const int* __restrict__ pidx, double relax,
                            double cvt, double kappa1, double kappa2,
                            double* __restrict__ lbx, double* __restrict__ ubx,
                            amigo_ipm_state S, int do_relax, int do_project) {
  for (int i = blockIdx.x * NT + threadIdx.x; i < np; i += gridDim.x * NT) {
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
__global__ void floor_slacks_kernel(amigo_ipm_problem P, amigo_ipm_state S,
                                    const double* __restrict__ lbx0,
                                    const double* __restrict__ ubx0, double mu,
                                    double s_min, double slack_move, double* ws) {
  __shared__ double red[(NT / 64) * 1];
  double v[1] = {0.0};
  const int ops[1] = {SUM};
  for (int i = blockIdx.x * NT + threadIdx.x; i < P.n_primal; i += gridDim.x * NT) {
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
  if (threadIdx.x == 0) ws[blockIdx.x] = v[0];
}

// ---- sparse matrix helpers -------------------------------------------------
// y = A x for the full symmetric CSR (one wave per row).
__global__ void spmv_kernel(int n, const int* __restrict__ rowp,
                            const int* __restrict__ cols,
                            const double* __restrict__ vals,
                            const double* __restrict__ x, double* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int row = (blockIdx.x * NT + threadIdx.x) >> 6;
  const int nw = (gridDim.x * NT) >> 6;
  for (int r = row; r < n; r += nw) {
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
                                 const double* __restrict__ sv, double* __restrict__ vals) {
  const int lane = threadIdx.x & 63;
  const int row = (blockIdx.x * NT + threadIdx.x) >> 6;
  const int nw = (gridDim.x * NT) >> 6;
  for (int r = row; r < n; r += nw) {
    const double d = sv[r];
    for (int p = rowp[r] + lane; p < rowp[r + 1]; p += 64) vals[p] *= d * sv[cols[p]];
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
__global__ void ir_residual_kernel(amigo_ipm_problem P, amigo_ipm_state S,
                                   double mu, const double* __restrict__ grad,
                                   const double* __restrict__ rhs_saved,
                                   const double* __restrict__ Kpx,
                                   const double* __restrict__ px,
                                   const double* __restrict__ dzl,
                                   const double* __restrict__ dzu,
                                   double* __restrict__ econd,
                                   double* __restrict__ ezl,
                                   double* __restrict__ ezu, double* ws) {
  __shared__ double red[(NT / 64) * 2];
  double v[2] = {0.0, 0.0};
  const int ops[2] = {MAX, MAX};
  const int total = P.n_primal + P.n_constraints;
  for (int t = blockIdx.x * NT + threadIdx.x; t < total; t += gridDim.x * NT) {
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
    ws[blockIdx.x * 2 + 0] = v[0];
    ws[blockIdx.x * 2 + 1] = v[1];
  }
}

// Norm of the 8-block RHS: max over |rhs_x|, |rhs_c|, |rhs_zl|, |rhs_zu|.
__global__ void ir_rhs_norm_kernel(amigo_ipm_problem P, amigo_ipm_state S, double mu,
                                   const double* __restrict__ grad,
                                   const double* __restrict__ rhs_saved, double* ws) {
  __shared__ double red[(NT / 64) * 1];
  double v[1] = {0.0};
  const int ops[1] = {MAX};
  const int total = P.n_primal + P.n_constraints;
  for (int t = blockIdx.x * NT + threadIdx.x; t < total; t += gridDim.x * NT) {
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
  if (threadIdx.x == 0) ws[blockIdx.x] = v[0];
}

// dzl += (e_zl - zl dc)/gl, dzu += (e_zu + zu dc)/gu, px += corr.
__global__ void ir_update_kernel(amigo_ipm_problem P, amigo_ipm_state S, int n,
                                 const double* __restrict__ corr,
                                 const double* __restrict__ ezl,
                                 const double* __restrict__ ezu,
                                 double* __restrict__ px, double* __restrict__ dzl,
                                 double* __restrict__ dzu) {
  for (int t = blockIdx.x * NT + threadIdx.x; t < n; t += gridDim.x * NT) {
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
__global__ void ir_init_dz_kernel(amigo_ipm_problem P, amigo_ipm_state S, double mu,
                                  const double* __restrict__ px,
                                  double* __restrict__ dzl, double* __restrict__ dzu) {
  for (int i = blockIdx.x * NT + threadIdx.x; i < P.n_primal; i += gridDim.x * NT) {
    const double dx = px[P.primal_indices[i]];
    const bool fl = finite_bound(P.lbx[i]), fu = finite_bound(P.ubx[i]);
    const double gl = fl ? S.sl[i] : 1.0, gu = fu ? S.su[i] : 1.0;
    dzl[i] = fl ? (mu - gl * S.zl[i] - S.zl[i] * dx) / gl : 0.0;
    dzu[i] = fu ? (mu - gu * S.zu[i] + S.zu[i] * dx) / gu : 0.0;
  }
}

__global__ void scatter_scale_kernel(int nc, const int* __restrict__ cidx,
                                     const double* __restrict__ dc,
                                     double* __restrict__ v, int mode) {
  for (int j = blockIdx.x * NT + threadIdx.x; j < nc; j += gridDim.x * NT) {
    if (mode == 0)
      v[cidx[j]] *= dc[j];
    else if (mode == 1)
      v[cidx[j]] /= dc[j];
    else
      v[j] *= dc[j];
  }
}

int grid_for(int n) { return nblocks(n); }

#define LAUNCH_CHECK(name)                                  \
  do {                                                      \
    hipError_t e__ = hipGetLastError();                     \
    if (e__ != hipSuccess) return set_err(name, e__);       \
  } while (0)

}  // namespace

extern "C" {

const char* amigo_ipm_last_error(void) { return g_ipm_err.c_str(); }

int amigo_ipm_workspace_doubles(void) { return MAXB * 4 + 16; }

int amigo_ipm_residual(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                       double mu, const double* grad, double* res, double* ws,
                       double* out3, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const int total = P->n_primal + P->n_constraints;
  const int nb = nblocks(total);
  hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(NT), 0, s, *P, *S, mu, grad, res, ws);
  LAUNCH_CHECK("residual_kernel");
  const int ops[3] = {SUM, SUM, SUM};
  return finish_and_copy(3, ops, ws, nb, out3, s);
}

int amigo_ipm_diagonal(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                       double* diag, void* stream) {
  const int total = P->n_primal + P->n_constraints;
  hipLaunchKernelGGL(diagonal_kernel, dim3(grid_for(total)), dim3(NT), 0, (hipStream_t)stream, *P, *S, diag);
  LAUNCH_CHECK("diagonal_kernel");
  return 0;
}

int amigo_ipm_bound_dual_step(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                              double mu, const double* px, double* dzl, double* dzu,
                              void* stream) {
  hipLaunchKernelGGL(bound_dual_step_kernel, dim3(grid_for(P->n_primal)), dim3(NT), 0,
                     (hipStream_t)stream, *P, *S, mu, px, dzl, dzu);
  LAUNCH_CHECK("bound_dual_step_kernel");
  return 0;
}

int amigo_ipm_max_step(const amigo_ipm_problem* P, const amigo_ipm_state* S,
                       double tau, const double* px, const double* dzl,
                       const double* dzu, double* ws, double* out4, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const int nb = nblocks(P->n_primal);
  hipLaunchKernelGGL(max_step_kernel, dim3(nb), dim3(NT), 0, s, *P, *S, tau, px, dzl, dzu, ws);
  double* out = ws + (size_t)MAXB * 4;
  hipLaunchKernelGGL(max_step_finish, dim3(1), dim3(64), 0, s, *P, ws, nb, out);
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
                     *P, *S, ax, az, dxlam, dzl, dzu, n, *T);
  LAUNCH_CHECK("apply_step_kernel");
  return 0;
}

int amigo_ipm_reduce(int kind, const amigo_ipm_problem* P, const amigo_ipm_state* S,
                     double mu, const double* a, const double* b, const double* c,
                     double* ws, double* out4, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const int total = P->n_primal + P->n_constraints;
  const int nb = nblocks(total);
  hipLaunchKernelGGL(reduce_kernel, dim3(nb), dim3(NT), 0, s, kind, *P, *S, mu, a, b, c, ws);
  LAUNCH_CHECK("reduce_kernel");
  int ops[4] = {SUM, SUM, SUM, SUM};
  if (kind == RK_COMPL) ops[2] = MIN;
  if (kind == RK_KKT_ERR) ops[0] = ops[1] = ops[2] = MAX;
  if (kind == RK_MAX_COMP_DEV) ops[0] = MAX;
  return finish_and_copy(4, ops, ws, nb, out4, s);
}

int amigo_ipm_reset_bound_multipliers(const amigo_ipm_problem* P,
                                      const amigo_ipm_state* S, double mu,
                                      double kappa, void* stream) {
  hipLaunchKernelGGL(reset_multipliers_kernel, dim3(grid_for(P->n_primal)), dim3(NT), 0,
                     (hipStream_t)stream, *P, *S, mu, kappa);
  LAUNCH_CHECK("reset_multipliers_kernel");
  return 0;
}

int amigo_ipm_dual_residual_vector(const amigo_ipm_problem* P,
                                   const amigo_ipm_state* S, const double* grad,
                                   double* out, int n, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(dual_residual_kernel, dim3(grid_for(n)), dim3(NT), 0, s, *P, *S, grad, out, n);
  hipLaunchKernelGGL(dual_residual_fill, dim3(grid_for(P->n_primal)), dim3(NT), 0, s, *P, *S, grad, out);
  LAUNCH_CHECK("dual_residual_kernel");
  return 0;
}

int amigo_ipm_initialize(const double* lbx0, const double* ubx0, int np,
                         const int* pidx, double relax, double cvt, double kappa1,
                         double kappa2, double* lbx, double* ubx,
                         const amigo_ipm_state* S, int do_relax, int do_project,
                         void* stream) {
  hipLaunchKernelGGL(init_kernel, dim3(grid_for(np)), dim3(NT), 0, (hipStream_t)stream, lbx0,
                     ubx0, np, pidx, relax, cvt, kappa1, kappa2, lbx, ubx, *S, do_relax,
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
  hipLaunchKernelGGL(floor_slacks_kernel, dim3(nb), dim3(NT), 0, s, *P, *S, lbx0, ubx0, mu,
                     s_min, slack_move, ws);
  LAUNCH_CHECK("floor_slacks_kernel");
  const int ops[1] = {SUM};
  return finish_and_copy(1, ops, ws, nb, out1, s);
}

int amigo_csr_spmv(int n, const int* rowp, const int* cols, const double* vals,
                   const double* x, double* y, void* stream) {
  const int waves = n;
  int nb = (waves * 64 + NT - 1) / NT;
  if (nb > 8192) nb = 8192;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(spmv_kernel, dim3(nb), dim3(NT), 0, (hipStream_t)stream, n, rowp, cols,
                     vals, x, y);
  LAUNCH_CHECK("spmv_kernel");
  return 0;
}

int amigo_csr_scale_symmetric(int n, const int* rowp, const int* cols,
                              const double* sv, double* vals, void* stream) {
  int nb = (n * 64 + NT - 1) / NT;
  if (nb > 8192) nb = 8192;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(scale_csr_kernel, dim3(nb), dim3(NT), 0, (hipStream_t)stream, n, rowp,
                     cols, sv, vals);
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
                     nc, cidx, dc, v, mode);
  LAUNCH_CHECK("scatter_scale_kernel");
  return 0;
}

int amigo_ipm_ir_init(const amigo_ipm_problem* P, const amigo_ipm_state* S, double mu,
                      const double* grad, const double* rhs_saved, const double* px,
                      double* dzl, double* dzu, double* ws, double* out1,
                      void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(ir_init_dz_kernel, dim3(grid_for(P->n_primal)), dim3(NT), 0, s, *P, *S,
                     mu, px, dzl, dzu);
  const int total = P->n_primal + P->n_constraints;
  const int nb = nblocks(total);
  hipLaunchKernelGGL(ir_rhs_norm_kernel, dim3(nb), dim3(NT), 0, s, *P, *S, mu, grad,
                     rhs_saved, ws);
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
  hipLaunchKernelGGL(ir_residual_kernel, dim3(nb), dim3(NT), 0, s, *P, *S, mu, grad,
                     rhs_saved, Kpx, px, dzl, dzu, econd, ezl, ezu, ws);
  LAUNCH_CHECK("ir_residual_kernel");
  const int ops[2] = {MAX, MAX};
  return finish_and_copy(2, ops, ws, nb, out2, s);
}

int amigo_ipm_ir_update(const amigo_ipm_problem* P, const amigo_ipm_state* S, int n,
                        const double* corr, const double* ezl, const double* ezu,
                        double* px, double* dzl, double* dzu, void* stream) {
  hipLaunchKernelGGL(ir_update_kernel, dim3(grid_for(n)), dim3(NT), 0, (hipStream_t)stream, *P,
                     *S, n, corr, ezl, ezu, px, dzl, dzu);
  LAUNCH_CHECK("ir_update_kernel");
  return 0;
}

}  // extern "C"
