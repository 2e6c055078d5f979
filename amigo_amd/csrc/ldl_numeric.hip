// Numeric multifrontal LDL^T for gfx950.
//
// Fronts are processed level by level over the relaxed-supernode tree
// (leaves first).  One launch per level assembles every front of the level
// (original entries + children's contribution blocks, coalesced scatter) and
// factors the fronts that fit a workgroup's LDS in place.  Larger fronts are
// factored by a blocked path: a single-workgroup panel kernel that selects
// pivots exactly like the reference's panel loop, then a tiled Schur-update
// kernel over the trailing lower triangle.
//
// Pivoting follows include/sparse_ldl.h:1298-1535 (threshold 1x1 / Duff
// 2x2 tests, delay to the parent).  Root fronts (no contribution rows) use
// Bunch-Kaufman partial pivoting, as LAPACK dsytrf does for the reference's
// root (sparse_ldl.h:1208-1230).  Inertia follows sparse_ldl.h:1917-1969.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <numeric>

#include "ldl_numeric.h"

namespace amigo_amd {

namespace {

constexpr int NT = 256;          // threads per workgroup (4 wave64)
constexpr int NB = 64;           // panel width of the blocked path
constexpr int LDS_MAX_M = 128;   // largest front held in LDS (128 KiB)
constexpr int BIG_M = 128;       // static fronts above this use the blocked path
constexpr int TILE = 64;         // Schur-update tile

struct VI {
  double v;
  int i;
};

__device__ __forceinline__ VI vi_best(VI a, VI b) {
  return (b.v > a.v || (b.v == a.v && b.i < a.i)) ? b : a;
}

// Block-wide argmax (largest value, smallest index on ties) -> all threads.
__device__ VI block_argmax(VI x, VI* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    VI o;
    o.v = __shfl_xor(x.v, off, 64);
    o.i = __shfl_xor(x.i, off, 64);
    x = vi_best(x, o);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = x;
  __syncthreads();
  VI r = red[0];
#pragma unroll
  for (int w = 1; w < NT / 64; w++) r = vi_best(r, red[w]);
  return r;
}

__device__ __forceinline__ void swapd(double& a, double& b) {
  double t = a;
  a = b;
  b = t;
}

// Symmetric interchange of rows/columns a < b of the lower triangle
// (sparse_ldl.h:1252-1274).  Ends with a barrier.
__device__ void symm_swap(double* F, int ld, int m, int* vars, int a, int b) {
  const int t0 = threadIdx.x;
  for (int t = t0; t < a; t += NT) swapd(F[a + (size_t)t * ld], F[b + (size_t)t * ld]);
  if (t0 == 0) {
    swapd(F[a + (size_t)a * ld], F[b + (size_t)b * ld]);
    int v = vars[a];
    vars[a] = vars[b];
    vars[b] = v;
  }
  for (int t = a + 1 + t0; t < b; t += NT)
    swapd(F[t + (size_t)a * ld], F[b + (size_t)t * ld]);
  for (int t = b + 1 + t0; t < m; t += NT)
    swapd(F[t + (size_t)a * ld], F[t + (size_t)b * ld]);
  __syncthreads();
}

// Swap rows a and b of W for columns [0, ncol).
__device__ void swap_w_rows(double* W, int ldw, int a, int b, int ncol) {
  for (int q = threadIdx.x; q < ncol; q += NT)
    swapd(W[a + (size_t)q * ldw], W[b + (size_t)q * ldw]);
  __syncthreads();
}

// Pivot decision shared by the unblocked and the panel paths.  colk(i) and
// coli(i) give the up-to-date values of column k and of column imax.
struct PivotChoice {
  int step;  // 0 = delay, 1 = 1x1, 2 = 2x2
  int kpiv;
  bool zero; // exact zero pivot (Bunch-Kaufman path only)
  bool need_imax_col;
};

constexpr double BK_ALPHA = 0.6403882032022076;  // (1 + sqrt(17)) / 8

// Inertia of a 1x1 / 2x2 pivot block (sparse_ldl.h:1917-1969): zero counts
// as positive like the reference; `zero` is reported separately.
__device__ void count_inertia(int step, double d11, double d21, double d22,
                              bool zero, int* cnt) {
  if (step == 1) {
    if (d11 >= 0.0)
      cnt[0]++;
    else
      cnt[1]++;
    if (zero || d11 == 0.0) cnt[2]++;
  } else {
    double det = d11 * d22 - d21 * d21;
    if (det >= 0.0) {
      if (d11 >= 0.0)
        cnt[0] += 2;
      else
        cnt[1] += 2;
    } else {
      cnt[0]++;
      cnt[1]++;
    }
  }
}

// ---------------------------------------------------------------------------
// Unblocked right-looking LDL^T of one front held at F (LDS or global).
// nfs = fully summed candidates; bk = root front (Bunch-Kaufman, no delay);
// chol = no pivoting (SolverType::CHOLESKY).  Returns the pivot count.
__device__ int factor_unblocked(double* F, int ld, int m, int nfs, int* vars,
                                bool bk, bool chol, double ustab,
                                double pivot_tol, VI* red, int* cnt) {
  const int tid = threadIdx.x;
  int k = 0, nc = nfs;
  while (k < nc) {
    int step = 1, kpiv = k;
    bool zero = false;
    if (!chol) {
      VI a = {-1.0, INT_MAX}, b = {-1.0, INT_MAX};
      for (int i = k + 1 + tid; i < m; i += NT) {
        VI c = {fabs(F[i + (size_t)k * ld]), i};
        if (i < nc)
          a = vi_best(a, c);
        else
          b = vi_best(b, c);
      }
      VI ca = block_argmax(a, red);
      VI cb = block_argmax(b, red);
      const double akk = fabs(F[k + (size_t)k * ld]);
      const bool has_cand = k + 1 < nc;
      const int imax = has_cand ? ca.i : 0;
      const double abswik = has_cand ? ca.v : 0.0;
      if (!bk) {
        const double colmax = fmax(abswik, nc < m ? cb.v : 0.0);
        if (akk >= pivot_tol && akk >= ustab * colmax) {
          step = 1;
          kpiv = k;
        } else if (has_cand) {
          VI r = {-1.0, INT_MAX};
          for (int i = k + tid; i < m; i += NT) {
            if (i == imax) continue;
            double g = i < imax ? F[imax + (size_t)i * ld] : F[i + (size_t)imax * ld];
            r = vi_best(r, VI{fabs(g), i});
          }
          const double rowmax = block_argmax(r, red).v;
          const double aii = fabs(F[imax + (size_t)imax * ld]);
          if (aii >= pivot_tol && aii >= ustab * rowmax) {
            step = 1;
            kpiv = imax;
          } else {
            const double wkk = F[k + (size_t)k * ld];
            const double wik = F[imax + (size_t)k * ld];
            const double wii = F[imax + (size_t)imax * ld];
            const double det = fabs(wkk * wii - wik * wik);
            step = 0;
            if (det >= pivot_tol) {
              const double delta = det / fmax(aii + abswik, akk + abswik);
              if (delta >= ustab * colmax && delta >= ustab * rowmax) {
                step = 2;
                kpiv = imax;
              }
            }
          }
        } else {
          step = 0;
        }
      } else {
        // Bunch-Kaufman partial pivoting over the whole front (nc == m).
        const double colmax = has_cand ? ca.v : 0.0;
        if (fmax(akk, colmax) == 0.0) {
          step = 1;
          kpiv = k;
          zero = true;
        } else if (akk >= BK_ALPHA * colmax) {
          step = 1;
          kpiv = k;
        } else {
          VI r = {-1.0, INT_MAX};
          for (int i = k + tid; i < m; i += NT) {
            if (i == imax) continue;
            double g = i < imax ? F[imax + (size_t)i * ld] : F[i + (size_t)imax * ld];
            r = vi_best(r, VI{fabs(g), i});
          }
          const double rowmax = block_argmax(r, red).v;
          if (akk >= BK_ALPHA * colmax * (colmax / rowmax)) {
            step = 1;
            kpiv = k;
          } else if (fabs(F[imax + (size_t)imax * ld]) >= BK_ALPHA * rowmax) {
            step = 1;
            kpiv = imax;
          } else {
            step = 2;
            kpiv = imax;
          }
        }
      }
    }

    if (step == 0) {
      if (k < nc - 1) symm_swap(F, ld, m, vars, k, nc - 1);
      nc -= 1;
      continue;
    }
    const int ksrc = k + step - 1;
    if (kpiv != ksrc) symm_swap(F, ld, m, vars, ksrc, kpiv);

    if (step == 1) {
      const double d = F[k + (size_t)k * ld];
      const int w = m - k - 1;
      if (!zero) {
        const double dinv = 1.0 / d;
        const int tot = w * w;
        for (int e = tid; e < tot; e += NT) {
          const int ii = e % w, jj = e / w;
          if (ii < jj) continue;
          const int i = k + 1 + ii, j = k + 1 + jj;
          F[i + (size_t)j * ld] -= (F[i + (size_t)k * ld] * dinv) * F[j + (size_t)k * ld];
        }
        __syncthreads();
        for (int i = k + 1 + tid; i < m; i += NT) F[i + (size_t)k * ld] *= dinv;
      }
      if (tid == 0) count_inertia(1, d, 0.0, 0.0, zero, cnt);
      __syncthreads();
      k += 1;
    } else {
      const double a = F[k + (size_t)k * ld];
      const double b = F[k + 1 + (size_t)k * ld];
      const double c = F[k + 1 + (size_t)(k + 1) * ld];
      const double detinv = 1.0 / (a * c - b * b);
      const double i00 = c * detinv, i01 = -b * detinv, i11 = a * detinv;
      const int w = m - k - 2;
      const int tot = w * w;
      for (int e = tid; e < tot; e += NT) {
        const int ii = e % w, jj = e / w;
        if (ii < jj) continue;
        const int i = k + 2 + ii, j = k + 2 + jj;
        const double w1 = F[i + (size_t)k * ld], w2 = F[i + (size_t)(k + 1) * ld];
        const double l1 = w1 * i00 + w2 * i01;
        const double l2 = w1 * i01 + w2 * i11;
        F[i + (size_t)j * ld] -= l1 * F[j + (size_t)k * ld] + l2 * F[j + (size_t)(k + 1) * ld];
      }
      __syncthreads();
      for (int i = k + 2 + tid; i < m; i += NT) {
        const double w1 = F[i + (size_t)k * ld], w2 = F[i + (size_t)(k + 1) * ld];
        F[i + (size_t)k * ld] = w1 * i00 + w2 * i01;
        F[i + (size_t)(k + 1) * ld] = w1 * i01 + w2 * i11;
      }
      if (tid == 0) {
        vars[k] = -(vars[k] + 1);
        vars[k + 1] = -(vars[k + 1] + 1);
        count_inertia(2, a, b, c, false, cnt);
      }
      __syncthreads();
      k += 2;
    }
  }
  return k;
}

// Position of row r of child `c`'s contribution block inside the parent's
// front.  Rows [0, dc) are the child's delayed pivots, placed at
// doff + r; the rest are static rows mapped through cmap.
__device__ __forceinline__ int child_row_pos(const DevPlan& P, int c, int dc,
                                             int doff, int ns, int D, int r) {
  if (r < dc) return doff + r;
  int s = P.cmap[P.cptr[c] + (r - dc)];
  return s < ns ? s : s + D;
}

// One workgroup per front of the level: allocate, assemble, and (unless the
// front is routed to the blocked path) factor.
__global__ __launch_bounds__(NT) void level_kernel(
    DevPlan P, DevState S, const double* __restrict__ vals,
    const double* __restrict__ shift, const int* __restrict__ lnodes,
    const int* __restrict__ big_mask, int lds_doubles, double ustab,
    double pivot_tol, int chol) {
  extern __shared__ double smem[];
  __shared__ VI red[NT / 64];
  __shared__ int sh[4];
  __shared__ int cnt[3];
  const int tid = threadIdx.x;
  const int node = lnodes[blockIdx.x];
  const int ns = P.ns[node], cs = P.cs[node];
  const int kb = P.kptr[node], ke = P.kptr[node + 1];

  if (tid == 0) {
    sh[0] = -1;
    cnt[0] = cnt[1] = cnt[2] = 0;
  }
  if (tid == 0 && atomicAdd(&S.flags[0], 0) == 0) {
    int D = 0;
    for (int q = kb; q < ke; q++) D += S.ndel[P.kids[q]];
    const int m = ns + D + cs;
    const unsigned long long need = (unsigned long long)m * (unsigned long long)m;
    const unsigned long long fo = atomicAdd(&S.tops[0], need);
    const unsigned long long vo = atomicAdd(&S.tops[1], (unsigned long long)m);
    if (fo + need > S.rcap || vo + m > S.icap) {
      atomicExch(&S.flags[0], -4);
      sh[0] = -1;
    } else {
      sh[0] = D;
      sh[1] = m;
      S.foff[node] = (int64_t)fo;
      S.voff[node] = (int64_t)vo;
      S.D[node] = D;
      S.m[node] = m;
    }
  }
  __syncthreads();
  if (sh[0] < 0) return;
  const int D = sh[0], m = sh[1];
  double* G = S.rpool + S.foff[node];
  int* vars = S.ipool + S.voff[node];
  const bool big = big_mask[node] != 0;
  const bool in_lds = !big && (long long)m * m <= lds_doubles;
  double* F = in_lds ? smem : G;
  const int ld = m;

  // Front variables: pivots of the supernode, delayed pivots of the
  // children (in child order), static contribution rows.
  for (int j = tid; j < ns; j += NT) vars[j] = P.nvars[P.nptr[node] + j];
  for (int j = tid; j < cs; j += NT) vars[ns + D + j] = P.crows[P.cptr[node] + j];
  {
    int off = ns;
    for (int q = kb; q < ke; q++) {
      const int c = P.kids[q];
      const int nd = S.ndel[c];
      const int* cv = S.ipool + S.voff[c] + S.npiv[c];
      for (int t = tid; t < nd; t += NT) vars[off + t] = cv[t];
      off += nd;
    }
  }
  const size_t mm = (size_t)m * m;
  for (size_t e = tid; e < mm; e += NT) F[e] = 0.0;
  __syncthreads();

  // Original entries of the supernode's columns.
  for (int t = P.aptr[node] + tid; t < P.aptr[node + 1]; t += NT) {
    int r = P.arow[t];
    r = r < ns ? r : r + D;
    const int c = P.acol[t];
    const double v = vals[P.aidx[t]];
    if (r >= c)
      F[r + (size_t)c * ld] += v;
    else
      F[c + (size_t)r * ld] += v;
  }
  __syncthreads();
  if (shift != nullptr) {
    for (int j = tid; j < ns; j += NT) F[j + (size_t)j * ld] += shift[vars[j]];
    __syncthreads();
  }

  // Extend-add of the children's contribution blocks, one child at a time.
  {
    int doff = ns;
    for (int q = kb; q < ke; q++) {
      const int c = P.kids[q];
      const int pc = S.npiv[c], dc = S.ndel[c], mc = S.m[c];
      const int w = mc - pc;
      const double* Cf = S.rpool + S.foff[c];
      const int tot = w * w;
      for (int e = tid; e < tot; e += NT) {
        const int ii = e % w, jj = e / w;
        if (ii < jj) continue;
        const int pi = child_row_pos(P, c, dc, doff, ns, D, ii);
        const int pj = child_row_pos(P, c, dc, doff, ns, D, jj);
        const double v = Cf[(pc + ii) + (size_t)(pc + jj) * mc];
        if (pi >= pj)
          F[pi + (size_t)pj * ld] += v;
        else
          F[pj + (size_t)pi * ld] += v;
      }
      doff += dc;
      __syncthreads();
    }
  }

  if (big) {
    if (tid == 0) {
      S.pstate[4 * node + 0] = 0;
      S.pstate[4 * node + 1] = ns + D;
      S.pstate[4 * node + 2] = 0;
      S.pstate[4 * node + 3] = 0;
    }
    return;
  }

  const int np = factor_unblocked(F, ld, m, ns + D, vars, cs == 0, chol != 0,
                                  ustab, pivot_tol, red, cnt);
  if (in_lds) {
    for (size_t e = tid; e < mm; e += NT) G[e] = F[e];
  }
  if (tid == 0) {
    S.npiv[node] = np;
    S.ndel[node] = ns + D - np;
    if (cnt[0]) atomicAdd(&S.flags[1], cnt[0]);
    if (cnt[1]) atomicAdd(&S.flags[2], cnt[1]);
    if (cnt[2]) atomicAdd(&S.flags[3], cnt[2]);
  }
}

// ---------------------------------------------------------------------------
// Blocked path for large fronts.  One panel: up to NB-1 pivots selected with
// the reference's left-looking panel loop (sparse_ldl.h:1309-1505), W holds
// the unscaled updated columns (m x NB, ld = m).
__global__ __launch_bounds__(NT) void panel_kernel(DevState S, int node,
                                                   double* __restrict__ W,
                                                   double ustab,
                                                   double pivot_tol, int root,
                                                   int chol) {
  __shared__ VI red[NT / 64];
  __shared__ int cnt[3];
  __shared__ int bad;
  const int tid = threadIdx.x;
  if (tid == 0) bad = atomicAdd(&S.flags[0], 0);
  __syncthreads();
  if (bad != 0) return;
  int* ps = S.pstate + 4 * node;
  const int m = S.m[node];
  double* F = S.rpool + S.foff[node];
  int* vars = S.ipool + S.voff[node];
  const int ld = m, ldw = m;
  int k = ps[0], nc = ps[1];
  if (tid == 0) cnt[0] = cnt[1] = cnt[2] = 0;
  __syncthreads();
  if (k >= nc) {
    if (tid == 0) {
      ps[2] = k;
      ps[3] = 0;
    }
    return;
  }
  const int kstart = k;
  int kw = 0;
  const bool bk = root != 0;
  while (kw < NB - 1 && k < nc) {
    // W[k:, kw] = F[k:, k] - F[k:, kstart:k] W[k, :kw]^T
    for (int i = k + tid; i < m; i += NT) {
      double s = F[i + (size_t)k * ld];
      for (int q = 0; q < kw; q++)
        s -= F[i + (size_t)(kstart + q) * ld] * W[k + (size_t)q * ldw];
      W[i + (size_t)kw * ldw] = s;
    }
    __syncthreads();
    double* wk = W + (size_t)kw * ldw;
    double* wi = W + (size_t)(kw + 1) * ldw;
    int step = 1, kpiv = k;
    bool zero = false;
    if (!chol) {
      VI a = {-1.0, INT_MAX}, b = {-1.0, INT_MAX};
      for (int i = k + 1 + tid; i < m; i += NT) {
        VI c = {fabs(wk[i]), i};
        if (i < nc)
          a = vi_best(a, c);
        else
          b = vi_best(b, c);
      }
      const VI ca = block_argmax(a, red);
      const VI cb = block_argmax(b, red);
      const double akk = fabs(wk[k]);
      const bool has_cand = k + 1 < nc;
      const int imax = has_cand ? ca.i : 0;
      const double abswik = has_cand ? ca.v : 0.0;
      const double colmax = bk ? abswik : fmax(abswik, nc < m ? cb.v : 0.0);
      bool ok1 = bk ? (fmax(akk, colmax) == 0.0 || akk >= BK_ALPHA * colmax)
                    : (akk >= pivot_tol && akk >= ustab * colmax);
      if (ok1) {
        step = 1;
        kpiv = k;
        zero = bk && fmax(akk, colmax) == 0.0;
      } else if (has_cand) {
        // Column imax, updated: W[k:, kw+1].
        for (int i = k + tid; i < m; i += NT) {
          double s = i < imax ? F[imax + (size_t)i * ld] : F[i + (size_t)imax * ld];
          for (int q = 0; q < kw; q++)
            s -= F[i + (size_t)(kstart + q) * ld] * W[imax + (size_t)q * ldw];
          wi[i] = s;
        }
        __syncthreads();
        VI r = {-1.0, INT_MAX};
        for (int i = k + tid; i < m; i += NT) {
          if (i == imax) continue;
          r = vi_best(r, VI{fabs(wi[i]), i});
        }
        const double rowmax = block_argmax(r, red).v;
        const double aii = fabs(wi[imax]);
        if (bk) {
          if (akk >= BK_ALPHA * colmax * (colmax / rowmax)) {
            step = 1;
            kpiv = k;
          } else if (aii >= BK_ALPHA * rowmax) {
            step = 1;
            kpiv = imax;
          } else {
            step = 2;
            kpiv = imax;
          }
        } else if (aii >= pivot_tol && aii >= ustab * rowmax) {
          step = 1;
          kpiv = imax;
        } else {
          const double wkk = wk[k], wik = wk[imax], wii = wi[imax];
          const double det = fabs(wkk * wii - wik * wik);
          step = 0;
          if (det >= pivot_tol) {
            const double delta = det / fmax(aii + abswik, akk + abswik);
            if (delta >= ustab * colmax && delta >= ustab * rowmax) {
              step = 2;
              kpiv = imax;
            }
          }
        }
        if (step == 1 && kpiv == imax) {
          for (int i = k + tid; i < m; i += NT) wk[i] = wi[i];
          __syncthreads();
        }
      } else {
        step = 0;
      }
    }

    if (step == 0) {
      if (k < nc - 1) {
        symm_swap(F, ld, m, vars, k, nc - 1);
        swap_w_rows(W, ldw, k, nc - 1, kw);
      }
      nc -= 1;
      continue;
    }
    const int ksrc = k + step - 1;
    if (kpiv != ksrc) {
      symm_swap(F, ld, m, vars, ksrc, kpiv);
      swap_w_rows(W, ldw, ksrc, kpiv, kw + step);
    }
    if (step == 1) {
      const double d = wk[k];
      if (zero) {
        for (int i = k + tid; i < m; i += NT) F[i + (size_t)k * ld] = 0.0;
      } else {
        const double dinv = 1.0 / d;
        for (int i = k + tid; i < m; i += NT)
          F[i + (size_t)k * ld] = (i == k) ? wk[i] : wk[i] * dinv;
      }
      if (tid == 0) count_inertia(1, d, 0.0, 0.0, zero, cnt);
    } else {
      const double d11 = wk[k], d21 = wk[k + 1], d22 = wi[k + 1];
      const double detinv = 1.0 / (d11 * d22 - d21 * d21);
      const double i00 = d22 * detinv, i01 = -d21 * detinv, i11 = d11 * detinv;
      for (int i = k + 2 + tid; i < m; i += NT) {
        const double w1 = wk[i], w2 = wi[i];
        F[i + (size_t)k * ld] = w1 * i00 + w2 * i01;
        F[i + (size_t)(k + 1) * ld] = w1 * i01 + w2 * i11;
      }
      if (tid == 0) {
        F[k + (size_t)k * ld] = d11;
        F[k + 1 + (size_t)k * ld] = d21;
        F[k + 1 + (size_t)(k + 1) * ld] = d22;
        vars[k] = -(vars[k] + 1);
        vars[k + 1] = -(vars[k + 1] + 1);
        count_inertia(2, d11, d21, d22, false, cnt);
      }
    }
    __syncthreads();
    k += step;
    kw += step;
  }
  if (tid == 0) {
    ps[0] = k;
    ps[1] = nc;
    ps[2] = kstart;
    ps[3] = kw;
    if (cnt[0]) atomicAdd(&S.flags[1], cnt[0]);
    if (cnt[1]) atomicAdd(&S.flags[2], cnt[1]);
    if (cnt[2]) atomicAdd(&S.flags[3], cnt[2]);
  }
}

// Trailing update F[k:, k:] -= F[k:, kstart:kstart+kw] W[k:, :kw]^T on the
// lower triangle, 64x64 tiles (sparse_ldl.h:1172-1197).
__global__ __launch_bounds__(NT) void trailing_kernel(DevState S, int node,
                                                      const double* __restrict__ W) {
  __shared__ double As[TILE][NB + 1];
  __shared__ double Bs[TILE][NB + 1];
  __shared__ int bad;
  if (threadIdx.x == 0) bad = atomicAdd(&S.flags[0], 0);
  __syncthreads();
  if (bad != 0) return;
  const int* ps = S.pstate + 4 * node;
  const int m = S.m[node];
  const int k = ps[0], kstart = ps[2], kw = ps[3];
  const int ti = blockIdx.x, tj = blockIdx.y;
  if (kw == 0 || ti < tj) return;
  const int i0 = k + ti * TILE, j0 = k + tj * TILE;
  if (i0 >= m || j0 >= m) return;
  const double* F0 = S.rpool + S.foff[node];
  double* F = S.rpool + S.foff[node];
  const int ld = m, ldw = m;
  const int tid = threadIdx.x;
  for (int e = tid; e < TILE * kw; e += NT) {
    const int r = e % TILE, q = e / TILE;
    const int i = i0 + r, j = j0 + r;
    As[r][q] = i < m ? F0[i + (size_t)(kstart + q) * ld] : 0.0;
    Bs[r][q] = j < m ? W[j + (size_t)q * ldw] : 0.0;
  }
  __syncthreads();
  const int tx = tid % 16, ty = tid / 16;
  double acc[4][4] = {};
  for (int q = 0; q < kw; q++) {
    double a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; r++) a[r] = As[tx + 16 * r][q];
#pragma unroll
    for (int s = 0; s < 4; s++) b[s] = Bs[ty + 16 * s][q];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int s = 0; s < 4; s++) acc[r][s] += a[r] * b[s];
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int i = i0 + tx + 16 * r;
    if (i >= m) continue;
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int j = j0 + ty + 16 * s;
      if (j < m && i >= j) F[i + (size_t)j * ld] -= acc[r][s];
    }
  }
}

__global__ void finalize_big_kernel(DevState S, int node, int nfs) {
  if (S.flags[0] != 0) return;
  const int nc = S.pstate[4 * node + 1];
  S.npiv[node] = nc;
  S.ndel[node] = nfs - nc;
}

__global__ void reset_kernel(DevState S) {
  if (threadIdx.x == 0) {
    S.tops[0] = 0;
    S.tops[1] = 0;
  }
  if (threadIdx.x < 8) S.flags[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// Level-scheduled solves (sparse_ldl.h:1807-1908).  Forward: t = x[piv];
// unit-lower solve; x[rows below] -= L21 t; t = D^-1 t; x[piv] = t.
__global__ __launch_bounds__(NT) void solve_fwd_kernel(DevState S,
                                                       const int* __restrict__ lnodes,
                                                       double* __restrict__ x,
                                                       int ldx, int nrhs) {
  extern __shared__ double t[];
  const int node = lnodes[blockIdx.x];
  const int p = S.npiv[node], m = S.m[node];
  if (p == 0) return;
  const double* L = S.rpool + S.foff[node];
  const int* vars = S.ipool + S.voff[node];
  const int ld = m, tid = threadIdx.x;
  for (int r = 0; r < nrhs; r++) {
    double* xr = x + (size_t)r * ldx;
    for (int j = tid; j < p; j += NT) {
      int v = vars[j];
      t[j] = xr[v >= 0 ? v : -v - 1];
    }
    __syncthreads();
    for (int j = 0; j < p;) {
      if (vars[j] >= 0) {
        const double tj = t[j];
        for (int i = j + 1 + tid; i < p; i += NT) t[i] -= L[i + (size_t)j * ld] * tj;
        j += 1;
      } else {
        const double t1 = t[j], t2 = t[j + 1];
        for (int i = j + 2 + tid; i < p; i += NT)
          t[i] -= L[i + (size_t)j * ld] * t1 + L[i + (size_t)(j + 1) * ld] * t2;
        j += 2;
      }
      __syncthreads();
    }
    for (int i = p + tid; i < m; i += NT) {
      double s = 0.0;
      for (int j = 0; j < p; j++) s += L[i + (size_t)j * ld] * t[j];
      atomicAdd(&xr[vars[i]], -s);
    }
    if (tid == 0) {
      for (int i = 0; i < p;) {
        if (vars[i] >= 0) {
          t[i] /= L[i + (size_t)i * ld];
          i++;
        } else {
          const double d11 = L[i + (size_t)i * ld], d21 = L[i + 1 + (size_t)i * ld];
          const double d22 = L[i + 1 + (size_t)(i + 1) * ld];
          const double inv = 1.0 / (d11 * d22 - d21 * d21);
          const double x1 = t[i], x2 = t[i + 1];
          t[i] = inv * (d22 * x1 - d21 * x2);
          t[i + 1] = inv * (d11 * x2 - d21 * x1);
          i += 2;
        }
      }
    }
    __syncthreads();
    for (int j = tid; j < p; j += NT) {
      int v = vars[j];
      xr[v >= 0 ? v : -v - 1] = t[j];
    }
    __syncthreads();
  }
}

// Backward: t = x[piv]; t -= L21^T x[rows below]; L11^T solve; x[piv] = t.
__global__ __launch_bounds__(NT) void solve_bwd_kernel(DevState S,
                                                       const int* __restrict__ lnodes,
                                                       double* __restrict__ x,
                                                       int ldx, int nrhs) {
  extern __shared__ double t[];  // m doubles: pivots then the other rows
  const int node = lnodes[blockIdx.x];
  const int p = S.npiv[node], m = S.m[node];
  if (p == 0) return;
  const double* L = S.rpool + S.foff[node];
  const int* vars = S.ipool + S.voff[node];
  const int ld = m, tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  for (int r = 0; r < nrhs; r++) {
    double* xr = x + (size_t)r * ldx;
    for (int j = tid; j < m; j += NT) {
      int v = vars[j];
      t[j] = xr[v >= 0 ? v : -v - 1];
    }
    __syncthreads();
    // One wave per pivot column: t[j] -= sum_{i>=p} L[i, j] t[i].
    for (int j = wid; j < p; j += NT / 64) {
      double s = 0.0;
      for (int i = p + lane; i < m; i += 64) s += L[i + (size_t)j * ld] * t[i];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (lane == 0) t[j] -= s;
    }
    __syncthreads();
    for (int i = p - 1; i >= 0;) {
      if (vars[i] >= 0) {
        const double ti = t[i];
        for (int q = tid; q < i; q += NT) t[q] -= L[i + (size_t)q * ld] * ti;
        i -= 1;
      } else {
        const double t2 = t[i], t1 = t[i - 1];
        for (int q = tid; q < i - 1; q += NT)
          t[q] -= L[i + (size_t)q * ld] * t2 + L[i - 1 + (size_t)q * ld] * t1;
        i -= 2;
      }
      __syncthreads();
    }
    for (int j = tid; j < p; j += NT) {
      int v = vars[j];
      xr[v >= 0 ? v : -v - 1] = t[j];
    }
    __syncthreads();
  }
}

#define HIPCHK(call)                                                   \
  do {                                                                 \
    hipError_t e__ = (call);                                           \
    if (e__ != hipSuccess) {                                           \
      err = std::string(#call) + ": " + hipGetErrorString(e__);        \
      return -2;                                                       \
    }                                                                  \
  } while (0)

template <class T>
int upload(std::vector<void*>& allocs, const std::vector<T>& h, const T** d,
           std::string& err) {
  size_t bytes = std::max<size_t>(1, h.size()) * sizeof(T);
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, bytes));
  allocs.push_back(p);
  if (!h.empty()) HIPCHK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  *d = static_cast<const T*>(p);
  return 0;
}

template <class T>
int dalloc(std::vector<void*>& allocs, size_t count, T** d, std::string& err) {
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T)));
  allocs.push_back(p);
  *d = static_cast<T*>(p);
  return 0;
}

}  // namespace

LdlNumeric::~LdlNumeric() {
  for (void* p : allocs_) (void)hipFree(p);
  if (S_.rpool) (void)hipFree(S_.rpool);
  if (S_.ipool) (void)hipFree(S_.ipool);
  if (W_) (void)hipFree(W_);
}

int LdlNumeric::setup(const SymbolicPlan& plan, const int* rowp,
                      const int* cols, double delay_growth, std::string& err) {
  n_ = plan.n;
  K_ = plan.num_nodes;
  delay_growth_ = delay_growth;
  const int K = K_, n = n_;
  const bool permuted = !plan.iperm.empty();
  auto pinv = [&](int v) { return permuted ? plan.iperm[v] : v; };

  std::vector<int> ns(K), cs(K), var_node(n), var_idx(n);
  for (int k = 0; k < K; k++) {
    ns[k] = plan.node_ptr[k + 1] - plan.node_ptr[k];
    cs[k] = plan.contrib_ptr[k + 1] - plan.contrib_ptr[k];
    for (int j = 0; j < ns[k]; j++) {
      int v = plan.node_to_var[plan.node_ptr[k] + j];
      var_node[v] = k;
      var_idx[v] = j;
    }
  }
  h_ns_ = ns;
  h_cs_ = cs;

  // Static position of every child contribution row inside its parent.
  std::vector<int> cmap(plan.contrib_rows.size(), 0);
  std::vector<int> cpos(n, -1);
  for (int p = 0; p < K; p++) {
    for (int j = 0; j < cs[p]; j++) cpos[plan.contrib_rows[plan.contrib_ptr[p] + j]] = j;
    for (int q = plan.children_ptr[p]; q < plan.children_ptr[p + 1]; q++) {
      int c = plan.children[q];
      for (int t = plan.contrib_ptr[c]; t < plan.contrib_ptr[c + 1]; t++) {
        int v = plan.contrib_rows[t];
        if (var_node[v] == p) {
          cmap[t] = var_idx[v];
        } else if (cpos[v] >= 0) {
          cmap[t] = ns[p] + cpos[v];
        } else {
          err = "symbolic plan inconsistency: child row outside parent front";
          return -1;
        }
      }
    }
    for (int j = 0; j < cs[p]; j++) cpos[plan.contrib_rows[plan.contrib_ptr[p] + j]] = -1;
  }

  // Assembly map of original entries (sparse_ldl.h:1004-1081): column j of
  // the supernode takes rows with permuted index >= its own.
  std::vector<int> aptr(K + 1, 0), aidx, arow, acol;
  for (int k = 0; k < K; k++) {
    for (int j = 0; j < cs[k]; j++) cpos[plan.contrib_rows[plan.contrib_ptr[k] + j]] = j;
    for (int j = 0; j < ns[k]; j++) {
      int var = plan.node_to_var[plan.node_ptr[k] + j];
      int pj = pinv(var);
      for (int ip = rowp[var]; ip < rowp[var + 1]; ip++) {
        int i = cols[ip];
        int fr = -1;
        if (var_node[i] == k)
          fr = var_idx[i];
        else if (cpos[i] >= 0)
          fr = ns[k] + cpos[i];
        if (fr >= 0 && pinv(i) >= pj) {
          aidx.push_back(ip);
          arow.push_back(fr);
          acol.push_back(j);
        }
      }
    }
    for (int j = 0; j < cs[k]; j++) cpos[plan.contrib_rows[plan.contrib_ptr[k] + j]] = -1;
    aptr[k + 1] = int(aidx.size());
  }

  // Levels by height (children have smaller indices in the postorder).
  std::vector<int> level(K, 0);
  int nlev = 0;
  for (int k = 0; k < K; k++) {
    int l = 0;
    for (int q = plan.children_ptr[k]; q < plan.children_ptr[k + 1]; q++)
      l = std::max(l, level[plan.children[q]] + 1);
    level[k] = l;
    nlev = std::max(nlev, l + 1);
  }
  level_ptr_.assign(nlev + 1, 0);
  for (int k = 0; k < K; k++) level_ptr_[level[k] + 1]++;
  for (int l = 0; l < nlev; l++) level_ptr_[l + 1] += level_ptr_[l];
  level_nodes_.assign(K, 0);
  {
    std::vector<int> fill(level_ptr_.begin(), level_ptr_.end() - 1);
    for (int k = 0; k < K; k++) level_nodes_[fill[level[k]]++] = k;
  }
  std::vector<int> big(K, 0);
  max_front_ = 0;
  double rneed = 0.0, ineed = 0.0;
  flops_ = 0.0;
  solve_bytes_ = 0.0;
  for (int k = 0; k < K; k++) {
    int m0 = ns[k] + cs[k];
    max_front_ = std::max(max_front_, m0);
    big[k] = m0 > BIG_M ? 1 : 0;
    rneed += double(m0) * m0;
    ineed += m0;
    for (int q = 0; q < ns[k]; q++) {
      double r = m0 - q - 1;
      flops_ += r + r * (m0 - q);
    }
    solve_bytes_ += 8.0 * 2.0 * double(m0) * ns[k] + 8.0 * 2.0 * m0 + 4.0 * m0;
  }
  level_lds_.assign(nlev, 0);
  big_ptr_.assign(nlev + 1, 0);
  big_nodes_.clear();
  for (int l = 0; l < nlev; l++) {
    int mx = 0;
    for (int t = level_ptr_[l]; t < level_ptr_[l + 1]; t++) {
      int k = level_nodes_[t];
      if (big[k]) {
        big_nodes_.push_back(k);
      } else {
        mx = std::max(mx, std::min(ns[k] + cs[k] + 16, LDS_MAX_M));
      }
    }
    level_lds_[l] = mx * mx;
    big_ptr_[l + 1] = int(big_nodes_.size());
  }

  std::vector<int> parent(plan.node_parent);
  if (upload(allocs_, ns, &P_.ns, err)) return -2;
  if (upload(allocs_, cs, &P_.cs, err)) return -2;
  if (upload(allocs_, plan.contrib_ptr, &P_.cptr, err)) return -2;
  if (upload(allocs_, plan.contrib_rows, &P_.crows, err)) return -2;
  if (upload(allocs_, cmap, &P_.cmap, err)) return -2;
  if (upload(allocs_, plan.children_ptr, &P_.kptr, err)) return -2;
  if (upload(allocs_, plan.children, &P_.kids, err)) return -2;
  if (upload(allocs_, plan.node_ptr, &P_.nptr, err)) return -2;
  if (upload(allocs_, plan.node_to_var, &P_.nvars, err)) return -2;
  if (upload(allocs_, parent, &P_.parent, err)) return -2;
  if (upload(allocs_, aptr, &P_.aptr, err)) return -2;
  if (upload(allocs_, aidx, &P_.aidx, err)) return -2;
  if (upload(allocs_, arow, &P_.arow, err)) return -2;
  if (upload(allocs_, acol, &P_.acol, err)) return -2;
  P_.n = n;
  P_.K = K;
  const int* dl = nullptr;
  if (upload(allocs_, level_nodes_, &dl, err)) return -2;
  d_level_nodes_ = const_cast<int*>(dl);
  const int* db = nullptr;
  if (upload(allocs_, big, &db, err)) return -2;
  d_big_mask_ = const_cast<int*>(db);

  if (dalloc(allocs_, K, &S_.D, err)) return -2;
  if (dalloc(allocs_, K, &S_.npiv, err)) return -2;
  if (dalloc(allocs_, K, &S_.ndel, err)) return -2;
  if (dalloc(allocs_, K, &S_.m, err)) return -2;
  if (dalloc(allocs_, K, &S_.foff, err)) return -2;
  if (dalloc(allocs_, K, &S_.voff, err)) return -2;
  if (dalloc(allocs_, 4 * (size_t)K, &S_.pstate, err)) return -2;
  if (dalloc(allocs_, 2, &S_.tops, err)) return -2;
  if (dalloc(allocs_, 8, &S_.flags, err)) return -2;

  {
    static bool attr_done = false;
    if (!attr_done) {
      HIPCHK(hipFuncSetAttribute((const void*)level_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 LDS_MAX_M * LDS_MAX_M * (int)sizeof(double)));
      HIPCHK(hipFuncSetAttribute((const void*)solve_fwd_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 150 * 1024));
      HIPCHK(hipFuncSetAttribute((const void*)solve_bwd_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 150 * 1024));
      attr_done = true;
    }
  }
  base_rcap_ = (unsigned long long)(std::max(1.0, delay_growth) * rneed) + 4096;
  base_icap_ = (unsigned long long)(std::max(1.0, delay_growth) * ineed) + 4096;
  return ensure_pools(base_rcap_, base_icap_, err);
}

int LdlNumeric::ensure_pools(unsigned long long rcap, unsigned long long icap,
                             std::string& err) {
  if (rcap > S_.rcap) {
    if (S_.rpool) HIPCHK(hipFree(S_.rpool));
    S_.rpool = nullptr;
    S_.rcap = 0;
    hipError_t e = hipMalloc(&S_.rpool, rcap * sizeof(double));
    if (e != hipSuccess) {
      err = "front pool allocation failed (" + std::to_string(rcap * 8) + " bytes)";
      return -5;
    }
    S_.rcap = rcap;
  }
  if (icap > S_.icap) {
    if (S_.ipool) HIPCHK(hipFree(S_.ipool));
    S_.ipool = nullptr;
    S_.icap = 0;
    hipError_t e = hipMalloc(&S_.ipool, icap * sizeof(int));
    if (e != hipSuccess) {
      err = "front index pool allocation failed";
      return -5;
    }
    S_.icap = icap;
  }
  return 0;
}

int LdlNumeric::run_factor(const double* vals, const double* shift,
                           FactorParams prm, bool cholesky,
                           hipStream_t stream, std::string& err) {
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, stream, S_);
  HIPCHK(hipGetLastError());
  const int nlev = num_levels();
  std::vector<int> hm, hD;
  for (int l = 0; l < nlev; l++) {
    const int cnt = level_ptr_[l + 1] - level_ptr_[l];
    const int lds = level_lds_[l];
    hipLaunchKernelGGL(level_kernel, dim3(cnt), dim3(NT),
                       size_t(lds) * sizeof(double), stream, P_, S_, vals,
                       shift, d_level_nodes_ + level_ptr_[l], d_big_mask_, lds,
                       prm.ustab, prm.pivot_tol, cholesky ? 1 : 0);
    HIPCHK(hipGetLastError());
    if (big_ptr_[l + 1] == big_ptr_[l]) continue;
    // Big fronts: sizes are known only after assembly (delayed pivots).
    HIPCHK(hipStreamSynchronize(stream));
    int flag = 0;
    HIPCHK(hipMemcpy(&flag, S_.flags, sizeof(int), hipMemcpyDeviceToHost));
    if (flag != 0) return flag;
    for (int t = big_ptr_[l]; t < big_ptr_[l + 1]; t++) {
      const int node = big_nodes_[t];
      int mD[1], Dv[1];
      HIPCHK(hipMemcpy(mD, S_.m + node, sizeof(int), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(Dv, S_.D + node, sizeof(int), hipMemcpyDeviceToHost));
      const int m = mD[0];
      const int nfs = h_ns_[node] + Dv[0];
      const size_t wneed = (size_t)m * NB;
      if (wneed > W_cap_) {
        if (W_) HIPCHK(hipFree(W_));
        W_ = nullptr;
        HIPCHK(hipMalloc(&W_, wneed * sizeof(double)));
        W_cap_ = wneed;
      }
      const int root = h_cs_[node] == 0 ? 1 : 0;
      const int npanels = nfs / (NB - 1) + 2;
      const int tiles = (m + TILE - 1) / TILE;
      for (int q = 0; q < npanels; q++) {
        hipLaunchKernelGGL(panel_kernel, dim3(1), dim3(NT), 0, stream, S_,
                           node, W_, prm.ustab, prm.pivot_tol, root,
                           cholesky ? 1 : 0);
        hipLaunchKernelGGL(trailing_kernel, dim3(tiles, tiles), dim3(NT), 0,
                           stream, S_, node, (const double*)W_);
      }
      hipLaunchKernelGGL(finalize_big_kernel, dim3(1), dim3(1), 0, stream, S_,
                         node, nfs);
      HIPCHK(hipGetLastError());
    }
  }
  HIPCHK(hipStreamSynchronize(stream));
  int flags[8];
  HIPCHK(hipMemcpy(flags, S_.flags, sizeof(flags), hipMemcpyDeviceToHost));
  if (flags[0] != 0) return flags[0];
  inertia_[0] = flags[1];
  inertia_[1] = flags[2];
  inertia_[2] = flags[3];
  return 0;
}

int LdlNumeric::factor(const double* vals, const double* shift, double ustab,
                       double pivot_tol, bool cholesky, hipStream_t stream,
                       std::string& err) {
  FactorParams prm{ustab, pivot_tol};
  factored_ = false;
  for (int attempt = 0; attempt < 6; attempt++) {
    int rc = run_factor(vals, shift, prm, cholesky, stream, err);
    if (rc == -4) {
      // Delayed pivots outgrew the pools: grow and refactor.
      unsigned long long used[2];
      HIPCHK(hipMemcpy(used, S_.tops, sizeof(used), hipMemcpyDeviceToHost));
      int r2 = ensure_pools(std::max(2 * S_.rcap, used[0] + used[0] / 2),
                            std::max(2 * S_.icap, used[1] + used[1] / 2), err);
      if (r2) return r2;
      continue;
    }
    if (rc) {
      if (err.empty()) err = "factorization failed";
      return rc;
    }
    // Max pivot block per level for the solve kernels' LDS sizing.
    std::vector<int> np(K_), mm(K_);
    HIPCHK(hipMemcpy(np.data(), S_.npiv, K_ * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(mm.data(), S_.m, K_ * sizeof(int), hipMemcpyDeviceToHost));
    max_p_ = 0;
    for (int k = 0; k < K_; k++) max_p_ = std::max(max_p_, mm[k]);
    factored_ = true;
    return 0;
  }
  err = "front pools could not hold the delayed pivots";
  return -4;
}

int LdlNumeric::solve(int nrhs, double* x, int ldx, hipStream_t stream,
                      std::string& err) {
  if (!factored_) {
    err = "solve called before a successful factor";
    return -3;
  }
  const size_t lds = size_t(max_p_) * sizeof(double);
  if (lds > 160 * 1024) {
    err = "front too large for the LDS solve kernels";
    return -1;
  }
  const int nlev = num_levels();
  for (int l = 0; l < nlev; l++) {
    const int cnt = level_ptr_[l + 1] - level_ptr_[l];
    hipLaunchKernelGGL(solve_fwd_kernel, dim3(cnt), dim3(NT), lds, stream, S_,
                       d_level_nodes_ + level_ptr_[l], x, ldx, nrhs);
  }
  for (int l = nlev - 1; l >= 0; l--) {
    const int cnt = level_ptr_[l + 1] - level_ptr_[l];
    hipLaunchKernelGGL(solve_bwd_kernel, dim3(cnt), dim3(NT), lds, stream, S_,
                       d_level_nodes_ + level_ptr_[l], x, ldx, nrhs);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int LdlNumeric::pivot_counts(int* np, int* nd, std::string& err) const {
  HIPCHK(hipMemcpy(np, S_.npiv, K_ * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(nd, S_.ndel, K_ * sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

}  // namespace amigo_amd
