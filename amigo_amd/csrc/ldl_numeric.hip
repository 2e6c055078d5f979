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

#include <cstring>

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
constexpr int SMALL_LDS_M = 64;  // fronts launched with at most 64^2 LDS doubles
constexpr int KID_BATCH = 64;    // children whose metadata a workgroup holds at once
constexpr int BIG_M = 128;       // static fronts above this use the blocked path
constexpr int TILE = 64;         // Schur-update tile
constexpr int DCAP_BIG = 64;     // delayed pivots a blocked front is sized for
constexpr int NTL = 512;         // level-kernel threads for a level's larger fronts
constexpr int DPAN_BIG = 16;     // delayed pivots the static panel count covers
// A front gets blocked-path launches in the static plan when DPLAN_BIG
// delayed pivots would take it past the LDS limit; one that receives more
// than that past the limit makes the factorization redo itself level-
// synchronously (fronts near 64 pivots then cost no empty panel launches).
constexpr int DPLAN_BIG = 16;
constexpr int PS = 8;            // ints of blocked-path state per front
constexpr int NTP = 1024;        // threads of the panel workgroup (16 wave64)
constexpr int NBW = NB + 2;      // W columns: NB pivots, a 2x2 overhang, imax
constexpr int NBR = 16;          // pivots per panel when the panel rows fit in registers

struct VI {
  double v;
  int i;
};

// A child front as its parent's workgroup reads it: pool offsets, pivots,
// delayed pivots, order and the start of its contribution-row map.
struct KidMeta {
  int64_t foff, voff;
  int pc, dc, mc, cptr;
};

__device__ __forceinline__ VI vi_best(VI a, VI b) {
  return (b.v > a.v || (b.v == a.v && b.i < a.i)) ? b : a;
}

// Block-wide argmax (largest value, smallest index on ties) -> all threads.
template <int NTH>
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
  for (int w = 1; w < NTH / 64; w++) r = vi_best(r, red[w]);
  return r;
}

__device__ __forceinline__ void swapd(double& a, double& b) {
  double t = a;
  a = b;
  b = t;
}

// Symmetric interchange of rows/columns a < b of the lower triangle
// (sparse_ldl.h:1252-1274).  Starts and ends with a barrier: the callers'
// pivot decisions read entries of F that the interchange moves (F[k, k],
// F[imax, imax], F[imax, k]), and every wave must have read them before any
// wave swaps.  Without the leading barrier a wave that falls behind (other
// kernels sharing its CU) can read a swapped value and take a different
// pivot step from the rest of its workgroup.
template <int NTH>
__device__ void symm_swap_t(double* F, int ld, int m, int* vars, int a, int b) {
  const int t0 = threadIdx.x;
  __syncthreads();
  for (int t = t0; t < a; t += NTH) swapd(F[a + (size_t)t * ld], F[b + (size_t)t * ld]);
  if (t0 == 0) {
    swapd(F[a + (size_t)a * ld], F[b + (size_t)b * ld]);
    int v = vars[a];
    vars[a] = vars[b];
    vars[b] = v;
  }
  for (int t = a + 1 + t0; t < b; t += NTH)
    swapd(F[t + (size_t)a * ld], F[b + (size_t)t * ld]);
  for (int t = b + 1 + t0; t < m; t += NTH)
    swapd(F[t + (size_t)a * ld], F[t + (size_t)b * ld]);
  __syncthreads();
}
__device__ void symm_swap(double* F, int ld, int m, int* vars, int a, int b) {
  symm_swap_t<NT>(F, ld, m, vars, a, b);
}
// The same interchange on an LDS-resident front (32-bit offsets).
template <int NTH>
__device__ __forceinline__ void symm_swap_lds(double* F, int ld, int m, int* vars, int a, int b) {
  const int t0 = threadIdx.x;
  __syncthreads();
  for (int t = t0; t < a; t += NTH) swapd(F[a + t * ld], F[b + t * ld]);
  if (t0 == 0) {
    swapd(F[a + a * ld], F[b + b * ld]);
    int v = vars[a];
    vars[a] = vars[b];
    vars[b] = v;
  }
  for (int t = a + 1 + t0; t < b; t += NTH) swapd(F[t + a * ld], F[b + t * ld]);
  for (int t = b + 1 + t0; t < m; t += NTH) swapd(F[t + a * ld], F[t + b * ld]);
  __syncthreads();
}

// Swap rows a and b of W for columns [0, ncol).
template <int NTH>
__device__ void swap_w_rows_t(double* W, int ldw, int a, int b, int ncol) {
  for (int q = threadIdx.x; q < ncol; q += NTH)
    swapd(W[a + (size_t)q * ldw], W[b + (size_t)q * ldw]);
  __syncthreads();
}

// Two block-wide argmax reductions in one pass (NTH threads; red holds
// 2 * NTH / 64 entries).  Results go to all threads.
template <int NTH>
__device__ void block_argmax2(VI& x, VI& y, VI* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    VI o, p;
    o.v = __shfl_xor(x.v, off, 64);
    o.i = __shfl_xor(x.i, off, 64);
    p.v = __shfl_xor(y.v, off, 64);
    p.i = __shfl_xor(y.i, off, 64);
    x = vi_best(x, o);
    y = vi_best(y, p);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    red[wid] = x;
    red[NTH / 64 + wid] = y;
  }
  __syncthreads();
  VI r = red[0], q = red[NTH / 64];
#pragma unroll
  for (int w = 1; w < NTH / 64; w++) {
    r = vi_best(r, red[w]);
    q = vi_best(q, red[NTH / 64 + w]);
  }
  x = r;
  y = q;
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

// A value every thread of the workgroup holds alike (the pivot decisions
// are taken by all threads from the same data), made scalar for the
// compiler: loop bounds, predicates and indices derived from it then live
// in scalar registers and branch without exec masks.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Inertia of a 1x1 / 2x2 pivot block (sparse_ldl.h:1917-1969): zero counts
// as positive like the reference; `zero` is reported separately.
__device__ void count_inertia(int step, double d11, double d21, double d22,
                              bool zero, int* cnt) {
  if (isnan(d11) || isnan(d21) || isnan(d22)) cnt[3]++;
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
template <int NTH>
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
      for (int i = k + 1 + tid; i < m; i += NTH) {
        VI c = {fabs(F[i + (size_t)k * ld]), i};
        if (i < nc)
          a = vi_best(a, c);
        else
          b = vi_best(b, c);
      }
      VI ca = block_argmax<NTH>(a, red);
      VI cb = block_argmax<NTH>(b, red);
      const double akk = fabs(F[k + (size_t)k * ld]);
      // A column of NaNs selects no row: treat it as having no candidate.
      const bool has_cand = k + 1 < nc && ca.i > k && ca.i < nc;
      const int imax = has_cand ? ca.i : 0;
      const double abswik = has_cand ? ca.v : 0.0;
      if (!bk) {
        const double colmax = fmax(abswik, nc < m ? cb.v : 0.0);
        if (akk >= pivot_tol && akk >= ustab * colmax) {
          step = 1;
          kpiv = k;
        } else if (has_cand) {
          VI r = {-1.0, INT_MAX};
          for (int i = k + tid; i < m; i += NTH) {
            if (i == imax) continue;
            double g = i < imax ? F[imax + (size_t)i * ld] : F[i + (size_t)imax * ld];
            r = vi_best(r, VI{fabs(g), i});
          }
          const double rowmax = block_argmax<NTH>(r, red).v;
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
          for (int i = k + tid; i < m; i += NTH) {
            if (i == imax) continue;
            double g = i < imax ? F[imax + (size_t)i * ld] : F[i + (size_t)imax * ld];
            r = vi_best(r, VI{fabs(g), i});
          }
          const double rowmax = block_argmax<NTH>(r, red).v;
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
    step = uni(step);
    kpiv = uni(kpiv);
    zero = uni(zero) != 0;

    if (step == 0) {
      if (k < nc - 1) symm_swap_t<NTH>(F, ld, m, vars, k, nc - 1);
      nc -= 1;
      continue;
    }
    const int ksrc = k + step - 1;
    if (kpiv != ksrc) symm_swap_t<NTH>(F, ld, m, vars, ksrc, kpiv);

    if (step == 1) {
      const double d = F[k + (size_t)k * ld];
      const int w = m - k - 1;
      if (!zero) {
        const double dinv = 1.0 / d;
        const int tot = w * w;
        for (int e = tid; e < tot; e += NTH) {
          const int ii = e % w, jj = e / w;
          if (ii < jj) continue;
          const int i = k + 1 + ii, j = k + 1 + jj;
          F[i + (size_t)j * ld] -= (F[i + (size_t)k * ld] * dinv) * F[j + (size_t)k * ld];
        }
        __syncthreads();
        for (int i = k + 1 + tid; i < m; i += NTH) F[i + (size_t)k * ld] *= dinv;
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
      for (int e = tid; e < tot; e += NTH) {
        const int ii = e % w, jj = e / w;
        if (ii < jj) continue;
        const int i = k + 2 + ii, j = k + 2 + jj;
        const double w1 = F[i + (size_t)k * ld], w2 = F[i + (size_t)(k + 1) * ld];
        const double l1 = w1 * i00 + w2 * i01;
        const double l2 = w1 * i01 + w2 * i11;
        F[i + (size_t)j * ld] -= l1 * F[j + (size_t)k * ld] + l2 * F[j + (size_t)(k + 1) * ld];
      }
      __syncthreads();
      for (int i = k + 2 + tid; i < m; i += NTH) {
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

// Order-preserving key of a candidate |value| for max reductions: 0 for
// "no candidate" (and NaN, which never wins, as in vi_best), otherwise the
// bit pattern of the non-negative double plus one.
__device__ __forceinline__ unsigned long long cand_key(VI c) {
  if (!(c.v >= 0.0)) return 0ull;
  return (unsigned long long)__double_as_longlong(c.v) + 1ull;
}

// Maxima of non-negative doubles (-1 for "none"; no NaN) as maxima of the
// keys bits(x) + 1 (0 for none, +0 and -0 alike): for non-negative values
// the unsigned order of the bit patterns is the order of the values.  A
// 64-bit key maximum is taken on the high words with 32-bit DPP maxima and
// then, only when several lanes share the largest high word, on their low
// words.
__device__ __forceinline__ unsigned long long fmax_key(double x) {
  return x > 0.0 ? (unsigned long long)__double_as_longlong(x) + 1ull : (x == 0.0 ? 1ull : 0ull);
}
__device__ __forceinline__ double fmax_key_value(unsigned long long k) {
  return k ? __longlong_as_double((long long)(k - 1ull)) : -1.0;
}
// Maximum of x over the wave (ROW = false) or over each row of 16 lanes
// (ROW = true), left in lane 63 (the row's lane 15): lanes without a source
// take the identity 0.
template <bool ROW>
__device__ __forceinline__ unsigned dpp_max_u32(unsigned x) {
#define AMIGO_DPP_UMAX(ctrl, rm, bm) \
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, ctrl, rm, bm, false));
  AMIGO_DPP_UMAX(0x111, 0xf, 0xf)  // row_shr:1
  AMIGO_DPP_UMAX(0x112, 0xf, 0xf)  // row_shr:2
  AMIGO_DPP_UMAX(0x113, 0xf, 0xf)  // row_shr:3
  AMIGO_DPP_UMAX(0x114, 0xf, 0xe)  // row_shr:4
  AMIGO_DPP_UMAX(0x118, 0xf, 0xc)  // row_shr:8
  if (!ROW) {
    AMIGO_DPP_UMAX(0x142, 0xa, 0xf)  // row_bcast:15
    AMIGO_DPP_UMAX(0x143, 0xc, 0xf)  // row_bcast:31
  }
#undef AMIGO_DPP_UMAX
  return x;
}
// Largest key of the wave (lanes 0..15 only with ROW), in every lane.  All
// lanes active.
template <bool ROW = false>
__device__ __forceinline__ unsigned long long wave_max_key64(unsigned long long key) {
  constexpr int last = ROW ? 15 : 63;
  const unsigned hi = (unsigned)(key >> 32), lo = (unsigned)key;
  if (__ballot(key != 0ull) == 0ull) return 0ull;
  const unsigned hm = (unsigned)__builtin_amdgcn_readlane((int)dpp_max_u32<ROW>(hi), last);
  const unsigned long long at = __ballot(hi == hm) & (ROW ? 0xffffull : ~0ull);
  unsigned lm;
  if ((at & (at - 1ull)) == 0ull)
    lm = (unsigned)__builtin_amdgcn_readlane((int)lo, __ffsll((long long)at) - 1);
  else
    lm = (unsigned)__builtin_amdgcn_readlane((int)dpp_max_u32<ROW>(hi == hm ? lo : 0u), last);
  return ((unsigned long long)hm << 32) | lm;
}
// Two block-wide argmax reductions (largest value, smallest index on ties;
// lanes hold at most one candidate each, and a lane's index grows with the
// lane id inside a wave) -> all threads.  red holds 2 * NTH / 64 entries.
// The caller guarantees a barrier between consecutive calls.
template <int NTH>
__device__ void block_argmax2_dpp(VI& x, VI& y, VI* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const unsigned long long kx = cand_key(x), ky = cand_key(y);
  const unsigned long long mx = wave_max_key64(kx), my = wave_max_key64(ky);
  const unsigned long long bx = __ballot(kx == mx && mx != 0ull);
  const unsigned long long by = __ballot(ky == my && my != 0ull);
  const int ix = bx ? __builtin_amdgcn_readlane(x.i, __ffsll((long long)bx) - 1) : INT_MAX;
  const int iy = by ? __builtin_amdgcn_readlane(y.i, __ffsll((long long)by) - 1) : INT_MAX;
  if (lane == 0) {
    red[wid] = VI{mx ? __longlong_as_double((long long)(mx - 1ull)) : -1.0, ix};
    red[NTH / 64 + wid] = VI{my ? __longlong_as_double((long long)(my - 1ull)) : -1.0, iy};
  }
  __syncthreads();
  // Across the waves, in every wave: the waves' x winners in lanes 0..NW-1
  // (row 0), their y winners in lanes 32..32+NW-1 (row 2), one DPP row
  // maximum each; equal keys go to the lowest lane, i.e. the lowest wave,
  // whose rows come first.
  constexpr int NW = NTH / 64;
  static_assert(NW <= 16, "one wave winner per lane of a DPP row");
  VI cx = {-1.0, INT_MAX}, cy = {-1.0, INT_MAX};
  if (lane < NW) {
    cx = red[lane];
    cy = red[NW + lane];
  }
  const unsigned long long ckx = cand_key(cx), cky = cand_key(cy);
  const unsigned long long kx2 = wave_max_key64<true>(ckx), ky2 = wave_max_key64<true>(cky);
  const unsigned long long wx = __ballot(ckx == kx2 && kx2 != 0ull) & 0xffffull;
  const unsigned long long wy = __ballot(cky == ky2 && ky2 != 0ull) & 0xffffull;
  x.v = kx2 ? __longlong_as_double((long long)(kx2 - 1ull)) : -1.0;
  x.i = wx ? __builtin_amdgcn_readlane(cx.i, __ffsll((long long)wx) - 1) : INT_MAX;
  y.v = ky2 ? __longlong_as_double((long long)(ky2 - 1ull)) : -1.0;
  y.i = wy ? __builtin_amdgcn_readlane(cy.i, __ffsll((long long)wy) - 1) : INT_MAX;
}

// ---- wave-local pivot search of an LDS-resident front ---------------------
// Every wave reads the column it needs from LDS (two rows per lane: fronts
// of at most LDS_MAX_M = 128 rows) and reduces it with DPP, so all waves
// hold the same pivot statistics without exchanging them through LDS.
static_assert(LDS_MAX_M <= 128, "two rows per lane in the wave-local pivot search");

__device__ __forceinline__ double lane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// One DPP step of a wave maximum of doubles: lanes without a source keep
// their own value.
#define AMIGO_DPP_FMAX(x, ctrl, rm, bm)                                                        \
  {                                                                                            \
    const long long b_ = __double_as_longlong(x);                                              \
    const int lo_ = __builtin_amdgcn_update_dpp((int)b_, (int)b_, ctrl, rm, bm, false);        \
    const int hi_ = __builtin_amdgcn_update_dpp((int)(b_ >> 32), (int)(b_ >> 32), ctrl, rm, bm, \
                                                false);                                        \
    x = fmax(x, __longlong_as_double(((long long)hi_ << 32) | (unsigned)lo_));                 \
  }
// Wave maximum of non-negative values (-1 for "none"; no NaN), every lane:
// a branch-free chain of DPP steps (the LDS fronts' pivot search, where a
// single wave's latency counts).  The panel's block maximum, where the
// VALU work of all waves counts, uses the key form above.
__device__ __forceinline__ double wave_fmax(double x) {
  AMIGO_DPP_FMAX(x, 0x111, 0xf, 0xf)  // row_shr:1
  AMIGO_DPP_FMAX(x, 0x112, 0xf, 0xf)  // row_shr:2
  AMIGO_DPP_FMAX(x, 0x113, 0xf, 0xf)  // row_shr:3
  AMIGO_DPP_FMAX(x, 0x114, 0xf, 0xe)  // row_shr:4
  AMIGO_DPP_FMAX(x, 0x118, 0xf, 0xc)  // row_shr:8
  AMIGO_DPP_FMAX(x, 0x142, 0xa, 0xf)  // row_bcast:15
  AMIGO_DPP_FMAX(x, 0x143, 0xc, 0xf)  // row_bcast:31
  return lane_d(x, 63);
}
#undef AMIGO_DPP_FMAX

// |v| as a candidate value: -1 for rows outside [lo, hi) and for NaN
// (which never wins a pivot search).
__device__ __forceinline__ double cand_abs(double v, int r, int lo, int hi) {
  const double a = fabs(v);
  return (r >= lo && r < hi && a >= 0.0) ? a : -1.0;
}

// Row of the largest candidate (smallest row on ties) of the two entries of
// each lane, rows r0 = base + lane and r1 = r0 + 64; mx is the wave maximum
// of those entries (>= 0).  Every row of the first half precedes every row
// of the second, so the first half is asked first.
__device__ __forceinline__ int wave_argmax_row(double c0, double c1, double mx, int r0, int r1) {
  const unsigned long long b0 = __ballot(c0 == mx), b1 = __ballot(c1 == mx);
  if (b0) return __builtin_amdgcn_readlane(r0, __ffsll((long long)b0) - 1);
  return __builtin_amdgcn_readlane(r1, __ffsll((long long)b1) - 1);
}

// LDS-resident front (m <= LDS_MAX_M): the pivot rules of factor_unblocked
// with one barrier per common step.  Lane l of every wave holds rows
// r = k + l (and r + 64 when ROWS = 2, fronts of more than 64 rows) of the
// pivot column; the largest entry below the diagonal is one DPP reduction,
// and only when the 1x1 test fails are the candidate row imax and row /
// column imax searched.  The decision is then the same in all waves.  The
// trailing update reads the pivot block's unscaled columns (its multipliers
// come from the lanes by readlane), so the columns are scaled to L in the
// next step's update pass by wave 0, whose lanes hold exactly the rows to
// scale (nothing reads those columns in between; the symmetric
// interchanges commute with the scaling, which is the same for every row),
// and after the last step.  Loads are clamped into the front so that they
// need no exec mask; a lane stores its row's entries of the trailing
// columns whenever the row lies in the front, entries above the diagonal
// included: the upper triangle of a front is never read (the interchanges
// and the pivot search touch the lower triangle only) and is not written
// back.
template <int ROWS, int NTH>
__device__ __forceinline__ int factor_front_lds(double* F, int ld, int m, int nfs, int* vars,
                                                bool bk, bool chol, double ustab,
                                                double pivot_tol, int* cnt) {
  static_assert(ROWS == 1 || ROWS == 2, "one or two rows per lane");
  // The wave index as a scalar: the update's column indices, bounds and
  // readlane positions derive from it and stay in scalar registers.
  const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int NW = NTH / 64;
  const int mlast = m - 1;
  int k = 0, nc = uni(nfs);
  // The pivot block whose columns still hold the unscaled values: its rows
  // to scale start at the current k.
  int pk = -1, pstep = 0;
  double q0 = 0.0, q1 = 0.0, q2 = 0.0;  // 1x1: 1/d; 2x2: i00, i01, i11
  __syncthreads();
  while (k < nc) {
    int step = 1, kpiv = k;
    bool zero = false, read_imax = false;
    const int r0 = k + lane, r1 = r0 + 64;
    const int c0 = min(r0, mlast), c1 = min(r1, mlast);
    // Column k from the diagonal down, rows r0 (and r1) of this lane.
    double v0 = F[c0 + k * ld], v1 = ROWS == 2 ? F[c1 + k * ld] : 0.0;
    if (!chol) {
      const double akk = fabs(F[k + k * ld]);
      const double a0 = cand_abs(v0, r0, k + 1, m);
      const double a1 = ROWS == 2 ? cand_abs(v1, r1, k + 1, m) : -1.0;
      // Largest entry below the diagonal: the 1x1 test's column maximum
      // (fronts without contribution rows have candidates only).
      const double allmax = wave_fmax(fmax(a0, a1));
      const double colmax0 = fmax(allmax, 0.0);
      const bool ok1 = bk ? (fmax(akk, colmax0) == 0.0 || akk >= BK_ALPHA * colmax0)
                          : (akk >= pivot_tol && akk >= ustab * colmax0);
      if (ok1) {
        zero = bk && fmax(akk, colmax0) == 0.0;
      } else {
        // Candidates: rows k + 1 .. nc - 1.
        const double e0 = cand_abs(v0, r0, k + 1, nc);
        const double e1 = ROWS == 2 ? cand_abs(v1, r1, k + 1, nc) : -1.0;
        const double cmx = k + 1 < nc ? wave_fmax(fmax(e0, e1)) : -1.0;
        if (cmx >= 0.0) {
          read_imax = true;
          const int imax = wave_argmax_row(e0, e1, cmx, r0, r1);
          const double abswik = cmx;
          const double colmax = colmax0;
          // Row / column imax from row k down, without the diagonal.
          auto g = [&](int i, int ic) {
            const double x = ic < imax ? F[imax + ic * ld] : F[ic + imax * ld];
            const double a = fabs(x);
            return (i < m && i != imax && a >= 0.0) ? a : -1.0;
          };
          const double g0 = g(r0, c0), g1 = ROWS == 2 ? g(r1, c1) : -1.0;
          const double rowmax = wave_fmax(fmax(g0, g1));
          const double aii = fabs(F[imax + imax * ld]);
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
            const double wkk = F[k + k * ld];
            const double wik = F[imax + k * ld];
            const double wii = F[imax + imax * ld];
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
      }
    }
    step = uni(step);
    kpiv = uni(kpiv);
    zero = uni(zero) != 0;
    if (step == 0) {
      // symm_swap_lds starts with a barrier: every wave has read what it moves.
      if (k < nc - 1) symm_swap_lds<NTH>(F, ld, m, vars, k, nc - 1);
      nc -= 1;
      continue;
    }
    const int ksrc = k + step - 1;
    if (kpiv != ksrc) {
      symm_swap_lds<NTH>(F, ld, m, vars, ksrc, kpiv);
      v0 = F[c0 + k * ld];
      if (ROWS == 2) v1 = F[c1 + k * ld];
    } else if (read_imax) {
      __syncthreads();  // row imax is read above and updated below
    }
    // The previous block's columns to L (wave 0: its lanes hold rows k..).
    if (pk >= 0 && wv == 0) {
      if (pstep == 1) {
        if (r0 < m) F[r0 + pk * ld] *= q0;
        if (ROWS == 2 && r1 < m) F[r1 + pk * ld] *= q0;
      } else {
        if (r0 < m) {
          const double w1 = F[r0 + pk * ld], w2 = F[r0 + (pk + 1) * ld];
          F[r0 + pk * ld] = w1 * q0 + w2 * q1;
          F[r0 + (pk + 1) * ld] = w1 * q1 + w2 * q2;
        }
        if (ROWS == 2 && r1 < m) {
          const double w1 = F[r1 + pk * ld], w2 = F[r1 + (pk + 1) * ld];
          F[r1 + pk * ld] = w1 * q0 + w2 * q1;
          F[r1 + (pk + 1) * ld] = w1 * q1 + w2 * q2;
        }
      }
    }
    pk = -1;
    // The trailing lower triangle from this block's unscaled columns:
    // each wave a contiguous block of columns, lanes down the rows, the
    // loads of UC columns in flight together; the pivot columns' entries at
    // row j are uniform LDS reads (broadcast), the column bounds scalar.
    constexpr int UC = 8;
    auto wave_cols = [&](int j0, int& ja, int& je) {
      const int per = (m - j0 + NW - 1) / NW;
      ja = j0 + wv * per;
      je = min(ja + per, m);
    };
    if (step == 1) {
      const double d = F[k + k * ld];
      if (!zero) {
        const double dinv = 1.0 / d;
        const double l0 = v0 * dinv, l1 = v1 * dinv;
        const double* ck = F + k * ld;
        int ja, je;
        wave_cols(k + 1, ja, je);
        for (int jb = ja; jb < je; jb += UC) {
          double cj[UC], f0[UC], f1[UC];
#pragma unroll
          for (int u = 0; u < UC; u++) {
            const int j = min(jb + u, je - 1);
            cj[u] = ck[j];
            f0[u] = F[c0 + j * ld];
            f1[u] = ROWS == 2 ? F[c1 + j * ld] : 0.0;
          }
          if (r0 < m) {
#pragma unroll
            for (int u = 0; u < UC; u++)
              if (jb + u < je) F[r0 + (jb + u) * ld] = f0[u] - l0 * cj[u];
          }
          if (ROWS == 2 && r1 < m) {
#pragma unroll
            for (int u = 0; u < UC; u++)
              if (jb + u < je) F[r1 + (jb + u) * ld] = f1[u] - l1 * cj[u];
          }
        }
        pk = k;
        pstep = 1;
        q0 = dinv;
      }
      if (tid == 0) count_inertia(1, d, 0.0, 0.0, zero, cnt);
      k += 1;
    } else {
      const double a = F[k + k * ld];
      const double b = F[k + 1 + k * ld];
      const double c = F[k + 1 + (k + 1) * ld];
      const double detinv = 1.0 / (a * c - b * b);
      const double i00 = c * detinv, i01 = -b * detinv, i11 = a * detinv;
      // Column k + 1 at this lane's rows; L of the block at rows >= k + 2.
      const double u0 = F[c0 + (k + 1) * ld], u1 = ROWS == 2 ? F[c1 + (k + 1) * ld] : 0.0;
      const double p0 = v0 * i00 + u0 * i01, s0 = v0 * i01 + u0 * i11;
      const double p1 = v1 * i00 + u1 * i01, s1 = v1 * i01 + u1 * i11;
      const double* ck = F + k * ld;
      const double* ck1 = F + (k + 1) * ld;
      int ja, je;
      wave_cols(k + 2, ja, je);
      for (int jb = ja; jb < je; jb += UC) {
        double a1[UC], a2[UC], f0[UC], f1[UC];
#pragma unroll
        for (int u = 0; u < UC; u++) {
          const int j = min(jb + u, je - 1);
          a1[u] = ck[j];
          a2[u] = ck1[j];
          f0[u] = F[c0 + j * ld];
          f1[u] = ROWS == 2 ? F[c1 + j * ld] : 0.0;
        }
        if (r0 < m) {
#pragma unroll
          for (int u = 0; u < UC; u++)
            if (jb + u < je) F[r0 + (jb + u) * ld] = f0[u] - (p0 * a1[u] + s0 * a2[u]);
        }
        if (ROWS == 2 && r1 < m) {
#pragma unroll
          for (int u = 0; u < UC; u++)
            if (jb + u < je) F[r1 + (jb + u) * ld] = f1[u] - (p1 * a1[u] + s1 * a2[u]);
        }
      }
      pk = k;
      pstep = 2;
      q0 = i00;
      q1 = i01;
      q2 = i11;
      if (tid == 0) {
        vars[k] = -(vars[k] + 1);
        vars[k + 1] = -(vars[k + 1] + 1);
        count_inertia(2, a, b, c, false, cnt);
      }
      k += 2;
    }
    __syncthreads();
  }
  // The last block's columns: its rows to scale are those from k down.
  if (pk >= 0 && wv == 0) {
    for (int r = k + lane; r < m; r += 64) {
      if (pstep == 1) {
        F[r + pk * ld] *= q0;
      } else {
        const double w1 = F[r + pk * ld], w2 = F[r + (pk + 1) * ld];
        F[r + pk * ld] = w1 * q0 + w2 * q1;
        F[r + (pk + 1) * ld] = w1 * q1 + w2 * q2;
      }
    }
  }
  __syncthreads();
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

// A node of a block left out of this factorization / solve (batches).
__device__ __forceinline__ bool node_on(const DevState& S, int node) {
  return S.blk_active == nullptr || S.blk_active[S.node_blk[node]] != 0;
}

// The factorization's failure flag, read without a read-modify-write (many
// workgroups check it; an atomic add of zero would serialize them at the
// cache).  A flag set by an earlier launch is visible at the kernel
// boundary; one set in the same launch only means the factorization is
// redone.
__device__ __forceinline__ int fail_flag(const DevState& S) {
  return __hip_atomic_load(&S.flags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One workgroup per front of the level: allocate, assemble, and (unless the
// front is routed to the blocked path) factor.
template <int NTH>
__global__ __launch_bounds__(NTH) void level_kernel(
    DevPlan P, DevState S, const double* __restrict__ vals,
    const double* __restrict__ shift, const int* __restrict__ lnodes,
    const int* __restrict__ big_mask, int lds_doubles, double ustab,
    double pivot_tol, int chol, int dcap_big, const double* __restrict__ sc,
    const int* __restrict__ node_cap) {
  extern __shared__ double smem[];  // F (lds_doubles) then the cached column
  __shared__ VI red[2 * NTH / 64];
  __shared__ VI cred[2 * NTH / 64];
  __shared__ int sh[4];
  __shared__ int cnt[4];
  __shared__ KidMeta km[KID_BATCH];
  __shared__ int rowpos[LDS_MAX_M];
  const int tid = threadIdx.x;
  const int node = lnodes[blockIdx.x];
  if (!node_on(S, node)) return;
  const int ns = P.ns[node], cs = P.cs[node];
  const int kb = P.kptr[node], ke = P.kptr[node + 1];

  if (tid == 0) {
    sh[0] = -1;
    sh[2] = 0;
    cnt[0] = cnt[1] = cnt[2] = cnt[3] = 0;
  }
  __syncthreads();
  // Delayed pivots of the children, summed in parallel (integer sum).  A
  // blocked child that the static panel launches did not finish has no
  // counts: the factorization is redone with per-level sizes.
  for (int q = kb + tid; q < ke; q += NTH) {
    const int c = P.kids[q];
    const int* pc = S.pstate + PS * c;
    if (pc[4] == 1 && pc[0] < pc[1]) atomicExch(&S.flags[0], -7);
    const int nd = S.ndel[c];
    if (nd) atomicAdd(&sh[2], nd);
  }
  __syncthreads();
  if (tid == 0 && fail_flag(S) == 0) {
    int D = sh[2];
    const int m = ns + D + cs;
    if (dcap_big >= 0 && m > LDS_MAX_M && (!big_mask[node] || D > dcap_big)) {
      // The statically planned blocked launches cannot hold this front:
      // the factorization is redone with per-level front sizes.
      atomicExch(&S.flags[0], -7);
      D = -1;
    }
    const unsigned long long need = (unsigned long long)m * (unsigned long long)m;
    unsigned long long fo, vo;
    bool over;
    if (S.node_blk) {
      // The block's own regions of the pools.
      const int b = S.node_blk[node];
      const unsigned long long fr = atomicAdd(&S.tops[2 * b], need);
      const unsigned long long vr = atomicAdd(&S.tops[2 * b + 1], (unsigned long long)m);
      over = fr + need > S.blk_rcap[b] || vr + m > S.blk_icap[b];
      fo = S.blk_rbase[b] + fr;
      vo = S.blk_ibase[b] + vr;
    } else {
      fo = atomicAdd(&S.tops[0], need);
      vo = atomicAdd(&S.tops[1], (unsigned long long)m);
      over = fo + need > S.rcap || vo + m > S.icap;
    }
    if (D < 0) {
      sh[0] = -1;
    } else if (over) {
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
  const int D = uni(sh[0]), m = uni(sh[1]);
  double* G = S.rpool + S.foff[node];
  int* vars = S.ipool + S.voff[node];
  // Fronts larger than the LDS limit go to the blocked path (decided on the
  // actual size, which includes delayed pivots).
  const bool big = m > LDS_MAX_M;
  const long long cap = node_cap ? (long long)node_cap[node] : (long long)lds_doubles;
  const bool in_lds = !big && (long long)m * m <= cap;
  const int ld = m;

  // Front variables: pivots of the supernode, delayed pivots of the
  // children (in child order), static contribution rows.
  for (int j = tid; j < ns; j += NTH) vars[j] = P.nvars[P.nptr[node] + j];
  for (int j = tid; j < cs; j += NTH) vars[ns + D + j] = P.crows[P.cptr[node] + j];
  // Children's metadata, KID_BATCH children at a time (one batch for
  // almost every front, kept for the extend-add below).
  auto load_kids = [&](int q0, int nb) {
    if (tid < nb) {
      const int c = P.kids[q0 + tid];
      km[tid].foff = S.foff[c];
      km[tid].voff = S.voff[c];
      km[tid].pc = S.npiv[c];
      km[tid].dc = S.ndel[c];
      km[tid].mc = S.m[c];
      km[tid].cptr = P.cptr[c];
    }
    __syncthreads();
  };
  const bool one_batch = ke - kb <= KID_BATCH;
  if (D > 0) {
    int off = ns;
    for (int q0 = kb; q0 < ke; q0 += KID_BATCH) {
      const int nb = min(KID_BATCH, ke - q0);
      load_kids(q0, nb);
      for (int qq = 0; qq < nb; qq++) {
        const int nd = km[qq].dc;
        const int* cv = S.ipool + km[qq].voff + km[qq].pc;
        for (int t = tid; t < nd; t += NTH) vars[off + t] = cv[t];
        off += nd;
      }
      __syncthreads();  // km is reloaded next
    }
  }
  if (big) {
    // Assembled by the grid-wide kernels of the blocked path.
    if (tid == 0) {
      S.pstate[PS * node + 0] = 0;
      S.pstate[PS * node + 1] = ns + D;
      S.pstate[PS * node + 2] = 0;
      S.pstate[PS * node + 3] = 0;
      S.pstate[PS * node + 4] = 1;
      S.pstate[PS * node + 5] = ns + D;  // fully summed rows, for the delay count
      for (int q = 0; q < 4; q++) S.ninert[4 * node + q] = 0;
    }
    return;
  }
  // Assembly into the front, instantiated for the LDS front (smem itself,
  // so that the accesses compile to LDS instructions) and for a front in
  // global memory.
  auto assemble = [&](double* F) __attribute__((always_inline)) {
    const size_t mm = (size_t)m * m;
    for (size_t e = tid; e < mm; e += NTH) F[e] = 0.0;
    __syncthreads();

    // Original entries of the supernode's columns.
    for (int t = P.aptr[node] + tid; t < P.aptr[node + 1]; t += NTH) {
      int r = P.arow[t];
      r = r < ns ? r : r + D;
      const int c = P.acol[t];
      double v = vals[P.aidx[t]];
      if (sc) v *= sc[vars[r]] * sc[vars[c]];
      if (r >= c)
        F[r + (size_t)c * ld] += v;
      else
        F[c + (size_t)r * ld] += v;
    }
    __syncthreads();
    if (shift != nullptr) {
      for (int j = tid; j < ns; j += NTH) {
        const int v = vars[j];
        F[j + (size_t)j * ld] += sc ? shift[v] * sc[v] * sc[v] : shift[v];
      }
      __syncthreads();
    }

    // Extend-add of the children's contribution blocks, one child at a time
    // (child order: deterministic sums).  The child's row positions in this
    // front go to LDS first (a child's update rows all lie in this front, so
    // there are at most m <= LDS_MAX_M of them); then each thread has 16 loads
    // of the contribution block in flight.
    {
      int doff = ns;
      for (int q0 = kb; q0 < ke; q0 += KID_BATCH) {
        const int nb = min(KID_BATCH, ke - q0);
        if (!one_batch || D == 0) load_kids(q0, nb);
        for (int qq = 0; qq < nb; qq++) {
          const int pc = km[qq].pc, dc = km[qq].dc, mc = km[qq].mc;
          const int w = mc - pc;
          const int* cm = P.cmap + km[qq].cptr - dc;  // cm[ii] for ii >= dc
          for (int ii = tid; ii < w; ii += NTH) {
            int pos;
            if (ii < dc) {
              pos = doff + ii;
            } else {
              const int sidx = cm[ii];
              pos = sidx < ns ? sidx : sidx + D;
            }
            rowpos[ii] = pos;
          }
          __syncthreads();
          const double* Cf = S.rpool + km[qq].foff + pc + (size_t)pc * mc;
          const int tot = w * w;
          for (int e0 = tid; e0 < tot; e0 += 16 * NTH) {
            double a[16];
  #pragma unroll
            for (int u = 0; u < 16; u++) {
              const int e = e0 + u * NTH, jj = e / w, ii = e - jj * w;
              a[u] = (e < tot && ii >= jj) ? Cf[ii + (size_t)jj * mc] : 0.0;
            }
  #pragma unroll
            for (int u = 0; u < 16; u++) {
              const int e = e0 + u * NTH, jj = e / w, ii = e - jj * w;
              if (e < tot && ii >= jj) {
                const int pi = rowpos[ii], pj = rowpos[jj];
                if (pi >= pj)
                  F[pi + (size_t)pj * ld] += a[u];
                else
                  F[pj + (size_t)pi * ld] += a[u];
              }
            }
          }
          doff += dc;
          __syncthreads();
        }
      }
    }
  };
  if (in_lds)
    assemble(smem);
  else
    assemble(G);

  // The LDS front is passed as smem itself, so that its accesses compile to
  // LDS instructions rather than generic (flat) ones.
  int np;
  if (in_lds) {
    if (m <= 64)
      np = factor_front_lds<1, NTH>(smem, m, m, ns + D, vars, cs == 0, chol != 0, ustab, pivot_tol,
                               cnt);
    else
      np = factor_front_lds<2, NTH>(smem, m, m, ns + D, vars, cs == 0, chol != 0, ustab, pivot_tol,
                               cnt);
    // The lower triangle (nothing reads a front above its diagonal).
    for (int e = tid; e < m * m; e += NTH) {
      const int j = e / m, i = e - j * m;
      if (i >= j) G[e] = smem[e];
    }
  } else {
    np = factor_unblocked<NTH>(G, ld, m, ns + D, vars, cs == 0, chol != 0, ustab,
                          pivot_tol, red, cnt);
  }
  if (tid == 0) {
    S.npiv[node] = np;
    S.ndel[node] = ns + D - np;
    // A planned blocked front that fit in a workgroup: the panel launches
    // of its level find nothing to do.
    S.pstate[PS * node + 0] = np;
    S.pstate[PS * node + 1] = np;
    S.pstate[PS * node + 3] = 0;
    S.pstate[PS * node + 4] = 0;
    // Inertia, delays and the largest front are summed over the nodes
    // after the factorization (factor_stats_kernel).
    for (int q = 0; q < 4; q++) S.ninert[4 * node + q] = cnt[q];
  }
}

// ---------------------------------------------------------------------------
// Grid-wide assembly of the blocked-path fronts of a level (the level kernel
// only allocated them), in one launch: workgroup g of a front owns the front
// columns g, g + G, g + 2G, ... and performs every addition to their lower
// triangle itself, in the order zero (and the diagonal shift of the
// supernode pivots), original entries, children in child order
// (sparse_ldl.h:1004-1156, :1084-1152).  No two workgroups write the same
// entry, and each entry's additions come in a fixed order, so the sums are
// deterministic.  For each child the workgroup places the child's row
// positions in the front (pos) and their inverse (inv) in LDS; the child
// entry (ii, jj), ii >= jj, goes to column pos[jj] when pos[ii] >= pos[jj]
// (read down child column jj) and otherwise, transposed, to column pos[ii]
// (read along child row ii; only a child's delayed rows, placed after the
// supernode rows, land there).  Dynamic LDS: 2 * lds_ints ints.
__device__ __forceinline__ bool big_active(const DevState& S, int node) {
  return node_on(S, node) && S.pstate[PS * node + 4] == 1 && fail_flag(S) == 0;
}

__global__ __launch_bounds__(NT) void big_assemble_kernel(DevPlan P, DevState S,
                                                          const int* __restrict__ bnodes,
                                                          const double* __restrict__ vals,
                                                          const double* __restrict__ shift,
                                                          const double* __restrict__ sc,
                                                          int lds_ints) {
  extern __shared__ int smap[];
  int* inv = smap;            // front position -> child row (or -1)
  int* pos = smap + lds_ints; // child row -> front position
  const int node = bnodes[blockIdx.y];
  if (!big_active(S, node)) return;
  const int m = S.m[node], ns = P.ns[node], D = S.D[node];
  const int tid = threadIdx.x, G = gridDim.x, g = blockIdx.x;
  if (m > lds_ints) {
    // More rows than the launch was sized for: redo the factorization with
    // per-level sizes.
    if (g == 0 && tid == 0) atomicExch(&S.flags[0], -7);
    return;
  }
  double* F = S.rpool + S.foff[node];
  const int* vars = S.ipool + S.voff[node];
  const int vb = P.nptr[node];
  // Zero, with the diagonal shift of the supernode pivots.
  for (int c = g; c < m; c += G) {
    double* Fc = F + (size_t)c * m;
    double dc = 0.0;
    if (c < ns && shift != nullptr) {
      const int v = vars[c];
      dc = sc ? shift[v] * sc[v] * sc[v] : shift[v];
    }
    for (int i = c + tid; i < m; i += NT) Fc[i] = i == c ? dc : 0.0;
  }
  __syncthreads();
  // Original entries of the supernode columns (each on or below the
  // diagonal of its column).
  for (int c = g; c < ns; c += G) {
    double* Fc = F + (size_t)c * m;
    const int sc_c = sc ? vars[c] : 0;
    for (int t = P.acstart[vb + c] + tid; t < P.acstart[vb + c + 1]; t += NT) {
      int r = P.arow[t];
      r = r < ns ? r : r + D;
      double v = vals[P.aidx[t]];
      if (sc) v *= sc[vars[r]] * sc[sc_c];
      Fc[r] += v;
    }
  }
  // Children, in child order.
  const int kb = P.kptr[node], ke = P.kptr[node + 1];
  int doff = ns;
  for (int q = kb; q < ke; q++) {
    const int ch = P.kids[q];
    const int pc = S.npiv[ch], dc = S.ndel[ch], mc = S.m[ch];
    const int w = mc - pc;
    __syncthreads();  // the previous child's additions and maps are done
    for (int i = tid; i < m; i += NT) inv[i] = -1;
    __syncthreads();
    for (int ii = tid; ii < w; ii += NT) {
      const int p = child_row_pos(P, ch, dc, doff, ns, D, ii);
      pos[ii] = p;
      inv[p] = ii;
    }
    __syncthreads();
    const double* Cf = S.rpool + S.foff[ch] + pc + (size_t)pc * mc;  // C[ii, jj] = Cf[ii + jj mc]
    for (int c = g; c < m; c += G) {
      const int x = inv[c];
      if (x < 0) continue;
      double* Fc = F + (size_t)c * m;
      // Down child column x: rows whose position is at or below c.
      const double* Cx = Cf + (size_t)x * mc;
      for (int ii = x + tid; ii < w; ii += NT) {
        const int pi = pos[ii];
        if (pi >= c) Fc[pi] += Cx[ii];
      }
      // Along child row x: earlier child columns placed below row c.
      for (int jj = tid; jj < x; jj += NT) {
        const int pj = pos[jj];
        if (pj > c) Fc[pj] += Cf[x + (size_t)jj * mc];
      }
    }
    doff += dc;
  }
}

// s - sum_q Fi[q * ld] * wr[q] for q < kw, four independent partial sums so
// that the strided loads of the panel columns overlap.
__device__ __forceinline__ double panel_dot(double s, const double* __restrict__ Fi,
                                            size_t ld, const double* wr, int kw) {
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int q = 0;
#pragma unroll 2
  for (; q + 4 <= kw; q += 4) {
    const double f0 = Fi[(size_t)q * ld], f1 = Fi[(size_t)(q + 1) * ld];
    const double f2 = Fi[(size_t)(q + 2) * ld], f3 = Fi[(size_t)(q + 3) * ld];
    s -= f0 * wr[q];
    s1 -= f1 * wr[q + 1];
    s2 -= f2 * wr[q + 2];
    s3 -= f3 * wr[q + 3];
  }
  for (; q < kw; q++) s -= Fi[(size_t)q * ld] * wr[q];
  return (s + s1) + (s2 + s3);
}

// LDS-resident panel (fronts with at most NTH rows from the panel's first
// pivot down, NTH = 256, 512 or 1024 threads by the level's largest planned
// front): thread t owns row kstart + t, and the panel's columns
// [kstart, kstart + NBR + 2) of all rows >= kstart live in LDS, column-major
// (Ls[q * NTH + t]): not yet updated columns, overwritten by L as they are
// factored.  A column update is conflict-free LDS reads against row k of
// W = L D (formed from the L rows and the D blocks) broadcast through LDS;
// the imax candidate column and symmetric interchanges inside the panel
// stay in LDS; the block goes back to global memory at the end.  Pivot
// rules and the W contents are those of the global-memory loop below.
constexpr int LS_COLS = NBR + 2;
constexpr size_t PANEL_LDS_BYTES = sizeof(double) * LS_COLS * NTP;

struct PanelShared {
  VI red[2 * NTP / 64];
  unsigned long long wkey[2][NTP / 64];  // per-wave column maxima as keys (double-buffered)
  double wrow[2][NBR + 2];   // row k of W (double-buffered)
  double wrow2[NBR + 2];
  double xch[4];
  double sval[2][4];      // pivot-row values (double-buffered)
  int cnt[4];
};

// Columns Q0 .. Q0 + NQ - 1 of s - sum_{q < kw} Ls[q][t] w[q] (even columns
// into s, odd into s1): all NQ LDS reads are issued before the first FMA
// waits on one.  In the last part (LAST) the columns past kw count as zero
// (their LDS entries are read, in bounds, and dropped).
template <int NTH, int Q0, int NQ, bool LAST>
__device__ __forceinline__ void lds_dot_part(double& s, double& s1, const double* Ls, int t,
                                             const double* w, int kw) {
  double a[NQ], b[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    a[q] = Ls[(Q0 + q) * NTH + t];
    b[q] = w[Q0 + q];
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    double aq = a[q], bq = b[q];
    if (LAST && q >= NQ - 4) {
      aq = Q0 + q < kw ? aq : 0.0;
      bq = Q0 + q < kw ? bq : 0.0;
    }
    if ((Q0 + q) & 1)
      s1 -= aq * bq;
    else
      s -= aq * bq;
  }
}

// s - sum_{q < kw} Ls[q][t] w[q] for a workgroup-uniform kw < NBR: one
// uniform branch on the number of 4-column chunks, then at most two
// batches of reads (8 columns each), so that the loads are not serialized
// behind per-column branches and the registers in flight stay bounded.
template <int NTH>
__device__ __forceinline__ double lds_dot(double s, const double* Ls, int t, const double* w,
                                          int kw) {
  static_assert(NBR <= 16 && LS_COLS >= 16, "four chunks of four panel columns");
  double s1 = 0.0;
  switch ((kw + 3) >> 2) {
    case 0:
      return s;
    case 1:
      lds_dot_part<NTH, 0, 4, true>(s, s1, Ls, t, w, kw);
      break;
    case 2:
      lds_dot_part<NTH, 0, 8, true>(s, s1, Ls, t, w, kw);
      break;
    case 3:
      lds_dot_part<NTH, 0, 8, false>(s, s1, Ls, t, w, kw);
      lds_dot_part<NTH, 8, 4, true>(s, s1, Ls, t, w, kw);
      break;
    default:
      lds_dot_part<NTH, 0, 8, false>(s, s1, Ls, t, w, kw);
      lds_dot_part<NTH, 8, 8, true>(s, s1, Ls, t, w, kw);
      break;
  }
  return s + s1;
}

// Exchange the current column values s, s2 of the owners of rows a, b
// (owner-relative ta, tb).  Ends with a barrier.
__device__ __forceinline__ void swap_lds_rows_values(double& s, double& s2, int tid, int ta,
                                                     int tb, PanelShared& sh) {
  if (tid == ta) {
    sh.xch[0] = s;
    sh.xch[1] = s2;
  }
  if (tid == tb) {
    sh.xch[2] = s;
    sh.xch[3] = s2;
  }
  __syncthreads();
  if (tid == ta) {
    s = sh.xch[2];
    s2 = sh.xch[3];
  }
  if (tid == tb) {
    s = sh.xch[0];
    s2 = sh.xch[1];
  }
  __syncthreads();
}

// Block-wide maximum of non-negative values (-1: none), every thread; buf
// selects the double buffer.  One barrier.
template <int NTH>
__device__ __forceinline__ double block_fmax(double x, PanelShared& sh, int buf) {
  constexpr int NW = NTH / 64;
  static_assert(NW <= 16, "one wave maximum per lane of a DPP row");
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const unsigned long long w = wave_max_key64(fmax_key(x));
  if (lane == 0) sh.wkey[buf][wid] = w;
  __syncthreads();
  // The waves' maxima in lanes 0..NW-1 of every wave, one row maximum.
  return fmax_key_value(wave_max_key64<true>(lane < NW ? sh.wkey[buf][lane] : 0ull));
}

// Panel block <-> global front: columns [kstart, kstart + pw), rows >= the
// column (lower triangle).
// The batched form (loads in flight together) is for the panel's first
// load; the reload after an interchange outside the panel, where registers
// are scarce, goes column by column.
template <bool BATCH, int NTH>
__device__ __forceinline__ void panel_load(const double* F, int ld, int m, int kstart, int pw,
                                           double* Ls) {
  const int t = threadIdx.x, rows = m - kstart;
  if (!BATCH) {
    if (t < rows)
      for (int q = 0; q < pw; q++)
        Ls[q * NTH + t] = t >= q ? F[kstart + t + (size_t)(kstart + q) * ld] : 0.0;
    return;
  }
  if (t < rows) {
    // The row's loads are in flight PB at a time before their LDS stores.
    constexpr int PB = 18;
    static_assert(LS_COLS % PB == 0, "panel columns in whole batches");
    const double* Fr = F + kstart + t + (size_t)kstart * ld;
#pragma unroll
    for (int q0 = 0; q0 < LS_COLS; q0 += PB) {
      double a[PB];
#pragma unroll
      for (int u = 0; u < PB; u++) {
        const int q = q0 + u;
        a[u] = (q < pw && t >= q) ? Fr[(size_t)q * ld] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < PB; u++)
        if (q0 + u < pw) Ls[(q0 + u) * NTH + t] = a[u];
    }
  }
}
template <int NTH>
__device__ __forceinline__ void panel_store(double* F, int ld, int m, int kstart, int pw,
                                            const double* Ls) {
  const int t = threadIdx.x, rows = m - kstart;
  if (t < rows)
    for (int q = 0; q < pw && q <= t; q++)
      F[kstart + t + (size_t)(kstart + q) * ld] = Ls[q * NTH + t];
}

// Symmetric interchange a < b of the front when both columns lie in the
// LDS panel (sparse_ldl.h:1252-1274 restricted to the stored lower
// triangle): rows a, b of the earlier panels' columns are swapped in global
// memory, everything else in LDS.  Ends with a barrier.
template <int NTH>
__device__ void panel_symm_swap(double* F, int ld, int m, int kstart, int* vars, int a, int b,
                                double* Ls) {
  const int t0 = threadIdx.x;
  const int qa = a - kstart, qb = b - kstart;
  // Row part, columns c < a.
  for (int c = t0; c < a; c += NTH) {
    if (c < kstart)
      swapd(F[a + (size_t)c * ld], F[b + (size_t)c * ld]);
    else
      swapd(Ls[(c - kstart) * NTH + qa], Ls[(c - kstart) * NTH + qb]);
  }
  if (t0 == 0) {
    swapd(Ls[qa * NTH + qa], Ls[qb * NTH + qb]);
    const int v = vars[a];
    vars[a] = vars[b];
    vars[b] = v;
  }
  // a < t < b: F[t, a] <-> F[b, t].
  for (int t = a + 1 + t0; t < b; t += NTH)
    swapd(Ls[qa * NTH + (t - kstart)], Ls[(t - kstart) * NTH + qb]);
  // t > b: F[t, a] <-> F[t, b].
  for (int t = b + 1 + t0; t < m; t += NTH)
    swapd(Ls[qa * NTH + (t - kstart)], Ls[qb * NTH + (t - kstart)]);
  __syncthreads();
}

// One LDS panel (NBR pivots) of a blocked front.  Column k is updated with
// row k of W (the unscaled updated columns, W[k, :kw], as the reference's
// panel loop uses it), read from global memory by the lanes of the last
// wave.  In the common step (a 1x1 pivot without interchange) the next
// pivot row is k + 1, so its W row is published for the next column before
// the step's one barrier (its earlier entries read while the column's dot
// products run, this column's entry the row's own value), and that barrier
// also carries the column maximum (a value-only reduction) and the pivot
// row's value; any other step publishes the next W row after an extra
// barrier.  Shared values used across that single barrier are
// double-buffered.
template <int NTH>
__device__ void panel_lds(double* F, double* W, int* vars, int m, int& k, int& nc, int& kw,
                          int kstart, bool bk, bool chol, double ustab, double pivot_tol,
                          double* Ls, PanelShared& sh) {
  const int ld = m, ldw = m;
  const int tid = threadIdx.x;
  const int r = kstart + tid;
  const bool own = r < m;
  const int pw = min(LS_COLS, m - kstart);
  // Lane q of the last wave reads W[row, q] of a row for the W-row buffers.
  const int wlane = tid - (NTH - 64);
  panel_load<true, NTH>(F, ld, m, kstart, pw, Ls);
  __syncthreads();
  // Interchange of a < b, in LDS when both columns are in the panel,
  // otherwise through global memory (panel written back and reloaded).
  auto interchange = [&](int a, int b, int wcols, double& s, double& s2) {
    if (b - kstart < pw) {
      panel_symm_swap<NTH>(F, ld, m, kstart, vars, a, b, Ls);
    } else {
      panel_store<NTH>(F, ld, m, kstart, pw, Ls);
      __syncthreads();
      symm_swap_t<NTH>(F, ld, m, vars, a, b);
      panel_load<false, NTH>(F, ld, m, kstart, pw, Ls);
    }
    swap_w_rows_t<NTH>(W, ldw, a, b, wcols);
    swap_lds_rows_values(s, s2, tid, a - kstart, b - kstart, sh);
  };
  // Publish row k of W into wrow[buf] (the rows' entries are in W).
  auto publish = [&](int buf) {
    if (wlane >= 0 && wlane < kw && k < m) sh.wrow[buf][wlane] = W[k + (size_t)wlane * ldw];
  };
  int wb = 0;
  k = uni(k);
  nc = uni(nc);
  kw = uni(kw);
  // The first pivot row's W row is empty (kw = 0 at the panel's start).
  while (kw < NBR && k < nc) {
    const double* wr = sh.wrow[wb];
    double s = 0.0, s2 = 0.0;
    double a_all = -1.0;
    // Row k + 1 of W over the earlier columns, in flight during the dots.
    const bool wnext = wlane >= 0 && wlane < kw && k + 1 < m;
    const double wn = wnext ? W[k + 1 + (size_t)wlane * ldw] : 0.0;
    if (own && r >= k) {
      s = lds_dot<NTH>(Ls[kw * NTH + tid], Ls, tid, wr, kw);
      W[r + (size_t)kw * ldw] = s;
      if (r == k) {
        sh.sval[wb][0] = s;
        sh.sval[wb][3] = 1.0 / s;  // the 1x1 pivot's reciprocal, off the others' path
      }
      if (r > k) {
        const double a = fabs(s);
        a_all = a >= 0.0 ? a : -1.0;
      }
      // The next column's W row if this step is a 1x1 pivot without
      // interchange: row k + 1, this column's entry included.
      if (r == k + 1) sh.wrow[wb ^ 1][kw] = s;
    }
    if (wnext) sh.wrow[wb ^ 1][wlane] = wn;
    int step = 1, kpiv = k;
    bool zero = false, plain = true;
    double colmax0 = 0.0, akk = 0.0;
    double wn2 = 0.0;  // W[k + 2, wlane]: the next W row after a 2x2 pivot at k, k + 1
    if (chol) {
      __syncthreads();
    } else {
      // Largest entry below the diagonal (candidates and contribution rows):
      // the 1x1 test's column maximum.
      colmax0 = fmax(block_fmax<NTH>(a_all, sh, wb), 0.0);
      akk = fabs(sh.sval[wb][0]);
      const bool ok1 = bk ? (fmax(akk, colmax0) == 0.0 || akk >= BK_ALPHA * colmax0)
                          : (akk >= pivot_tol && akk >= ustab * colmax0);
      if (ok1) {
        zero = bk && fmax(akk, colmax0) == 0.0;
      } else {
        plain = false;
        // The candidate with the largest entry (smallest row on ties).
        VI a = {-1.0, INT_MAX}, b = {-1.0, INT_MAX};
        if (own && r > k && r < nc) a = VI{fabs(s), r};
        block_argmax2_dpp<NTH>(a, b, sh.red);
        const bool has_cand = k + 1 < nc && a.i > k && a.i < nc;
        const int imax = has_cand ? a.i : 0;
        const double abswik = has_cand ? a.v : 0.0;
        const double colmax = colmax0;
        if (has_cand) {
          if (wlane >= 0 && wlane < kw) sh.wrow2[wlane] = W[imax + (size_t)wlane * ldw];
          if (imax == k + 1 && wlane >= 0 && wlane < kw && k + 2 < m)
            wn2 = W[k + 2 + (size_t)wlane * ldw];
          // Not yet updated column imax, rows >= k: rows below imax from its
          // column, rows k..imax-1 from row imax (the columns r < imax).
          double fi = 0.0;
          if (own && r >= k) {
            const int qi = imax - kstart;
            if (r >= imax)
              fi = qi < pw ? Ls[qi * NTH + tid] : F[r + (size_t)imax * ld];
            else
              fi = tid < pw ? Ls[tid * NTH + qi] : F[imax + (size_t)r * ld];
          }
          __syncthreads();
          VI rr = {-1.0, INT_MAX}, dummy = {-1.0, INT_MAX};
          if (own && r >= k) {
            s2 = lds_dot<NTH>(fi, Ls, tid, sh.wrow2, kw);
            W[r + (size_t)(kw + 1) * ldw] = s2;
            if (r == imax) {
              sh.sval[wb][1] = s2;
              sh.sval[wb][2] = s;
            } else {
              rr = VI{fabs(s2), r};
            }
          }
          block_argmax2_dpp<NTH>(rr, dummy, sh.red);
          const double rowmax = rr.v;
          const double aii = fabs(sh.sval[wb][1]);
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
            const double wkk = sh.sval[wb][0], wii = sh.sval[wb][1], wik = sh.sval[wb][2];
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
          if (step == 1 && kpiv == imax && own && r >= k) {
            s = s2;
            W[r + (size_t)kw * ldw] = s2;
          }
        } else {
          step = 0;
        }
      }
    }
    step = uni(step);
    kpiv = uni(kpiv);
    zero = uni(zero) != 0;
    plain = uni(plain) != 0;

    if (step == 0) {
      // The candidate goes to the end of the candidates (sparse_ldl.h:1441).
      __syncthreads();
      if (k < nc - 1) interchange(k, nc - 1, kw, s, s2);
      nc -= 1;
      wb ^= 1;
      publish(wb);
      __syncthreads();
      continue;
    }
    const int ksrc = k + step - 1;
    // A 2x2 pivot on rows k, k + 1 (no interchange): the next W row, row
    // k + 2, is published now (its earlier entries read above, this step's
    // two entries its own values), so the step ends with one barrier.
    const bool next2 = step == 2 && kpiv == k + 1;
    if (next2) {
      if (wlane >= 0 && wlane < kw && k + 2 < m) sh.wrow[wb ^ 1][wlane] = wn2;
      if (own && r == k + 2) {
        sh.wrow[wb ^ 1][kw] = s;
        sh.wrow[wb ^ 1][kw + 1] = s2;
      }
    }
    if (kpiv != ksrc) {
      __syncthreads();
      interchange(ksrc, kpiv, kw + step, s, s2);
      plain = false;
    }
    const int q = kw;
    if (step == 1) {
      const double d = kpiv == k ? sh.sval[wb][0] : sh.sval[wb][1];
      const double dinv = kpiv == k ? sh.sval[wb][3] : 1.0 / d;
      if (own && r >= k) {
        double v;
        if (r == k)
          v = d;
        else
          v = zero ? 0.0 : s * dinv;
        Ls[q * NTH + tid] = v;
      }
      if (tid == 0) count_inertia(1, d, 0.0, 0.0, zero, sh.cnt);
    } else {
      const double d11 = sh.sval[wb][0], d21 = sh.sval[wb][2], d22 = sh.sval[wb][1];
      const double detinv = 1.0 / (d11 * d22 - d21 * d21);
      const double i00 = d22 * detinv, i01 = -d21 * detinv, i11 = d11 * detinv;
      if (own && r >= k + 2) {
        Ls[q * NTH + tid] = s * i00 + s2 * i01;
        Ls[(q + 1) * NTH + tid] = s * i01 + s2 * i11;
      }
      if (tid == 0) {
        Ls[q * NTH + q] = d11;
        Ls[q * NTH + q + 1] = d21;
        Ls[(q + 1) * NTH + q + 1] = d22;
        vars[k] = -(vars[k] + 1);
        vars[k + 1] = -(vars[k + 1] + 1);
        count_inertia(2, d11, d21, d22, false, sh.cnt);
      }
    }
    k += step;
    kw += step;
    wb ^= 1;
    if (next2) {
      __syncthreads();
    } else if (!plain || step != 1) {
      // The next pivot row's W row was not published speculatively.
      __syncthreads();
      publish(wb);
      __syncthreads();
    }
  }
  __syncthreads();
  panel_store<NTH>(F, ld, m, kstart, pw, Ls);
}

// ---------------------------------------------------------------------------
// Blocked path for large fronts.  One panel: up to NB pivots selected with
// the reference's left-looking panel loop (sparse_ldl.h:1309-1505); W holds
// the unscaled updated columns (m x NBW, ld = m).  The reference's panel
// holds nblock - 1 = 63 pivots; the choice of panel width does not change
// any pivot decision.  One 1024-thread workgroup per front: each column
// update W[k:, kw] = F[k:, k] - F[k:, kstart:k] W[k, :kw]^T reads row k of W
// from LDS, and the candidate / contribution maxima come out of the same
// pass with one fused reduction.
template <bool REG, int NTH>
__global__ __launch_bounds__(REG ? NTH : NTP) void panel_kernel(DevState S,
                                                                const int* __restrict__ bnodes,
                                                                const int* __restrict__ broot,
                                                                double* __restrict__ Wbase,
                                                                const int64_t* __restrict__ woff,
                                                                double ustab,
                                                                double pivot_tol, int chol,
                                                                int lds_rows) {
  const int node = bnodes[blockIdx.x];
  const int root = broot[blockIdx.x];
  double* W = Wbase + woff[blockIdx.x];
  __shared__ VI red[2 * NTP / 64];
  __shared__ double wrow[NBW];
  __shared__ double sval[2];
  __shared__ PanelShared psh;
  __shared__ int bad;
  int* cnt = psh.cnt;
  const int tid = threadIdx.x;
  if (tid == 0) bad = fail_flag(S);
  __syncthreads();
  if (bad != 0) return;
  if (!node_on(S, node)) return;
  int* ps = S.pstate + PS * node;
  if (ps[4] != 1) return;  // factored in a workgroup of the level kernel
  const int m = uni(S.m[node]);
  double* F = S.rpool + S.foff[node];
  int* vars = S.ipool + S.voff[node];
  const int ld = m, ldw = m;
  int k = uni(ps[0]), nc = uni(ps[1]);
  if (tid == 0) cnt[0] = cnt[1] = cnt[2] = cnt[3] = 0;
  __syncthreads();
  if (k >= nc) {
    // Finished in an earlier step: make the following trailing launch a
    // no-op.  Only the first of the two instantiations launched per step
    // (REG) does this; the second may see a front the first just finished,
    // whose trailing update is still to come.
    if (REG && tid == 0) {
      ps[2] = k;
      ps[3] = 0;
    }
    return;
  }
  const int kstart = k;
  int kw = 0;
  const bool bk = root != 0;
  // Fronts whose panel rows fit one row per thread take the LDS path; the
  // two paths are separate instantiations so that each gets its own
  // register allocation, and each returns on the other's fronts.
  if ((m - kstart <= lds_rows) != REG) return;
  if constexpr (REG) {
    extern __shared__ double Ls[];
    panel_lds<NTH>(F, W, vars, m, k, nc, kw, kstart, bk, chol != 0, ustab, pivot_tol, Ls, psh);
  } else {
  while (kw < NB && k < nc) {
    double* wk = W + (size_t)kw * ldw;
    double* wi = W + (size_t)(kw + 1) * ldw;
    // Row k of W (the multipliers of this column's update) into LDS.
    if (tid < kw) wrow[tid] = W[k + (size_t)tid * ldw];
    __syncthreads();
    VI a = {-1.0, INT_MAX}, b = {-1.0, INT_MAX};
    for (int i = k + tid; i < m; i += NTP) {
      const double s = panel_dot(F[i + (size_t)k * ld], F + i + (size_t)kstart * ld,
                                 (size_t)ld, wrow, kw);
      wk[i] = s;
      if (i == k) sval[0] = s;
      if (i > k) {
        VI c = {fabs(s), i};
        if (i < nc)
          a = vi_best(a, c);
        else
          b = vi_best(b, c);
      }
    }
    int step = 1, kpiv = k;
    bool zero = false;
    if (chol) {
      __syncthreads();
    } else {
      block_argmax2<NTP>(a, b, red);
      const VI ca = a, cb = b;
      const double akk = fabs(sval[0]);
      const bool has_cand = k + 1 < nc && ca.i > k && ca.i < nc;
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
        // Column imax, updated: W[k:, kw+1]; its largest off-diagonal entry
        // comes out of the same pass.
        if (tid < kw) wrow[tid] = W[imax + (size_t)tid * ldw];
        __syncthreads();
        VI r = {-1.0, INT_MAX}, dummy = {-1.0, INT_MAX};
        for (int i = k + tid; i < m; i += NTP) {
          const double s0 = i < imax ? F[imax + (size_t)i * ld] : F[i + (size_t)imax * ld];
          const double s = panel_dot(s0, F + i + (size_t)kstart * ld, (size_t)ld, wrow, kw);
          wi[i] = s;
          if (i == imax)
            sval[1] = s;
          else
            r = vi_best(r, VI{fabs(s), i});
        }
        block_argmax2<NTP>(r, dummy, red);
        const double rowmax = r.v;
        const double aii = fabs(sval[1]);
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
          const double wkk = sval[0], wik = wk[imax], wii = sval[1];
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
          for (int i = k + tid; i < m; i += NTP) wk[i] = wi[i];
          __syncthreads();
        }
      } else {
        step = 0;
      }
    }
    step = uni(step);
    kpiv = uni(kpiv);
    zero = uni(zero) != 0;

    if (step == 0) {
      if (k < nc - 1) {
        symm_swap_t<NTP>(F, ld, m, vars, k, nc - 1);
        swap_w_rows_t<NTP>(W, ldw, k, nc - 1, kw);
      }
      nc -= 1;
      continue;
    }
    const int ksrc = k + step - 1;
    if (kpiv != ksrc) {
      symm_swap_t<NTP>(F, ld, m, vars, ksrc, kpiv);
      swap_w_rows_t<NTP>(W, ldw, ksrc, kpiv, kw + step);
    }
    if (step == 1) {
      const double d = wk[k];
      if (zero) {
        for (int i = k + tid; i < m; i += NTP) F[i + (size_t)k * ld] = 0.0;
      } else {
        const double dinv = 1.0 / d;
        for (int i = k + tid; i < m; i += NTP)
          F[i + (size_t)k * ld] = (i == k) ? wk[i] : wk[i] * dinv;
      }
      if (tid == 0) count_inertia(1, d, 0.0, 0.0, zero, cnt);
    } else {
      const double d11 = wk[k], d21 = wk[k + 1], d22 = wi[k + 1];
      const double detinv = 1.0 / (d11 * d22 - d21 * d21);
      const double i00 = d22 * detinv, i01 = -d21 * detinv, i11 = d11 * detinv;
      for (int i = k + 2 + tid; i < m; i += NTP) {
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
  }
  if (tid == 0) {
    ps[0] = k;
    ps[1] = nc;
    ps[2] = kstart;
    ps[3] = kw;
    if (k >= nc) {
      // Finished: pivot and delay counts for the parent and the solves.
      S.npiv[node] = nc;
      S.ndel[node] = ps[5] - nc;
    }
    for (int q = 0; q < 4; q++)
      if (cnt[q]) atomicAdd(&S.ninert[4 * node + q], cnt[q]);
  }
}

// Trailing update on MFMA (sparse_ldl.h:1172-1197):
// F[k:, k:] -= F[k:, kstart:kstart+kw] W[k:, :kw]^T on the lower triangle, 64x64 tiles staged through LDS in precision T
// (double: v_mfma_f64_16x16x4f64; float, the mixed-precision form:
// v_mfma_f32_16x16x4f32 with FP32 accumulation, result subtracted from the
// FP64 front).  Four waves per tile, each a 32x32 quadrant of 2x2 MFMA
// 16x16 blocks; K padded to a multiple of 4 with zeros.
// Operand layout of the 16x16x4 MFMAs: lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16].  Result register j of lane l holds
// D[4 (l / 16) + j][l % 16] for f32 and D[l / 16 + 4 j][l % 16] for f64
// (checked on gfx950).
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
typedef double mfma_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mfma_f4 mfma_16x16x4(float a, float b, mfma_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ mfma_d4 mfma_16x16x4(double a, double b, mfma_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int TK_PAD = ((NBW + 3) / 4) * 4;  // padded K of a panel

template <typename T, typename V4>
__global__ __launch_bounds__(NT) void trailing_mfma_kernel(DevState S,
                                                           const int* __restrict__ bnodes,
                                                           const double* __restrict__ Wbase,
                                                           const int64_t* __restrict__ woff) {
  const int node = bnodes[blockIdx.z];
  const double* W = Wbase + woff[blockIdx.z];
  __shared__ T As[TILE][TK_PAD + 1];
  __shared__ T Bs[TILE][TK_PAD + 1];
  __shared__ int bad;
  if (threadIdx.x == 0) bad = fail_flag(S);
  __syncthreads();
  if (bad != 0) return;
  if (!node_on(S, node)) return;
  const int* ps = S.pstate + PS * node;
  const int m = uni(S.m[node]);
  const int k = uni(ps[0]), kstart = uni(ps[2]), kw = uni(ps[3]);
  const int ti = blockIdx.x, tj = blockIdx.y;
  if (kw == 0 || ti < tj) return;
  const int i0 = k + ti * TILE, j0 = k + tj * TILE;
  if (i0 >= m || j0 >= m) return;
  double* F = S.rpool + S.foff[node];
  const int ld = m, ldw = m;
  const int tid = threadIdx.x;
  const int kp = (kw + 3) & ~3;
  const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // scalar
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  // This thread's 16 entries of the tile, read first so that their latency
  // overlaps the operand loads (the tile lies in columns >= k, the operands
  // in the panel's columns and in W).
  double old[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int jj = 0; jj < 4; jj++) {
        const int i = i0 + wr + a * 16 + (sizeof(T) == 8 ? lk + 4 * jj : lk * 4 + jj);
        const int j = j0 + wc + b * 16 + lr;
        old[a][b][jj] = (i < m && j < m && i >= j) ? F[i + (size_t)j * ld] : 0.0;
      }
  // Operand tiles into LDS, 2 x 8 loads per thread in flight.
  const int tot = TILE * kp;
  for (int e0 = tid; e0 < tot; e0 += 8 * NT) {
    double av[8], bv[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int e = e0 + u * NT, r = e % TILE, q = e / TILE;
      const int i = i0 + r, j = j0 + r;
      const bool ok = e < tot && q < kw;
      av[u] = (ok && i < m) ? F[i + (size_t)(kstart + q) * ld] : 0.0;
      bv[u] = (ok && j < m) ? W[j + (size_t)q * ldw] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int e = e0 + u * NT, r = e % TILE, q = e / TILE;
      if (e < tot) {
        As[r][q] = (T)av[u];
        Bs[r][q] = (T)bv[u];
      }
    }
  }
  __syncthreads();
  V4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = V4{0, 0, 0, 0};
  for (int q = 0; q < kp; q += 4) {
    T av[2], bv[2];
#pragma unroll
    for (int a = 0; a < 2; a++) av[a] = As[wr + a * 16 + lr][q + lk];
#pragma unroll
    for (int b = 0; b < 2; b++) bv[b] = Bs[wc + b * 16 + lr][q + lk];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) acc[a][b] = mfma_16x16x4(av[a], bv[b], acc[a][b]);
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int jj = 0; jj < 4; jj++) {
        const int i = i0 + wr + a * 16 + (sizeof(T) == 8 ? lk + 4 * jj : lk * 4 + jj);
        const int j = j0 + wc + b * 16 + lr;
        if (i < m && j < m && i >= j) F[i + (size_t)j * ld] = old[a][b][jj] - (double)acc[a][b][jj];
      }
}

// flags[1..4] = inertia counts (+, -, 0, non-finite), flags[5] = delayed
// pivots, flags[6] = largest front, over the fronts this factorization
// took part in, from the per-node counts (integer sums: exact in any order).
__global__ __launch_bounds__(NT) void factor_stats_kernel(DevState S, int K) {
  __shared__ int red[6][NT / 64];
  if (fail_flag(S) != 0) return;
  int c[6] = {0, 0, 0, 0, 0, 0};
  for (int node = blockIdx.x * NT + threadIdx.x; node < K; node += gridDim.x * NT) {
    if (!node_on(S, node)) continue;
    const int* ps = S.pstate + PS * node;
    if (ps[4] == 1 && ps[0] < ps[1]) atomicExch(&S.flags[0], -7);  // a root left unfinished
    for (int q = 0; q < 4; q++) c[q] += S.ninert[4 * node + q];
    c[4] += S.ndel[node];
    c[5] = max(c[5], S.m[node]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < 5; q++) c[q] += __shfl_xor(c[q], off, 64);
    c[5] = max(c[5], __shfl_xor(c[5], off, 64));
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int q = 0; q < 6; q++) red[q][w] = c[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int ww = 1; ww < NT / 64; ww++) {
      for (int q = 0; q < 5; q++) c[q] += red[q][ww];
      c[5] = max(c[5], red[5][ww]);
    }
    for (int q = 0; q < 4; q++)
      if (c[q]) atomicAdd(&S.flags[1 + q], c[q]);
    if (c[4]) atomicAdd(&S.flags[5], c[4]);
    atomicMax(&S.flags[6], c[5]);
  }
}

// Per-block inertia of a block-diagonal matrix: block b owns nodes
// seg[b]..seg[b+1]; one thread per block sums its nodes' pivot signs in
// node order (integer sums: exact).
// Inertia counts of each block of a batch: workgroup b sums the per-node
// counts of block b (integer sums: any order gives the same result).
__global__ __launch_bounds__(256) void batch_inertia_kernel(int nb, const int* __restrict__ seg,
                                                            const int* __restrict__ ninert,
                                                            int* __restrict__ out) {
  __shared__ int red[4][256 / 64];
  const int b = blockIdx.x, t = threadIdx.x;
  if (b >= nb) return;
  int c[4] = {0, 0, 0, 0};
  for (int k = seg[b] + t; k < seg[b + 1]; k += 256)
#pragma unroll
    for (int q = 0; q < 4; q++) c[q] += ninert[4 * k + q];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    int v = c[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((t & 63) == 0) red[q][t >> 6] = v;
  }
  __syncthreads();
  if (t < 4) out[4 * b + t] = red[t][0] + red[t][1] + red[t][2] + red[t][3];
}

// Symmetric equilibration by Ruiz iteration: each pass replaces
// s_i <- s_i / sqrt(max_j |s_i a_ij s_j|) over row i of the shifted matrix
// (Jacobi style: sin -> sout), so that every row and column maximum of the
// scaled matrix tends to 1.  One wave per row.  (The reference factors the
// unscaled matrix; the scaling leaves the inertia unchanged and the solves
// unscale.)
__global__ void equilibrate_kernel(int n, const int* __restrict__ rowp,
                                   const int* __restrict__ cols,
                                   const double* __restrict__ vals,
                                   const double* __restrict__ shift,
                                   const double* __restrict__ sin,
                                   double* __restrict__ sout,
                                   const int* __restrict__ row_blk,
                                   const int* __restrict__ blk_active) {
  // One thread per row: KKT rows hold a few entries, and a maximum is exact
  // in any order.
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    // Rows of blocks left out keep their scale factors.
    if (row_blk && !blk_active[row_blk[r]]) continue;
    double m = 0.0;
    const double sr = sin ? sin[r] : 1.0;
    for (int p = rowp[r]; p < rowp[r + 1]; p++) {
      const int c = cols[p];
      double v = vals[p];
      if (c == r && shift) v += shift[r];
      m = fmax(m, fabs(v * sr * (sin ? sin[c] : 1.0)));
    }
    sout[r] = m > 0.0 && m < INFINITY ? sr / sqrt(m) : sr;
  }
}

__global__ void scale_vec_kernel(int n, int nrhs, int ldx, const double* __restrict__ sc,
                                 double* __restrict__ x, const int* __restrict__ row_blk,
                                 const int* __restrict__ blk_active) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < (size_t)n * nrhs;
       t += (size_t)gridDim.x * blockDim.x) {
    const int i = int(t % n), r = int(t / n);
    if (row_blk && !blk_active[row_blk[i]]) continue;
    x[i + (size_t)r * ldx] *= sc[i];
  }
}

__global__ void reset_kernel(DevState S) {
  if (S.node_blk) {
    // Only the blocks that take part start their regions afresh.
    for (int b = threadIdx.x; b < S.nblk; b += blockDim.x)
      if (S.blk_active[b]) {
        S.tops[2 * b] = 0;
        S.tops[2 * b + 1] = 0;
      }
  } else if (threadIdx.x == 0) {
    S.tops[0] = 0;
    S.tops[1] = 0;
  }
  if (threadIdx.x < 8) S.flags[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// Pivoting permutes the fully summed rows of a front; the forward solve maps
// the children's update rows (laid out in assembly order: supernode
// variables, then the children's delayed pivots in child order) to their
// final positions through fpos.  One workgroup per front.
__global__ __launch_bounds__(NT) void fpos_kernel(DevPlan P, DevState S, int K) {
  // After a failed (to be redone) factorization the fronts are not all
  // allocated: nothing to map.
  if (fail_flag(S) != 0) return;
  for (int node = blockIdx.x; node < K; node += gridDim.x) {
    if (!node_on(S, node)) continue;
    const int ns = P.ns[node], D = S.D[node], nfs = ns + D;
    const int* vars = S.ipool + S.voff[node];
    int* fpos = S.fpos + S.voff[node];
    const int* sv = P.nvars + P.nptr[node];
    const int kb = P.kptr[node], ke = P.kptr[node + 1];
    // A blocked child that its panel launches did not finish (the root
    // levels have no parent level kernel to see it): redo, map nothing.
    bool kids_done = true;
    for (int t = kb; t < ke; t++) {
      const int* pc = S.pstate + PS * P.kids[t];
      if (pc[4] == 1 && pc[0] < pc[1]) kids_done = false;
    }
    if (!kids_done) {
      if (threadIdx.x == 0) atomicExch(&S.flags[0], -7);
      continue;
    }
    for (int j = threadIdx.x; j < nfs; j += NT) {
      int v = vars[j];
      v = v >= 0 ? v : -v - 1;
      int a = -1;
      for (int q = 0; q < ns && a < 0; q++)
        if (sv[q] == v) a = q;
      int doff = ns;
      for (int t = kb; t < ke && a < 0; t++) {
        const int c = P.kids[t];
        const int pc = S.npiv[c], dc = S.ndel[c];
        const int* cv = S.ipool + S.voff[c] + pc;
        for (int q = 0; q < dc && a < 0; q++)
          if (cv[q] == v) a = doff + q;
        doff += dc;
      }
      if (a >= 0) fpos[a] = j;
    }
  }
}

// ---------------------------------------------------------------------------
// Wave-synchronous pivot-block solves (p <= 64 pivots): lane l holds t[l],
// the p x p pivot block L11 (with the D blocks on its diagonal) is staged in
// LDS as a packed lower triangle, and each pivot step broadcasts one (two)
// value(s) with readlane instead of a workgroup barrier.
__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Pivot blocks L[0:p, 0:p] (p <= 64) are staged in LDS as packed lower
// triangles: column j starts at tri_off(j, p).
constexpr int PIVOT_TRI = 64 * 65 / 2;
// Forward-solve workgroups: NT threads, or NTS on levels of at most
// SOLVE_WIDE_FRONTS fronts with one of at least SOLVE_WIDE_ROWS rows (the
// arithmetic of a front does not depend on the width).
constexpr int NTS = 1024;
constexpr int SOLVE_WIDE_FRONTS = 256;
constexpr int SOLVE_WIDE_ROWS = 256;
__device__ __forceinline__ int tri_off(int j, int p) { return j * p - ((j * (j - 1)) >> 1); }

// All threads; 16 loads per thread are in flight before any is stored.
// Ends with a barrier when SYNC (otherwise the caller's next barrier
// publishes the block).
template <int NTH, bool SYNC = true>
__device__ __forceinline__ void stage_pivot_block(const double* __restrict__ L, int ld, int p,
                                                  double* Ls) {
  const int tot = p * p;
  for (int e0 = threadIdx.x; e0 < tot; e0 += 16 * NTH) {
    double a[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int e = e0 + u * NTH, j = e / p, i = e - j * p;
      a[u] = (e < tot && i >= j) ? L[i + (size_t)j * ld] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int e = e0 + u * NTH, j = e / p, i = e - j * p;
      if (e < tot && i >= j) Ls[tri_off(j, p) + i - j] = a[u];
    }
  }
  if (SYNC) __syncthreads();
}

// Forward: unit-lower solve of t[0:p) (in place, lane l <- t[l]) and the
// roles of the lanes (0: 1x1, 1: first of a 2x2, 2: second).  Wave 0 only.
__device__ __forceinline__ double wave_lower_solve(double tl, int l, int p, const int* vars,
                                                   const double* Ls, int& role) {
  const int vl = l < p ? vars[l] : 0;
  role = 0;
  for (int j = 0; j < p;) {
    const bool two = __builtin_amdgcn_readlane(vl, j) < 0;
    const int cj = tri_off(j, p) - j;  // Ls[cj + i] = L[i, j]
    if (!two) {
      const double tj = readlane_d(tl, j);
      if (l > j && l < p) tl -= Ls[cj + l] * tj;
      j += 1;
    } else {
      const double t1 = readlane_d(tl, j), t2 = readlane_d(tl, j + 1);
      const int cj1 = tri_off(j + 1, p) - (j + 1);
      if (l > j + 1 && l < p) tl -= Ls[cj + l] * t1 + Ls[cj1 + l] * t2;
      if (l == j) role = 1;
      if (l == j + 1) role = 2;
      j += 2;
    }
  }
  return tl;
}

// D solve of the lane's pivot (1x1, or its half of a 2x2 block).
__device__ __forceinline__ double wave_d_solve(double tl, int l, int p, int role,
                                               const double* Ls) {
  const double up = __shfl(tl, min(l + 1, 63), 64);
  const double dn = __shfl(tl, max(l - 1, 0), 64);
  if (l >= p) return tl;
  if (role == 0) return tl / Ls[tri_off(l, p)];
  if (role == 1) {
    const int c = tri_off(l, p);
    const double d11 = Ls[c], d21 = Ls[c + 1], d22 = Ls[tri_off(l + 1, p)];
    const double inv = 1.0 / (d11 * d22 - d21 * d21);
    return inv * (d22 * tl - d21 * up);
  }
  const int c = tri_off(l - 1, p);
  const double d11 = Ls[c], d21 = Ls[c + 1], d22 = Ls[tri_off(l, p)];
  const double inv = 1.0 / (d11 * d22 - d21 * d21);
  return inv * (d11 * tl - d21 * dn);
}

// Backward: unit-upper (L11^T) solve of t[0:p), wave 0 only.
__device__ __forceinline__ double wave_upper_solve(double tl, int l, int p, const int* vars,
                                                   const double* Ls) {
  const int vl = l < p ? vars[l] : 0;
  const int cl = tri_off(l, p) - l;  // Ls[cl + i] = L[i, l]
  for (int i = p - 1; i >= 0;) {
    if (__builtin_amdgcn_readlane(vl, i) >= 0) {
      const double ti = readlane_d(tl, i);
      if (l < i) tl -= Ls[cl + i] * ti;
      i -= 1;
    } else {
      const double t2 = readlane_d(tl, i), t1 = readlane_d(tl, i - 1);
      if (l < i - 1) tl -= Ls[cl + i] * t2 + Ls[cl + i - 1] * t1;
      i -= 2;
    }
  }
  return tl;
}

// Front update of the forward solve: uo[i] = t[i] - sum_{j<p} L[i, j] t[j]
// for the rows i in [p, m).  A thread takes one row and one of G column
// groups (columns g, g + G, ...), with 16 loads in flight; the G partial sums
// of a row are added in group order, so the result is deterministic.  red
// holds NT doubles of LDS.  All threads; R = m - p is uniform.
template <int NTH>
__device__ __forceinline__ void front_update(const double* __restrict__ L, int ld, int p, int m,
                                             const double* t, double* __restrict__ uo,
                                             double* red) {
  constexpr int CH = 16;  // loads in flight per thread
  const int tid = threadIdx.x, R = m - p;
  if (R <= 0) return;
  if (R < NTH) {
    // G from NT, not NTH: a row's partial sums (and so its rounding) do not
    // depend on the workgroup size, so a front solves to the same bits
    // whichever width its level is launched with (alone or in a batch).
    const int G = max(1, min(NT / R, p));
    double s = 0.0;
    if (tid < G * R) {
      const int g = tid / R, i = p + tid - g * R;
      const double* Li = L + i;
      for (int j0 = g; j0 < p; j0 += CH * G) {
        double a[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
          const int j = j0 + u * G;
          a[u] = j < p ? Li[(size_t)j * ld] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
          const int j = j0 + u * G;
          if (j < p) s += a[u] * t[j];
        }
      }
    }
    red[tid] = s;
    __syncthreads();
    if (tid < R) {
      double a = t[p + tid];
      for (int g = 0; g < G; g++) a -= red[g * R + tid];
      uo[p + tid] = a;
    }
  } else {
    for (int i = p + tid; i < m; i += NTH) {
      const double* Li = L + i;
      double s = 0.0;
      for (int j0 = 0; j0 < p; j0 += CH) {
        double a[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) a[u] = j0 + u < p ? Li[(size_t)(j0 + u) * ld] : 0.0;
#pragma unroll
        for (int u = 0; u < CH; u++)
          if (j0 + u < p) s += a[u] * t[j0 + u];
      }
      uo[i] = t[i] - s;
    }
  }
}

// Sum of v[u] over the 64 lanes for 16 values at once: four exchange steps
// halve the values a lane holds (lane bits 5..2 select which one it keeps),
// two more add over lane bits 1..0.  Lane l with (l & 3) == 0 returns the sum
// of value ((l >> 2) & 15) read as bits 5..2 of l: index 8*b5 + 4*b4 + 2*b3 + b2.
__device__ __forceinline__ double wave_reduce16(double (&v)[16], int lane) {
#pragma unroll
  for (int step = 0; step < 4; step++) {
    const int half = 8 >> step, off = 32 >> step;
    const bool hi = (lane & off) != 0;
#pragma unroll
    for (int u = 0; u < half; u++) {
      double lo = v[u], up = v[u + half];
      // Opaque copies: without them the two selects fold into a lane-varying
      // index into v[], which is then lowered as a chain of selects over
      // all 16 values.
      asm volatile("" : "+v"(lo), "+v"(up));
      const double keep = hi ? up : lo;
      const double give = hi ? lo : up;
      v[u] = keep + __shfl_xor(give, off, 64);
    }
  }
  double s = v[0];
  s += __shfl_xor(s, 2, 64);
  s += __shfl_xor(s, 1, 64);
  return s;
}

// Front update of the backward solve: t[j] -= sum_{i in [p,m)} L[i, j] t[i]
// for j < p.  Wave w takes columns jb + w + NW u (u < 16) of each block of
// 16 NW columns and its lanes run over all the rows 64 at a time, with 16
// loads in flight.  All threads; ends with a barrier.
template <int NTH>
__device__ __forceinline__ void front_update_t(const double* __restrict__ L, int ld, int p,
                                               int m, double* t) {
  constexpr int NW = NTH / 64;
  const int lane = threadIdx.x & 63;
  // The wave index is uniform: the column addresses stay in scalar registers.
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int jb = 0; jb < p; jb += 16 * NW) {
    if (jb + w >= p) continue;  // no column of this block for this wave
    double s[16];
#pragma unroll
    for (int u = 0; u < 16; u++) s[u] = 0.0;
    // Columns past p read column p - 1 (in bounds) and are dropped below.
    const double* Lc = L + (size_t)(jb + w) * ld;
    const int ju = (p - 1 - jb - w) / NW;  // last u with a column below p
    for (unsigned i = p + lane; i < (unsigned)m; i += 64) {
      const double ti = t[i];
      double a[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        // Uniform column base (scalar registers) plus the lane's row.
        const double* cu = Lc + __builtin_amdgcn_readfirstlane(min(u, ju) * NW * ld);
        a[u] = cu[i];
      }
#pragma unroll
      for (int u = 0; u < 16; u++) s[u] += a[u] * ti;
    }
    const double tot = wave_reduce16(s, lane);
    if ((lane & 3) == 0) {
      const int u = 8 * ((lane >> 5) & 1) + 4 * ((lane >> 4) & 1) + 2 * ((lane >> 3) & 1) +
                    ((lane >> 2) & 1);
      const int j = jb + w + NW * u;
      if (j < p) t[j] -= tot;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Level-scheduled solves (sparse_ldl.h:1807-1908).  Forward, multifrontal
// form: the front vector t holds the right-hand side on the pivot rows plus
// the children's update vectors (extend-added in child order, so the sums
// are deterministic and no atomics are needed); the unit-lower solve on the
// pivots follows, the front's own update vector u = t[p:m] - L21 t[0:p] is
// stored for the parent, then t = D^-1 t goes to x.  Update vectors live in
// ub at the front's index-pool offset (ld of one right-hand side = ubld).
// Dynamic LDS: t (tcap >= m doubles), the packed pivot block, NT doubles.
template <int NTH>
__global__ __launch_bounds__(NTH) void solve_fwd_kernel(DevPlan P, DevState S,
                                                       const int* __restrict__ lnodes,
                                                       double* __restrict__ x,
                                                       int ldx, int nrhs,
                                                       double* __restrict__ ub,
                                                       size_t ubld, int tcap) {
  extern __shared__ double smem[];
  double* t = smem;
  double* Ls = smem + tcap;
  double* red = Ls + PIVOT_TRI;
  __shared__ int kmeta[4][NT];  // per child: delayed, update rows, offset, cmap start
  const int node = lnodes[blockIdx.x];
  if (!node_on(S, node)) return;
  const int p = S.npiv[node], m = S.m[node];
  const double* L = S.rpool + S.foff[node];
  const int* vars = S.ipool + S.voff[node];
  const int ld = m, tid = threadIdx.x;
  const int ns = P.ns[node], D = S.D[node];
  const int kb = P.kptr[node], ke = P.kptr[node + 1];
  const int* fpos = S.fpos + S.voff[node];
  const bool wave = p <= 64;
  // The pivot block, the children's metadata (when one batch holds them
  // all) and the right-hand side are read together, before one barrier.
  const bool one_batch = ke - kb <= NT;
  auto load_kmeta = [&](int q0, int nb) {
    if (tid < nb) {
      const int c = P.kids[q0 + tid];
      const int pc = S.npiv[c];
      kmeta[0][tid] = S.ndel[c];
      kmeta[1][tid] = S.m[c] - pc;
      kmeta[2][tid] = S.voff[c] + pc;
      kmeta[3][tid] = P.cptr[c];
    }
  };
  if (wave) stage_pivot_block<NTH, false>(L, ld, p, Ls);
  if (one_batch) load_kmeta(kb, ke - kb);
  for (int r = 0; r < nrhs; r++) {
    double* xr = x + (size_t)r * ldx;
    double* ur = ub + (size_t)r * ubld;
    for (int j = tid; j < m; j += NTH) {
      double v = 0.0;
      if (j < p) {
        const int q = vars[j];
        v = xr[q >= 0 ? q : -q - 1];
      }
      t[j] = v;
    }
    __syncthreads();
    int doff = ns;
    for (int q0 = kb; q0 < ke; q0 += NT) {  // children in batches of NT
      const int nb = min(NT, ke - q0);
      if (!one_batch) {
        load_kmeta(q0, nb);
        __syncthreads();
      }
      for (int qq = 0; qq < nb; qq++) {
        const int dc = kmeta[0][qq], wc = kmeta[1][qq];
        const double* uc = ur + kmeta[2][qq];
        const int* cm = P.cmap + kmeta[3][qq] - dc;  // cm[ii] for ii >= dc
        for (int ii = tid; ii < wc; ii += NTH) {
          int pos;
          if (ii < dc) {
            pos = doff + ii;
          } else {
            const int s = cm[ii];
            pos = s < ns ? s : s + D;
          }
          if (pos < ns + D) pos = fpos[pos];
          t[pos] += uc[ii];
        }
        doff += kmeta[0][qq];
        __syncthreads();
      }
    }
    double* uo = ur + S.voff[node];
    if (wave) {
      double tl = 0.0;
      int role = 0;
      if (tid < 64) {
        tl = wave_lower_solve(tid < p ? t[tid] : 0.0, tid, p, vars, Ls, role);
        if (tid < p) t[tid] = tl;
      }
      __syncthreads();
      front_update<NTH>(L, ld, p, m, t, uo, red);
      if (tid < 64) {
        tl = wave_d_solve(tl, tid, p, role, Ls);
        if (tid < p) {
          const int v = vars[tid];
          xr[v >= 0 ? v : -v - 1] = tl;
        }
      }
      __syncthreads();
      continue;
    }
    for (int j = 0; j < p;) {
      if (vars[j] >= 0) {
        const double tj = t[j];
        for (int i = j + 1 + tid; i < p; i += NTH) t[i] -= L[i + (size_t)j * ld] * tj;
        j += 1;
      } else {
        const double t1 = t[j], t2 = t[j + 1];
        for (int i = j + 2 + tid; i < p; i += NTH)
          t[i] -= L[i + (size_t)j * ld] * t1 + L[i + (size_t)(j + 1) * ld] * t2;
        j += 2;
      }
      __syncthreads();
    }
    front_update<NTH>(L, ld, p, m, t, uo, red);
    __syncthreads();  // t[0:p) is read above and rescaled below
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
    for (int j = tid; j < p; j += NTH) {
      int v = vars[j];
      xr[v >= 0 ? v : -v - 1] = t[j];
    }
    __syncthreads();
  }
}

// Backward: t = x[piv]; t -= L21^T x[rows below]; L11^T solve; x[piv] = t.
// Dynamic LDS: t (tcap >= m doubles) then the packed pivot block.
template <int NTH>
__global__ __launch_bounds__(NTH) void solve_bwd_kernel(DevState S,
                                                       const int* __restrict__ lnodes,
                                                       double* __restrict__ x,
                                                       int ldx, int nrhs, int tcap) {
  extern __shared__ double smem[];
  double* t = smem;  // pivots then the other rows
  double* Ls = smem + tcap;
  const int node = lnodes[blockIdx.x];
  if (!node_on(S, node)) return;
  const int p = S.npiv[node], m = S.m[node];
  if (p == 0) return;
  const double* L = S.rpool + S.foff[node];
  const int* vars = S.ipool + S.voff[node];
  const int ld = m, tid = threadIdx.x;
  const bool wave = p <= 64;
  // The pivot block and the right-hand side are read before one barrier.
  if (wave) stage_pivot_block<NTH, false>(L, ld, p, Ls);
  for (int r = 0; r < nrhs; r++) {
    double* xr = x + (size_t)r * ldx;
    for (int j = tid; j < m; j += NTH) {
      int v = vars[j];
      t[j] = xr[v >= 0 ? v : -v - 1];
    }
    __syncthreads();
    front_update_t<NTH>(L, ld, p, m, t);
    if (wave) {
      if (tid < 64) {
        const double tl = wave_upper_solve(tid < p ? t[tid] : 0.0, tid, p, vars, Ls);
        if (tid < p) {
          const int v = vars[tid];
          xr[v >= 0 ? v : -v - 1] = tl;
        }
      }
      __syncthreads();
      continue;
    }
    for (int i = p - 1; i >= 0;) {
      if (vars[i] >= 0) {
        const double ti = t[i];
        for (int q = tid; q < i; q += NTH) t[q] -= L[i + (size_t)q * ld] * ti;
        i -= 1;
      } else {
        const double t2 = t[i], t1 = t[i - 1];
        for (int q = tid; q < i - 1; q += NTH)
          t[q] -= L[i + (size_t)q * ld] * t2 + L[i - 1 + (size_t)q * ld] * t1;
        i -= 2;
      }
      __syncthreads();
    }
    for (int j = tid; j < p; j += NTH) {
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
  drop_graphs();
  if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
  for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
  for (void* p : allocs_) (void)hipFree(p);
  if (S_.rpool) (void)hipFree(S_.rpool);
  if (S_.ipool) (void)hipFree(S_.ipool);
  if (S_.fpos) (void)hipFree(S_.fpos);
  if (W_) (void)hipFree(W_);
  if (ub_) (void)hipFree(ub_);
  if (d_batch_seg_) (void)hipFree(d_batch_seg_);
  if (d_batch_inert_) (void)hipFree(d_batch_inert_);
  if (d_node_blk_) (void)hipFree(d_node_blk_);
  if (d_blk_active_) (void)hipFree(d_blk_active_);
  if (d_blk_layout_) (void)hipFree(d_blk_layout_);
  if (d_tops_batch_) (void)hipFree(d_tops_batch_);
  if (h_flags_pin_) (void)hipHostFree(h_flags_pin_);
  if (h_binert_pin_) (void)hipHostFree(h_binert_pin_);
  if (h_stage_pin_) (void)hipHostFree(h_stage_pin_);
  for (void* q : {(void*)d_m_level_nodes_, (void*)d_m_big_nodes_, (void*)d_m_big_root_,
                  (void*)d_m_big_woff_, (void*)d_row_blk_,
                  (void*)d_node_cap_})
    if (q) (void)hipFree(q);
}

void LdlNumeric::tk_begin(hipStream_t s) {
  if (!timing_) return;
  if (ev_used_ + 2 > (int)ev_.size()) {
    for (int i = 0; i < 512; i++) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      ev_.push_back(e);
      ev_class_.push_back(-1);
    }
  }
  (void)hipEventRecord(ev_[ev_used_], s);
}

void LdlNumeric::tk_end(hipStream_t s, int cls) {
  if (!timing_ || ev_used_ + 2 > (int)ev_.size()) return;
  (void)hipEventRecord(ev_[ev_used_ + 1], s);
  ev_class_[ev_used_] = cls;
  ev_used_ += 2;
}

// Called after the stream has been synchronized.
void LdlNumeric::tk_collect() {
  for (int i = 0; i + 1 < ev_used_; i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev_[i], ev_[i + 1]) == hipSuccess) {
      tk_ms_[ev_class_[i]] += ms;
      tk_n_[ev_class_[i]] += 1;
    }
  }
  ev_used_ = 0;
}

int LdlNumeric::setup(const SymbolicPlan& plan, const int* rowp,
                      const int* cols, double delay_growth, std::string& err) {
  n_ = plan.n;
  K_ = plan.num_nodes;
  delay_growth_ = delay_growth;
  h_nptr_ = plan.node_ptr;
  h_nvars_ = plan.node_to_var;
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
  // Original entries of each supernode column j (all on or below the
  // diagonal, so their front column is j), column by column: acstart gives
  // each column's first entry.
  std::vector<int> aptr(K + 1, 0), aidx, arow, acol, acstart(n_ + 1, 0);
  for (int k = 0; k < K; k++) {
    for (int j = 0; j < cs[k]; j++) cpos[plan.contrib_rows[plan.contrib_ptr[k] + j]] = j;
    for (int j = 0; j < ns[k]; j++) {
      acstart[plan.node_ptr[k] + j] = int(aidx.size());
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
  acstart[n_] = int(aidx.size());

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
  // Inside a level, fronts whose workgroup LDS fits SMALL_LDS_M^2 come
  // first and are launched with that smaller LDS size (several workgroups
  // per CU); the others follow in a second launch.
  auto lds_side = [&](int k) { return std::min(ns[k] + cs[k] + 32, LDS_MAX_M); };
  level_split_.assign(nlev, 0);
  level_lds_small_.assign(nlev, 0);
  for (int l = 0; l < nlev; l++) {
    auto b = level_nodes_.begin() + level_ptr_[l], e = level_nodes_.begin() + level_ptr_[l + 1];
    auto mid = std::stable_partition(b, e, [&](int k) { return lds_side(k) <= SMALL_LDS_M; });
    level_split_[l] = int(mid - level_nodes_.begin());
    int mx = 0;
    for (auto it = b; it != mid; ++it) mx = std::max(mx, lds_side(*it));
    level_lds_small_[l] = mx * mx;
  }
  std::vector<int> big(K, 0);
  max_front_ = 0;
  double rneed = 0.0, ineed = 0.0;
  flops_ = 0.0;
  flops_small_ = flops_big_ = 0.0;
  solve_bytes_ = 0.0;
  for (int k = 0; k < K; k++) {
    int m0 = ns[k] + cs[k];
    max_front_ = std::max(max_front_, m0);
    big[k] = m0 + DPLAN_BIG > LDS_MAX_M ? 1 : 0;
    rneed += double(m0) * m0;
    ineed += m0;
    double fk = 0.0;
    for (int q = 0; q < ns[k]; q++) {
      double r = m0 - q - 1;
      fk += r + r * (m0 - q);
    }
    flops_ += fk;
    if (m0 > LDS_MAX_M)
      flops_big_ += fk;
    else
      flops_small_ += fk;
    solve_bytes_ += 8.0 * 2.0 * double(m0) * ns[k] + 8.0 * 2.0 * m0 + 4.0 * m0;
  }
  level_lds_.assign(nlev, 0);
  big_ptr_.assign(nlev + 1, 0);
  big_nodes_.clear();
  big_npanels_.assign(nlev, 0);
  big_tiles_.assign(nlev, 0);
  big_wide_.assign(nlev, 0);
  big_mmax_.assign(nlev, 0);
  h_kptr_ = plan.children_ptr;
  std::vector<int> big_root;
  std::vector<int64_t> big_woff;
  int64_t wmax = 0;
  big_nth_.assign(nlev, NTP);
  solve_wide_.assign(nlev, 0);
  for (int l = 0; l < nlev; l++) {
    int mx = 0;
    int big_mcap = 0;
    int64_t wtot = 0;
    for (int t = level_ptr_[l]; t < level_ptr_[l + 1]; t++) {
      int k = level_nodes_[t];
      if (big[k]) {
        big_nodes_.push_back(k);
        big_root.push_back(cs[k] == 0 ? 1 : 0);
        const int mcap = ns[k] + cs[k] + DCAP_BIG;
        big_woff.push_back(wtot);
        wtot += (int64_t)mcap * NBW;
        big_npanels_[l] = std::max(big_npanels_[l], (ns[k] + DPAN_BIG + NBR - 1) / NBR);
        if (mcap > NTP) big_wide_[l] = 1;
        big_mcap = std::max(big_mcap, mcap);
        big_mmax_[l] = std::max(big_mmax_[l], mcap);
        big_tiles_[l] = std::max(big_tiles_[l], (mcap + TILE - 1) / TILE);
      }
      if (t >= level_split_[l]) mx = std::max(mx, std::min(ns[k] + cs[k] + 32, LDS_MAX_M));
    }
    level_lds_[l] = mx * mx;
    big_nth_[l] = big_mcap <= NTP / 4 ? NTP / 4 : big_mcap <= NTP / 2 ? NTP / 2 : NTP;
    // Solves: levels of few tall fronts take 1,024-thread workgroups (a
    // front's update reads then keep four times the loads in flight).
    int smax = 0;
    for (int t = level_ptr_[l]; t < level_ptr_[l + 1]; t++)
      smax = std::max(smax, ns[level_nodes_[t]] + cs[level_nodes_[t]]);
    solve_wide_[l] = (level_ptr_[l + 1] - level_ptr_[l] <= SOLVE_WIDE_FRONTS &&
                      smax >= SOLVE_WIDE_ROWS) ? 1 : 0;
    big_ptr_[l + 1] = int(big_nodes_.size());
    wmax = std::max(wmax, wtot);
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
  if (upload(allocs_, acstart, &P_.acstart, err)) return -2;
  {
    std::vector<int> rp(rowp, rowp + n + 1), cl(cols, cols + rowp[n]);
    if (upload(allocs_, rp, &d_rowp_, err)) return -2;
    if (upload(allocs_, cl, &d_cols_, err)) return -2;
    if (dalloc(allocs_, std::max(n, 1), &scale_, err)) return -2;
    if (dalloc(allocs_, std::max(n, 1), &scale_tmp_, err)) return -2;
  }
  P_.n = n;
  P_.K = K;
  const int* dl = nullptr;
  if (upload(allocs_, level_nodes_, &dl, err)) return -2;
  d_level_nodes_ = const_cast<int*>(dl);
  const int* db = nullptr;
  if (upload(allocs_, big, &db, err)) return -2;
  d_big_mask_ = const_cast<int*>(db);
  {
    const int* t1 = nullptr;
    const int* t2 = nullptr;
    const int64_t* t4 = nullptr;
    h_big_root_ = big_root;
    h_big_woff_ = big_woff;
    if (upload(allocs_, big_nodes_, &t1, err)) return -2;
    if (upload(allocs_, big_root, &t2, err)) return -2;
    if (upload(allocs_, big_woff, &t4, err)) return -2;
    d_big_nodes_ = t1;
    d_big_root_ = t2;
    d_big_woff_ = t4;
    if (dalloc(allocs_, std::max<size_t>(1, (size_t)K), &d_big_woff_sync_, err)) return -2;
    if (wmax > 0) {
      HIPCHK(hipMalloc(&W_, wmax * sizeof(double)));
      W_cap_ = wmax;
    }
  }

  if (dalloc(allocs_, K, &S_.D, err)) return -2;
  if (dalloc(allocs_, K, &S_.npiv, err)) return -2;
  if (dalloc(allocs_, K, &S_.ndel, err)) return -2;
  if (dalloc(allocs_, K, &S_.m, err)) return -2;
  if (dalloc(allocs_, K, &S_.foff, err)) return -2;
  if (dalloc(allocs_, K, &S_.voff, err)) return -2;
  if (dalloc(allocs_, PS * (size_t)K, &S_.pstate, err)) return -2;
  if (dalloc(allocs_, K, &d_sync_nodes_, err)) return -2;
  if (dalloc(allocs_, K, &d_sync_root_, err)) return -2;
  if (dalloc(allocs_, 2, &S_.tops, err)) return -2;
  if (dalloc(allocs_, 8, &S_.flags, err)) return -2;
  if (dalloc(allocs_, 4 * (size_t)K, &S_.ninert, err)) return -2;

  {
    static bool attr_done = false;
    if (!attr_done) {
      HIPCHK(hipFuncSetAttribute((const void*)level_kernel<NT>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (LDS_MAX_M * LDS_MAX_M + 2 * LDS_MAX_M) * (int)sizeof(double)));
      HIPCHK(hipFuncSetAttribute((const void*)level_kernel<NTL>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (LDS_MAX_M * LDS_MAX_M + 2 * LDS_MAX_M) * (int)sizeof(double)));
      HIPCHK(hipFuncSetAttribute((const void*)panel_kernel<true, NTP>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)PANEL_LDS_BYTES));
      HIPCHK(hipFuncSetAttribute((const void*)panel_kernel<true, NTP / 2>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)PANEL_LDS_BYTES / 2));
      HIPCHK(hipFuncSetAttribute((const void*)panel_kernel<true, NTP / 4>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)PANEL_LDS_BYTES / 4));
      HIPCHK(hipFuncSetAttribute((const void*)big_assemble_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 150 * 1024));
      HIPCHK(hipFuncSetAttribute((const void*)solve_fwd_kernel<NT>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 150 * 1024));
      HIPCHK(hipFuncSetAttribute((const void*)solve_bwd_kernel<NT>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 150 * 1024));
      HIPCHK(hipFuncSetAttribute((const void*)solve_fwd_kernel<NTS>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 150 * 1024));
      attr_done = true;
    }
  }
  base_rcap_ = (unsigned long long)(std::max(1.0, delay_growth) * rneed) + 4096;
  base_icap_ = (unsigned long long)(std::max(1.0, delay_growth) * ineed) + 4096;
  const int rc = ensure_pools(base_rcap_, base_icap_, err);
  // The plan was uploaded through the null stream: make it visible to
  // kernels on any (non-blocking) stream before the first factorization.
  HIPCHK(hipDeviceSynchronize());
  return rc;
}

int LdlNumeric::ensure_pools(unsigned long long rcap, unsigned long long icap,
                             std::string& err) {
  if (rcap > S_.rcap || icap > S_.icap) drop_graphs();
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
    if (S_.fpos) HIPCHK(hipFree(S_.fpos));
    S_.ipool = nullptr;
    S_.fpos = nullptr;
    S_.icap = 0;
    hipError_t e = hipMalloc(&S_.ipool, icap * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&S_.fpos, icap * sizeof(int));
    if (e != hipSuccess) {
      err = "front index pool allocation failed";
      return -5;
    }
    S_.icap = icap;
  }
  return 0;
}

int LdlNumeric::run_factor(const double* vals, const double* shift,
                           FactorParams prm, bool cholesky, bool sync_mode,
                           hipStream_t stream, std::string& err) {
  const int rc = enqueue_factor(vals, shift, prm, cholesky, sync_mode, stream, err);
  if (rc != 0) return rc;
  return finish_factor(stream, err);
}

// Launch sequence of one factorization.  Without sync_mode it makes no
// host synchronization, so it can be captured into a HIP graph.
int LdlNumeric::enqueue_factor(const double* vals, const double* shift,
                               FactorParams prm, bool cholesky, bool sync_mode,
                               hipStream_t stream, std::string& err) {
  const bool mixed = prm.mixed;
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, stream, S_);
  HIPCHK(hipGetLastError());
  if (scaling_) {
    int nb = (n_ + NT - 1) / NT;
    nb = std::min(std::max(nb, 1), 65535);
    // Five Ruiz passes, alternating between the two scale buffers, the
    // last one writing scale_.
    constexpr int kPasses = 5;
    const double* sin = nullptr;
    for (int it = 0; it < kPasses; it++) {
      double* sout = ((kPasses - 1 - it) % 2 == 0) ? scale_ : scale_tmp_;
      hipLaunchKernelGGL(equilibrate_kernel, dim3(nb), dim3(NT), 0, stream, n_,
                         d_rowp_, d_cols_, vals, shift, sin, sout,
                         (const int*)(masked() ? d_row_blk_ : nullptr),
                         (const int*)d_blk_active_);
      sin = sout;
    }
    HIPCHK(hipGetLastError());
  }
  const int nlev = num_levels();
  // With a mask, the compacted lists of the blocks that take part.
  const bool ml = masked();
  const std::vector<int>& LP = ml ? m_level_ptr_ : level_ptr_;
  const std::vector<int>& LS = ml ? m_level_split_ : level_split_;
  const int* LN = ml ? d_m_level_nodes_ : d_level_nodes_;
  const std::vector<int>& BP = ml ? m_big_ptr_ : big_ptr_;
  for (int l = 0; l < nlev; l++) {
    const int split = LS[l];
    const int parts[2][3] = {{LP[l], split - LP[l], level_lds_small_[l]},
                             {split, LP[l + 1] - split, level_lds_[l]}};
    tk_begin(stream);
    for (const auto& pt : parts) {
      if (pt[1] == 0) continue;
      const int lds = pt[2];
      // The larger fronts (one workgroup per CU by their LDS) take NTL
      // threads: more waves share each pivot step's update.
      if (&pt == &parts[0])
        hipLaunchKernelGGL(level_kernel<NT>, dim3(pt[1]), dim3(NT),
                         size_t(lds + 2 * LDS_MAX_M) * sizeof(double), stream, P_, S_, vals,
                         shift, LN + pt[0], d_big_mask_, lds,
                         prm.ustab, prm.pivot_tol, cholesky ? 1 : 0,
                         sync_mode ? -1 : DCAP_BIG,
                         (const double*)(scaling_ ? scale_ : nullptr),
                         (const int*)(nbatch_ ? d_node_cap_ : nullptr));
      else
        hipLaunchKernelGGL(level_kernel<NTL>, dim3(pt[1]), dim3(NTL),
                         size_t(lds + 2 * LDS_MAX_M) * sizeof(double), stream, P_, S_, vals,
                         shift, LN + pt[0], d_big_mask_, lds,
                         prm.ustab, prm.pivot_tol, cholesky ? 1 : 0,
                         sync_mode ? -1 : DCAP_BIG,
                         (const double*)(scaling_ ? scale_ : nullptr),
                         (const int*)(nbatch_ ? d_node_cap_ : nullptr));
    }
    tk_end(stream, TK_LEVEL);
    HIPCHK(hipGetLastError());
    const int b0 = BP[l], nb = BP[l + 1] - BP[l];
    if (nb == 0 && !sync_mode) continue;
    int npanels = big_npanels_[l], tiles = big_tiles_[l];
    const int* bnodes = (ml ? d_m_big_nodes_ : d_big_nodes_) + b0;
    const int* broot = (ml ? d_m_big_root_ : d_big_root_) + b0;
    const int64_t* woff = (ml ? d_m_big_woff_ : d_big_woff_) + b0;
    int nlaunch = nb;
    if (sync_mode) {
      // Front sizes are known only after assembly: read them back and list
      // the fronts of this level that exceed the workgroup limit.
      HIPCHK(hipStreamSynchronize(stream));
      int flag = 0;
      HIPCHK(hipMemcpyAsync(&flag, S_.flags, sizeof(int), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      if (flag != 0) return flag;
      std::vector<int>& hm = h_m_;
      std::vector<int> hD(K_);
      hm.resize(K_);
      HIPCHK(hipMemcpyAsync(hm.data(), S_.m, K_ * sizeof(int), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipMemcpyAsync(hD.data(), S_.D, K_ * sizeof(int), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      std::vector<int> ln, lr;
      std::vector<int64_t> wo;
      int64_t tot = 0;
      npanels = 0;
      tiles = 0;
      for (int t = level_ptr_[l]; t < level_ptr_[l + 1]; t++) {
        const int node = level_nodes_[t];
        if (nbatch_ && !h_blk_active_[h_node_blk_[node]]) continue;
        if (hm[node] <= LDS_MAX_M) continue;
        ln.push_back(node);
        lr.push_back(h_cs_[node] == 0 ? 1 : 0);
        wo.push_back(tot);
        tot += (int64_t)hm[node] * NBW;
        npanels = std::max(npanels, (h_ns_[node] + hD[node] + NBR - 1) / NBR);
        tiles = std::max(tiles, (hm[node] + TILE - 1) / TILE);
      }
      nlaunch = int(ln.size());
      if (nlaunch == 0) continue;
      if ((size_t)tot > W_cap_) {
        if (W_) HIPCHK(hipFree(W_));
        W_ = nullptr;
        HIPCHK(hipMalloc(&W_, tot * sizeof(double)));
        W_cap_ = tot;
        drop_graphs();
      }
      HIPCHK(hipMemcpyAsync(d_big_woff_sync_, wo.data(), nlaunch * sizeof(int64_t),
                            hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(d_sync_nodes_, ln.data(), nlaunch * sizeof(int),
                            hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(d_sync_root_, lr.data(), nlaunch * sizeof(int),
                            hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));  // the host lists are loop locals
      bnodes = d_sync_nodes_;
      broot = d_sync_root_;
      woff = d_big_woff_sync_;
    }
    {
      // Rows of the level's largest blocked front (planned, or level-
      // synchronously the actual ones): the assembly's position maps.
      int mmax = big_mmax_[l];
      if (sync_mode) {
        mmax = 0;
        for (int t = level_ptr_[l]; t < level_ptr_[l + 1]; t++) {
          const int node = level_nodes_[t];
          if (nbatch_ && !h_blk_active_[h_node_blk_[node]]) continue;
          if (h_m_[node] > LDS_MAX_M) mmax = std::max(mmax, h_m_[node]);
        }
      }
      const double* scp = scaling_ ? scale_ : nullptr;
      tk_begin(stream);
      hipLaunchKernelGGL(big_assemble_kernel, dim3(128, nlaunch), dim3(NT),
                         2 * sizeof(int) * (size_t)mmax, stream, P_, S_, bnodes, vals, shift,
                         scp, mmax);
      tk_end(stream, TK_ASSEMBLE);
      HIPCHK(hipGetLastError());
    }
    for (int q = 0; q < npanels; q++) {
      tk_begin(stream);
      // One row per thread: the level's planned fronts (or, level-
      // synchronously, any front up to NTP rows) in LDS; taller ones in
      // global memory.
      const int nth = sync_mode ? NTP : big_nth_[l];
      const size_t plds = sizeof(double) * LS_COLS * nth;
      if (nth == NTP / 4)
        hipLaunchKernelGGL((panel_kernel<true, NTP / 4>), dim3(nlaunch), dim3(nth), plds,
                           stream, S_, bnodes, broot, W_, woff, prm.ustab, prm.pivot_tol,
                           cholesky ? 1 : 0, nth);
      else if (nth == NTP / 2)
        hipLaunchKernelGGL((panel_kernel<true, NTP / 2>), dim3(nlaunch), dim3(nth), plds,
                           stream, S_, bnodes, broot, W_, woff, prm.ustab, prm.pivot_tol,
                           cholesky ? 1 : 0, nth);
      else
        hipLaunchKernelGGL((panel_kernel<true, NTP>), dim3(nlaunch), dim3(nth), plds, stream,
                           S_, bnodes, broot, W_, woff, prm.ustab, prm.pivot_tol,
                           cholesky ? 1 : 0, nth);
      if (big_wide_[l] || sync_mode)
        hipLaunchKernelGGL((panel_kernel<false, NTP>), dim3(nlaunch), dim3(NTP), 0, stream, S_,
                           bnodes, broot, W_, woff, prm.ustab, prm.pivot_tol,
                           cholesky ? 1 : 0, nth);
      tk_end(stream, TK_PANEL);
      tk_begin(stream);
      if (mixed)
        hipLaunchKernelGGL((trailing_mfma_kernel<float, mfma_f4>), dim3(tiles, tiles, nlaunch),
                           dim3(NT), 0, stream, S_, bnodes, (const double*)W_, woff);
      else
        hipLaunchKernelGGL((trailing_mfma_kernel<double, mfma_d4>), dim3(tiles, tiles, nlaunch),
                           dim3(NT), 0, stream, S_, bnodes, (const double*)W_, woff);
      tk_end(stream, TK_TRAILING);
    }
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(fpos_kernel, dim3(std::min(K_, 4096)), dim3(NT), 0, stream, P_, S_, K_);
  hipLaunchKernelGGL(factor_stats_kernel, dim3(std::max(1, std::min((K_ + NT - 1) / NT, 1024))),
                     dim3(NT), 0, stream, S_, K_);
  HIPCHK(hipGetLastError());
  return 0;
}

// Wait for a factorization and read its status, inertia and statistics.
int LdlNumeric::finish_factor(hipStream_t stream, std::string& err) {
  HIPCHK(hipStreamSynchronize(stream));
  if (timing_) tk_collect();
  if (!h_flags_pin_) HIPCHK(hipHostMalloc((void**)&h_flags_pin_, 8 * sizeof(int)));
  int flags[8];
  HIPCHK(hipMemcpyAsync(h_flags_pin_, S_.flags, sizeof(flags), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  for (int q = 0; q < 8; q++) flags[q] = h_flags_pin_[q];
  if (flags[0] != 0) return flags[0];
  if (nbatch_ > 0) {
    // Per-block signs; a non-finite pivot fails only its own block.
    hipLaunchKernelGGL(batch_inertia_kernel, dim3(nbatch_), dim3(256), 0, stream,
                       nbatch_, (const int*)d_batch_seg_, (const int*)S_.ninert, d_batch_inert_);
    HIPCHK(hipGetLastError());
    binert_.resize(4 * (size_t)nbatch_);
    HIPCHK(hipMemcpyAsync(h_binert_pin_, d_batch_inert_, 4 * sizeof(int) * nbatch_,
                          hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    std::copy(h_binert_pin_, h_binert_pin_ + 4 * (size_t)nbatch_, binert_.begin());
  } else if (flags[4] != 0) {
    err = "non-finite pivot (the matrix holds NaN or Inf)";
    return -6;
  }
  inertia_[0] = flags[1];
  inertia_[1] = flags[2];
  inertia_[2] = flags[3];
  delayed_ = flags[5];
  max_front_seen_ = flags[6];
  return 0;
}

int LdlNumeric::factor(const double* vals, const double* shift, double ustab,
                       double pivot_tol, bool cholesky, bool mixed, hipStream_t stream,
                       std::string& err) {
  FactorParams prm{ustab, pivot_tol, mixed};
  factored_ = false;
  bool sync_mode = false;
  // Replay of the captured launch sequence (same buffers, stream and
  // parameters); a replay that needs the level-synchronous mode or larger
  // pools falls through to the eager path below.
  const bool graph_ok = graphs_ && !timing_ && !masked();
  if (graph_ok && fgraph_ && fg_vals_ == vals && fg_shift_ == shift && fg_stream_ == stream &&
      fg_ustab_ == ustab && fg_tol_ == pivot_tol && fg_chol_ == cholesky && fg_mixed_ == mixed) {
    HIPCHK(hipGraphLaunch(fgraph_, stream));
    const int rc = finish_factor(stream, err);
    if (rc == 0) {
      // A batch keeps the factors of the blocks a masked factorization
      // left out: the solves' LDS sizing covers every factor kept.
      max_p_ = nbatch_ ? std::max(max_p_, max_front_seen_) : max_front_seen_;
      factored_ = true;
      return 0;
    }
    if (rc != -7 && rc != -4) return rc;
    sync_mode = rc == -7;
  }
  for (int attempt = 0; attempt < 8; attempt++) {
    int rc = run_factor(vals, shift, prm, cholesky, sync_mode, stream, err);
    if (rc == -7) {
      // More delayed pivots than the static launch plan holds: redo the
      // factorization with per-level front sizes.
      sync_mode = true;
      continue;
    }
    if (rc == -4 && nbatch_) {
      // A block's delayed pivots outgrew its regions: double every
      // region, re-lay the pools out and refactor every block (the
      // re-layout loses the factors the mask kept).
      for (int b = 0; b < nbatch_; b++) {
        blk_rcap_h_[b] *= 2;
        blk_icap_h_[b] *= 2;
      }
      int r2 = layout_batch(err);
      if (r2) return r2;
      r2 = set_batch_mask(nullptr, stream, err);
      if (r2) return r2;
      continue;
    }
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
    // Largest front of this factorization (reduced on the device): the
    // solve kernels' LDS sizing (in a batch, of every factor kept).
    max_p_ = nbatch_ ? std::max(max_p_, max_front_seen_) : max_front_seen_;
    factored_ = true;
    // Capture once the same buffers have been factored eagerly twice (a
    // handle factored once does not pay for a capture).
    if (graph_ok && !sync_mode) {
      if (fseen_vals_ == vals && fseen_shift_ == shift && fseen_stream_ == stream)
        capture_factor(vals, shift, prm, cholesky, stream);
      fseen_vals_ = vals;
      fseen_shift_ = shift;
      fseen_stream_ = stream;
    }
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
  const size_t ubneed = (size_t)S_.icap * std::max(nrhs, 1);
  if (ubneed > ub_cap_) {
    if (ub_) HIPCHK(hipFree(ub_));
    ub_ = nullptr;
    ub_cap_ = 0;
    drop_graphs();
    if (hipMalloc(&ub_, ubneed * sizeof(double)) != hipSuccess) {
      err = "solve workspace allocation failed";
      return -5;
    }
    ub_cap_ = ubneed;
  }
  const bool graph_ok = graphs_ && !timing_ && !masked();
  if (graph_ok) {
    for (SolveGraph& g : sgraphs_) {
      if (g.x == x && g.nrhs == nrhs && g.ldx == ldx && g.stream == stream && g.maxp == max_p_) {
        HIPCHK(hipGraphLaunch(g.exec, stream));
        g.last_use = ++sgraph_clock_;
        return 0;
      }
    }
  }
  const int rc = enqueue_solve(nrhs, x, ldx, stream, err);
  if (rc != 0) return rc;
  if (timing_) {
    HIPCHK(hipStreamSynchronize(stream));
    tk_collect();
  } else if (graph_ok) {
    // Capture on the second eager solve of the same buffer.
    bool seen = false;
    for (const SolveGraph& g : sseen_)
      if (g.x == x && g.nrhs == nrhs && g.ldx == ldx && g.stream == stream && g.maxp == max_p_)
        seen = true;
    if (seen) {
      capture_solve(nrhs, x, ldx, stream);
    } else {
      if (sseen_.size() >= 2 * kMaxSolveGraphs) sseen_.erase(sseen_.begin());
      sseen_.push_back(SolveGraph{x, nrhs, ldx, stream, max_p_, nullptr, 0});
    }
  }
  return 0;
}

int LdlNumeric::enqueue_solve(int nrhs, double* x, int ldx, hipStream_t stream,
                              std::string& err) {
  const int tcap = max_p_;
  const size_t lds = size_t(tcap + PIVOT_TRI + NT) * sizeof(double);
  const size_t lds_wide = size_t(tcap + PIVOT_TRI + NTS) * sizeof(double);
  if (lds_wide > 150 * 1024) {
    err = "front too large for the LDS solve kernels";
    return -1;
  }
  const int nlev = num_levels();
  const int sb = std::min(4096, std::max(1, (n_ * std::max(nrhs, 1) + NT - 1) / NT));
  if (scaling_)
    hipLaunchKernelGGL(scale_vec_kernel, dim3(sb), dim3(NT), 0, stream, n_, nrhs, ldx,
                       (const double*)scale_, x, (const int*)(masked() ? d_row_blk_ : nullptr),
                       (const int*)d_blk_active_);
  const bool ml = masked();
  const std::vector<int>& LP = ml ? m_level_ptr_ : level_ptr_;
  const int* LN = ml ? d_m_level_nodes_ : d_level_nodes_;
  for (int l = 0; l < nlev; l++) {
    const int cnt = LP[l + 1] - LP[l];
    if (cnt == 0) continue;
    tk_begin(stream);
    if (solve_wide_[l])
      hipLaunchKernelGGL(solve_fwd_kernel<NTS>, dim3(cnt), dim3(NTS), lds_wide, stream, P_, S_,
                         LN + LP[l], x, ldx, nrhs, ub_, (size_t)S_.icap, tcap);
    else
      hipLaunchKernelGGL(solve_fwd_kernel<NT>, dim3(cnt), dim3(NT), lds, stream, P_, S_,
                         LN + LP[l], x, ldx, nrhs, ub_, (size_t)S_.icap, tcap);
    tk_end(stream, TK_SOLVE_FWD);
  }
  for (int l = nlev - 1; l >= 0; l--) {
    const int cnt = LP[l + 1] - LP[l];
    if (cnt == 0) continue;
    tk_begin(stream);
    hipLaunchKernelGGL(solve_bwd_kernel<NT>, dim3(cnt), dim3(NT), lds, stream, S_, LN + LP[l], x,
                       ldx, nrhs, tcap);
    tk_end(stream, TK_SOLVE_BWD);
  }
  if (scaling_)
    hipLaunchKernelGGL(scale_vec_kernel, dim3(sb), dim3(NT), 0, stream, n_, nrhs, ldx,
                       (const double*)scale_, x, (const int*)(masked() ? d_row_blk_ : nullptr),
                       (const int*)d_blk_active_);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---- HIP graphs of the factorization and solve launch sequences ----------
void LdlNumeric::drop_graphs() {
  if (fgraph_) (void)hipGraphExecDestroy(fgraph_);
  fgraph_ = nullptr;
  for (SolveGraph& g : sgraphs_) (void)hipGraphExecDestroy(g.exec);
  sgraphs_.clear();
}

// Record the factorization's launches (async plan) as a graph; on any
// failure the eager path simply stays in use.
// Captures run on a private non-blocking stream: recording executes
// nothing, the graph is launched on the caller's stream, and a capturing
// non-blocking stream never makes other threads' legacy-stream work depend
// on it.
bool LdlNumeric::capture_stream() {
  if (cap_stream_) return true;
  if (hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking) != hipSuccess) {
    cap_stream_ = nullptr;
    (void)hipGetLastError();
    return false;
  }
  return true;
}

void LdlNumeric::capture_factor(const double* vals, const double* shift, FactorParams prm,
                                bool cholesky, hipStream_t stream) {
  if (!capture_stream()) return;
  if (fgraph_) (void)hipGraphExecDestroy(fgraph_);
  fgraph_ = nullptr;
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) return;
  std::string err;
  const int rc = enqueue_factor(vals, shift, prm, cholesky, false, cap_stream_, err);
  if (hipStreamEndCapture(cap_stream_, &g) != hipSuccess || rc != 0 || g == nullptr) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return;
  }
  hipGraphExec_t ex = nullptr;
  if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess) {
    fgraph_ = ex;
    fg_vals_ = vals;
    fg_shift_ = shift;
    fg_stream_ = stream;
    fg_ustab_ = prm.ustab;
    fg_tol_ = prm.pivot_tol;
    fg_chol_ = cholesky;
    fg_mixed_ = prm.mixed;
  }
  (void)hipGraphDestroy(g);
  (void)hipGetLastError();
}

void LdlNumeric::capture_solve(int nrhs, double* x, int ldx, hipStream_t stream) {
  if (!capture_stream()) return;
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) return;
  std::string err;
  const int rc = enqueue_solve(nrhs, x, ldx, cap_stream_, err);
  if (hipStreamEndCapture(cap_stream_, &g) != hipSuccess || rc != 0 || g == nullptr) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return;
  }
  hipGraphExec_t ex = nullptr;
  if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess) {
    if (sgraphs_.size() >= kMaxSolveGraphs) {
      size_t lru = 0;
      for (size_t i = 1; i < sgraphs_.size(); i++)
        if (sgraphs_[i].last_use < sgraphs_[lru].last_use) lru = i;
      (void)hipGraphExecDestroy(sgraphs_[lru].exec);
      sgraphs_.erase(sgraphs_.begin() + lru);
    }
    sgraphs_.push_back(SolveGraph{x, nrhs, ldx, stream, max_p_, ex, ++sgraph_clock_});
  }
  (void)hipGraphDestroy(g);
  (void)hipGetLastError();
}

int LdlNumeric::set_batch(int nblocks, const int* node_seg, std::string& err) {
  if (nblocks < 1 || node_seg[0] != 0 || node_seg[nblocks] != K_) {
    err = "block segments must cover the nodes";
    return -1;
  }
  auto fr = [](auto*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
  };
  fr(d_batch_seg_);
  fr(d_batch_inert_);
  fr(d_node_blk_);
  fr(d_blk_active_);
  HIPCHK(hipMalloc(&d_batch_seg_, sizeof(int) * (nblocks + 1)));
  HIPCHK(hipMalloc(&d_batch_inert_, sizeof(int) * 4 * (size_t)nblocks));
  HIPCHK(hipMemcpy(d_batch_seg_, node_seg, sizeof(int) * (nblocks + 1), hipMemcpyHostToDevice));
  nbatch_ = nblocks;
  binert_.assign(4 * (size_t)nblocks, 0);
  if (h_binert_pin_) (void)hipHostFree(h_binert_pin_);
  h_binert_pin_ = nullptr;
  HIPCHK(hipHostMalloc((void**)&h_binert_pin_, 4 * sizeof(int) * (size_t)nblocks));
  // Node -> block, and each block's pool needs at delay_growth x its
  // static front sizes.
  h_node_blk_.assign(K_, 0);
  blk_rcap_h_.assign(nblocks, 0);
  blk_icap_h_.assign(nblocks, 0);
  for (int b = 0; b < nblocks; b++) {
    double rn = 0.0, in = 0.0;
    for (int k = node_seg[b]; k < node_seg[b + 1]; k++) {
      h_node_blk_[k] = b;
      const double m0 = double(h_ns_[k] + h_cs_[k]);
      rn += m0 * m0;
      in += m0;
    }
    blk_rcap_h_[b] = (unsigned long long)(std::max(1.0, delay_growth_) * rn) + 64;
    blk_icap_h_[b] = (unsigned long long)(std::max(1.0, delay_growth_) * in) + 64;
  }
  h_blk_active_.assign(nblocks, 1);
  mask_all_ = true;
  {
    std::vector<int> rb(n_, 0);
    for (int k = 0; k < K_; k++)
      for (int q = h_nptr_[k]; q < h_nptr_[k + 1]; q++) rb[h_nvars_[q]] = h_node_blk_[k];
    fr(d_row_blk_);
    HIPCHK(hipMalloc(&d_row_blk_, sizeof(int) * std::max(n_, 1)));
    HIPCHK(hipMemcpy(d_row_blk_, rb.data(), sizeof(int) * n_, hipMemcpyHostToDevice));
  }
  {
    // LDS slot each front gets with its own block alone.
    std::vector<int> cap(K_, 0);
    const int nlev = num_levels();
    auto side = [&](int k) { return std::min(h_ns_[k] + h_cs_[k] + 32, LDS_MAX_M); };
    std::vector<int> mx(nblocks, 0);
    for (int l = 0; l < nlev; l++) {
      const int parts[2][2] = {{level_ptr_[l], level_split_[l]},
                               {level_split_[l], level_ptr_[l + 1]}};
      for (const auto& pt : parts) {
        for (int t = pt[0]; t < pt[1]; t++) mx[h_node_blk_[level_nodes_[t]]] = 0;
        for (int t = pt[0]; t < pt[1]; t++) {
          const int k = level_nodes_[t];
          int& v = mx[h_node_blk_[k]];
          v = std::max(v, side(k));
        }
        for (int t = pt[0]; t < pt[1]; t++) {
          const int k = level_nodes_[t];
          const int v = mx[h_node_blk_[k]];
          cap[k] = v * v;
        }
      }
    }
    fr(d_node_cap_);
    HIPCHK(hipMalloc(&d_node_cap_, sizeof(int) * std::max(K_, 1)));
    HIPCHK(hipMemcpy(d_node_cap_, cap.data(), sizeof(int) * K_, hipMemcpyHostToDevice));
  }
  fr(d_m_level_nodes_);
  fr(d_m_big_nodes_);
  fr(d_m_big_root_);
  fr(d_m_big_woff_);
  HIPCHK(hipMalloc(&d_m_level_nodes_, sizeof(int) * std::max(K_, 1)));
  const size_t nbig = std::max<size_t>(1, big_nodes_.size());
  HIPCHK(hipMalloc(&d_m_big_nodes_, sizeof(int) * nbig));
  HIPCHK(hipMalloc(&d_m_big_root_, sizeof(int) * nbig));
  HIPCHK(hipMalloc(&d_m_big_woff_, sizeof(int64_t) * nbig));
  HIPCHK(hipMalloc(&d_node_blk_, sizeof(int) * std::max(K_, 1)));
  HIPCHK(hipMemcpy(d_node_blk_, h_node_blk_.data(), sizeof(int) * K_, hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&d_blk_active_, sizeof(int) * nblocks));
  HIPCHK(hipMemcpy(d_blk_active_, h_blk_active_.data(), sizeof(int) * nblocks,
                   hipMemcpyHostToDevice));
  return layout_batch(err);
}

// Regions of every block in the pools (rpool doubles, ipool ints) from the
// per-block capacities; grows the pools when they no longer hold them.
int LdlNumeric::layout_batch(std::string& err) {
  const int B = nbatch_;
  std::vector<unsigned long long> lay(4 * (size_t)B);
  unsigned long long r = 0, i = 0;
  for (int b = 0; b < B; b++) {
    lay[b] = r;
    lay[B + b] = i;
    lay[2 * B + b] = blk_rcap_h_[b];
    lay[3 * B + b] = blk_icap_h_[b];
    r += blk_rcap_h_[b];
    i += blk_icap_h_[b];
  }
  int rc = ensure_pools(r, i, err);
  if (rc) return rc;
  if (d_blk_layout_) (void)hipFree(d_blk_layout_);
  d_blk_layout_ = nullptr;
  if (d_tops_batch_) (void)hipFree(d_tops_batch_);
  d_tops_batch_ = nullptr;
  HIPCHK(hipMalloc(&d_blk_layout_, sizeof(unsigned long long) * 4 * (size_t)B));
  HIPCHK(hipMemcpy(d_blk_layout_, lay.data(), sizeof(unsigned long long) * 4 * (size_t)B,
                   hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&d_tops_batch_, sizeof(unsigned long long) * 2 * (size_t)B));
  HIPCHK(hipMemset(d_tops_batch_, 0, sizeof(unsigned long long) * 2 * (size_t)B));
  S_.nblk = B;
  S_.node_blk = d_node_blk_;
  S_.blk_active = d_blk_active_;
  S_.blk_rbase = d_blk_layout_;
  S_.blk_ibase = d_blk_layout_ + B;
  S_.blk_rcap = d_blk_layout_ + 2 * B;
  S_.blk_icap = d_blk_layout_ + 3 * B;
  S_.tops = d_tops_batch_;
  // Kernel arguments changed: captured launch sequences are stale, and
  // the re-laid-out pools hold no factor.
  drop_graphs();
  factored_ = false;
  // Tables written through the null stream must be complete before kernels
  // on the caller's (possibly non-blocking) stream read them.
  HIPCHK(hipDeviceSynchronize());
  return 0;
}

int LdlNumeric::set_batch_mask(const int* mask, hipStream_t stream, std::string& err) {
  if (!nbatch_) {
    err = "no batch segments set";
    return -3;
  }
  int on = 0;
  for (int b = 0; b < nbatch_; b++) {
    h_blk_active_[b] = mask ? (mask[b] != 0) : 1;
    on += h_blk_active_[b];
  }
  mask_all_ = on == nbatch_;
  // The uploads go through one pinned staging buffer, so that each copy is
  // asynchronous and its source outlives it (drained at the end).
  {
    size_t need = sizeof(int) * nbatch_ + 64;
    if (!mask_all_)
      need += sizeof(int) * (size_t)K_ + (3 * sizeof(int) + sizeof(int64_t)) * big_nodes_.size() + 256;
    if (need > h_stage_cap_) {
      if (h_stage_pin_) (void)hipHostFree(h_stage_pin_);
      h_stage_pin_ = nullptr;
      h_stage_cap_ = 0;
      HIPCHK(hipHostMalloc((void**)&h_stage_pin_, need));
      h_stage_cap_ = need;
    }
  }
  size_t soff = 0;
  const void* src = nullptr;
  if (stage(h_blk_active_.data(), sizeof(int) * nbatch_, soff, &src, err)) return -2;
  HIPCHK(hipMemcpyAsync(d_blk_active_, src, sizeof(int) * nbatch_, hipMemcpyHostToDevice, stream));
  if (!mask_all_) {
    // Compacted launch lists (same order, same LDS buckets).
    const int nlev = num_levels();
    m_level_ptr_.assign(nlev + 1, 0);
    m_level_split_.assign(nlev, 0);
    m_level_nodes_.clear();
    m_big_ptr_.assign(nlev + 1, 0);
    m_big_nodes_.clear();
    m_big_root_.clear();
    m_big_woff_.clear();
    auto on_node = [&](int node) { return h_blk_active_[h_node_blk_[node]] != 0; };
    for (int l = 0; l < nlev; l++) {
      for (int t = level_ptr_[l]; t < level_split_[l]; t++)
        if (on_node(level_nodes_[t])) m_level_nodes_.push_back(level_nodes_[t]);
      m_level_split_[l] = int(m_level_nodes_.size());
      for (int t = level_split_[l]; t < level_ptr_[l + 1]; t++)
        if (on_node(level_nodes_[t])) m_level_nodes_.push_back(level_nodes_[t]);
      m_level_ptr_[l + 1] = int(m_level_nodes_.size());
      for (int t = big_ptr_[l]; t < big_ptr_[l + 1]; t++) {
        if (!on_node(big_nodes_[t])) continue;
        m_big_nodes_.push_back(big_nodes_[t]);
        m_big_root_.push_back(h_big_root_[t]);
        m_big_woff_.push_back(h_big_woff_[t]);
      }
      m_big_ptr_[l + 1] = int(m_big_nodes_.size());
    }
    auto up = [&](void* dst, const void* h, size_t bytes) -> int {
      const void* sp = nullptr;
      if (stage(h, bytes, soff, &sp, err)) return -2;
      if (hipMemcpyAsync(dst, sp, bytes, hipMemcpyHostToDevice, stream) != hipSuccess) {
        err = "hipMemcpyAsync";
        return -2;
      }
      return 0;
    };
    if (!m_level_nodes_.empty() &&
        up(d_m_level_nodes_, m_level_nodes_.data(), sizeof(int) * m_level_nodes_.size()))
      return -2;
    if (!m_big_nodes_.empty()) {
      const size_t nb = m_big_nodes_.size();
      if (up(d_m_big_nodes_, m_big_nodes_.data(), sizeof(int) * nb) ||
          up(d_m_big_root_, m_big_root_.data(), sizeof(int) * nb) ||
          up(d_m_big_woff_, m_big_woff_.data(), sizeof(int64_t) * nb))
        return -2;
    }
  }
  HIPCHK(hipStreamSynchronize(stream));
  return 0;
}

// Copy `bytes` of host data into the pinned staging buffer at `off` (8-byte
// aligned); *out is the staged address.
int LdlNumeric::stage(const void* src, size_t bytes, size_t& off, const void** out,
                      std::string& err) {
  off = (off + 7) & ~size_t(7);
  if (off + bytes > h_stage_cap_) {
    err = "staging buffer too small";
    return -2;
  }
  std::memcpy(h_stage_pin_ + off, src, bytes);
  *out = h_stage_pin_ + off;
  off += bytes;
  return 0;
}

// Sign test of a factorization without pivoting of a symmetric
// quasi-definite matrix [A B^T; B -C] (SparseCholesky,
// include/sparse_cholesky.h:9-63): the pivot of a variable below sqdef
// (the A block) must be positive and that of the others negative.  The
// smallest 1-based index of a variable that fails goes to out[0] (atomicMin
// over an INT_MAX start); a non-finite pivot fails too.
__global__ void sqd_sign_kernel(DevState S, int K, int sqdef, int* __restrict__ out) {
  const int node = blockIdx.x;
  if (node >= K || !node_on(S, node)) return;
  const int p = S.npiv[node], m = S.m[node];
  const double* F = S.rpool + S.foff[node];
  const int* vars = S.ipool + S.voff[node];
  for (int j = threadIdx.x; j < p; j += blockDim.x) {
    const int q = vars[j];
    const int v = q >= 0 ? q : -q - 1;
    const double d = F[j + (size_t)j * m];
    const bool ok = v < sqdef ? (d > 0.0) : (d < 0.0);
    if (!ok) atomicMin(out, v + 1);
  }
}

int LdlNumeric::sqd_check(int sqdef, int* first_bad, hipStream_t stream, std::string& err) {
  *first_bad = 0;
  if (!factored_) {
    err = "no factorization";
    return -3;
  }
  if (!d_sqd_) {
    HIPCHK(hipMalloc((void**)&d_sqd_, sizeof(int)));
    allocs_.push_back(d_sqd_);
  }
  if (!h_flags_pin_) HIPCHK(hipHostMalloc((void**)&h_flags_pin_, 8 * sizeof(int)));
  h_flags_pin_[7] = INT_MAX;
  HIPCHK(hipMemcpyAsync(d_sqd_, &h_flags_pin_[7], sizeof(int), hipMemcpyHostToDevice, stream));
  if (K_ > 0)
    hipLaunchKernelGGL(sqd_sign_kernel, dim3(K_), dim3(64), 0, stream, S_, K_, sqdef, d_sqd_);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&h_flags_pin_[7], d_sqd_, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  const int h = h_flags_pin_[7];
  *first_bad = h == INT_MAX ? 0 : h;
  return 0;
}

int LdlNumeric::pivot_counts(int* np, int* nd, std::string& err) const {
  HIPCHK(hipMemcpy(np, S_.npiv, K_ * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(nd, S_.ndel, K_ * sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

}  // namespace amigo_amd
