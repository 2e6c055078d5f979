// Element-wise helper kernels of the native interior-point driver
// (ipm_driver.cpp): diagonal shifts of the inertia correction, the shifted
// KKT product of iterative refinement, multiplier updates of the
// initialization and of the second-order correction, and plain vector
// fills, gathers and scatters.  Every kernel touches only the rows of the
// problems that take part (active[k] != 0; active == nullptr: all).
#include <hip/hip_runtime.h>

#include "ipm_driver.h"

namespace amigo_amd {
namespace drv {
namespace {

constexpr int NT = 256;

__device__ __forceinline__ bool on(const int* active, const int* row_prob, int r) {
  return !active || active[row_prob[r]] != 0;
}

int grid(long long n) {
  long long b = (n + NT - 1) / NT;
  if (b < 1) b = 1;
  return b > 65535 ? 65535 : int(b);
}

// The part of a launch's rows that one thread visits: lo, lo + step, ... <
// hi, with `unit` threads per row.  Without a list the whole grid strides
// over rows 0..n-1; with one, block b serves listed problem b / chunks and
// strides over that problem's segment with the other chunks of it.
struct Span {
  int lo, hi, step;
};
__device__ __forceinline__ Span span(const Rows& R, int chunks, int n, int unit) {
  const int per = NT / unit;
  const int t = threadIdx.x / unit;
  if (!R.list) return Span{int(blockIdx.x) * per + t, n, int(gridDim.x) * per};
  const int e = blockIdx.x / chunks, c = blockIdx.x - e * chunks;
  const int k = R.list[e];
  return Span{R.seg[k] + c * per + t, R.seg[k + 1], chunks * per};
}

// Blocks per listed problem: enough for its longest segment, at most 64.
int chunks_of(const Rows& R, int unit) {
  const long long need = ((long long)R.max_len * unit + NT - 1) / NT;
  return need < 1 ? 1 : (need > 64 ? 64 : int(need));
}

__global__ void fill_kernel(long long n, double v, double* __restrict__ x) {
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT)
    x[i] = v;
}

// reg = base + eps on primal rows (the numerical_eps of
// inertia_correction.py:295), base on multiplier rows.
__global__ void reg_kernel(int n, const int* __restrict__ is_mult, const int* __restrict__ row_prob,
                           const int* __restrict__ active, const double* __restrict__ base,
                           double eps, double* __restrict__ reg, Rows R, int chunks) {
  const Span sp = span(R, chunks, n, 1);
  for (int r = sp.lo; r < sp.hi; r += sp.step) {
    if (!on(active, row_prob, r)) continue;
    reg[r] = is_mult[r] ? base[r] : __dadd_rn(base[r], eps);
  }
}

// shift = reg + delta_w on primal rows and reg - delta_c on multiplier rows
// of each problem (perturbations that are zero leave reg as it is).
__global__ void shift_kernel(int n, const int* __restrict__ is_mult, const int* __restrict__ row_prob,
                             const int* __restrict__ active, const double* __restrict__ reg,
                             const double* __restrict__ dw, const double* __restrict__ dc,
                             double* __restrict__ shift, Rows R, int chunks) {
  const Span sp = span(R, chunks, n, 1);
  for (int r = sp.lo; r < sp.hi; r += sp.step) {
    if (!on(active, row_prob, r)) continue;
    const int k = row_prob[r];
    double v = reg[r];
    if (is_mult[r]) {
      if (dc[k] > 0.0) v = __dsub_rn(v, dc[k]);
    } else {
      if (dw[k] > 0.0) v = __dadd_rn(v, dw[k]);
    }
    shift[r] = v;
  }
}

__global__ void negate_kernel(int n, double* __restrict__ x) {
  for (int r = blockIdx.x * NT + threadIdx.x; r < n; r += gridDim.x * NT) x[r] = -x[r];
}

// 1 on primal rows, 0 on multiplier rows (the diagonal of the least-squares
// multiplier system, multiplier_initialization.py:33-36).
__global__ void unit_primal_kernel(int n, const int* __restrict__ is_mult, double* __restrict__ d) {
  for (int r = blockIdx.x * NT + threadIdx.x; r < n; r += gridDim.x * NT) d[r] = is_mult[r] ? 0.0 : 1.0;
}

// y = (A + diag(shift)) x on the rows of the active problems, one wave per
// row: the matrix that was factored.
__global__ void kkt_matvec_kernel(int n, const int* __restrict__ rowp, const int* __restrict__ cols,
                                  const double* __restrict__ vals, const double* __restrict__ shift,
                                  const double* __restrict__ x, double* __restrict__ y,
                                  const int* __restrict__ row_prob, const int* __restrict__ active,
                                  Rows R, int chunks) {
  const int lane = threadIdx.x & 63;
  const Span sp = span(R, chunks, n, 64);
  for (int r = sp.lo; r < sp.hi; r += sp.step) {
    if (!on(active, row_prob, r)) continue;
    double s = 0.0;
    for (int p = rowp[r] + lane; p < rowp[r + 1]; p += 64) s += vals[p] * x[cols[p]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[r] = __dadd_rn(s, __dmul_rn(shift[r], x[r]));
  }
}

__global__ void gather_kernel(int n, const long long* __restrict__ idx, const double* __restrict__ src,
                              double* __restrict__ dst) {
  for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) dst[i] = src[idx[i]];
}

// x[c] = keep[k] ? src[c] : 0 on the multiplier rows c of problem k.
__global__ void set_multipliers_kernel(int nc, const long long* __restrict__ cidx,
                                       const int* __restrict__ row_prob, const int* __restrict__ keep,
                                       const int* __restrict__ active,
                                       const double* __restrict__ src, double* __restrict__ x,
                                       Rows R, int chunks) {
  const Span sp = span(R, chunks, nc, 1);
  for (int j = sp.lo; j < sp.hi; j += sp.step) {
    const long long c = cidx[j];
    const int k = row_prob[c];
    if (active && !active[k]) continue;
    x[c] = keep[k] ? src[c] : 0.0;
  }
}

__global__ void zero_rows_kernel(int nc, const long long* __restrict__ cidx, double* __restrict__ x,
                                 Rows R, int chunks) {
  const Span sp = span(R, chunks, nc, 1);
  for (int j = sp.lo; j < sp.hi; j += sp.step) x[cidx[j]] = 0.0;
}

// Second-order correction right-hand side (filter_line_search.py:229-238):
// c[m] = res[m] + alpha_soc * c[m] on the multiplier rows, then res = c on
// every row of the active problems.
__global__ void soc_rhs_kernel(int nc, const long long* __restrict__ cidx,
                               const int* __restrict__ row_prob, const int* __restrict__ active,
                               const double* __restrict__ alpha, const double* __restrict__ res,
                               double* __restrict__ c, Rows R, int chunks) {
  const Span sp = span(R, chunks, nc, 1);
  for (int j = sp.lo; j < sp.hi; j += sp.step) {
    const long long m = cidx[j];
    const int k = row_prob[m];
    if (active && !active[k]) continue;
    c[m] = __dadd_rn(res[m], __dmul_rn(alpha[k], c[m]));
  }
}

__global__ void copy_rows_kernel(int n, const int* __restrict__ row_prob, const int* __restrict__ active,
                                 const double* __restrict__ src, double* __restrict__ dst, Rows R,
                                 int chunks) {
  const Span sp = span(R, chunks, n, 1);
  for (int r = sp.lo; r < sp.hi; r += sp.step)
    if (on(active, row_prob, r)) dst[r] = src[r];
}

// sv = 1 on every row, dc[j] on multiplier row cidx[j]; lbh = lbh0 * dc.
__global__ void scale_vec_kernel(int n, double* __restrict__ sv) {
  for (int r = blockIdx.x * NT + threadIdx.x; r < n; r += gridDim.x * NT) sv[r] = 1.0;
}
__global__ void scatter_dc_kernel(int nc, const long long* __restrict__ cidx, const double* __restrict__ dc,
                                  const double* __restrict__ lbh0, double* __restrict__ sv,
                                  double* __restrict__ lbh) {
  for (int j = blockIdx.x * NT + threadIdx.x; j < nc; j += gridDim.x * NT) {
    sv[cidx[j]] = dc[j];
    lbh[j] = __dmul_rn(lbh0[j], dc[j]);
  }
}

}  // namespace

#define DRV_LAUNCH(kernel, blocks, ...)                                                   \
  do {                                                                                   \
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(NT), 0, s, __VA_ARGS__);               \
    hipError_t e__ = hipGetLastError();                                                  \
    if (e__ != hipSuccess) return e__;                                                   \
  } while (0)

// Grid of a launch over n rows with `unit` threads per row, or over the
// listed problems' segments.
int blocks(const Rows& R, int chunks, int n, int unit) {
  return R.list ? R.nlist * chunks : grid((long long)n * unit);
}

hipError_t fill(long long n, double v, double* x, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  DRV_LAUNCH(fill_kernel, grid(n), n, v, x);
  return hipSuccess;
}

hipError_t regularized_base(int n, const int* is_mult, const int* row_prob, const int* active,
                            const double* base, double eps, double* reg, hipStream_t s,
                            const Rows& R) {
  const int ch = chunks_of(R, 1);
  DRV_LAUNCH(reg_kernel, blocks(R, ch, n, 1), n, is_mult, row_prob, active, base, eps, reg, R,
             ch);
  return hipSuccess;
}

hipError_t diagonal_shift(int n, const int* is_mult, const int* row_prob, const int* active,
                          const double* reg, const double* dw, const double* dc, double* shift,
                          hipStream_t s, const Rows& R) {
  const int ch = chunks_of(R, 1);
  DRV_LAUNCH(shift_kernel, blocks(R, ch, n, 1), n, is_mult, row_prob, active, reg, dw, dc, shift,
             R, ch);
  return hipSuccess;
}

hipError_t negate(int n, double* x, hipStream_t s) {
  DRV_LAUNCH(negate_kernel, grid(n), n, x);
  return hipSuccess;
}

hipError_t unit_primal(int n, const int* is_mult, double* d, hipStream_t s) {
  DRV_LAUNCH(unit_primal_kernel, grid(n), n, is_mult, d);
  return hipSuccess;
}

hipError_t kkt_matvec(int n, const int* rowp, const int* cols, const double* vals,
                      const double* shift, const double* x, double* y, const int* row_prob,
                      const int* active, hipStream_t s, const Rows& R) {
  const int ch = chunks_of(R, 64);
  DRV_LAUNCH(kkt_matvec_kernel, blocks(R, ch, n, 64), n, rowp, cols, vals, shift, x, y, row_prob,
             active, R, ch);
  return hipSuccess;
}

hipError_t gather(int n, const long long* idx, const double* src, double* dst, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  DRV_LAUNCH(gather_kernel, grid(n), n, idx, src, dst);
  return hipSuccess;
}

hipError_t set_multipliers(int nc, const long long* cidx, const int* row_prob, const int* keep,
                           const int* active, const double* src, double* x, hipStream_t s,
                           const Rows& R) {
  if (nc <= 0) return hipSuccess;
  const int ch = chunks_of(R, 1);
  DRV_LAUNCH(set_multipliers_kernel, blocks(R, ch, nc, 1), nc, cidx, row_prob, keep, active, src,
             x, R, ch);
  return hipSuccess;
}

hipError_t zero_rows(int nc, const long long* cidx, double* x, hipStream_t s, const Rows& R) {
  if (nc <= 0) return hipSuccess;
  const int ch = chunks_of(R, 1);
  DRV_LAUNCH(zero_rows_kernel, blocks(R, ch, nc, 1), nc, cidx, x, R, ch);
  return hipSuccess;
}

hipError_t soc_rhs(int nc, const long long* cidx, const int* row_prob, const int* active,
                   const double* alpha, const double* res, double* c, hipStream_t s,
                   const Rows& R) {
  if (nc <= 0) return hipSuccess;
  const int ch = chunks_of(R, 1);
  DRV_LAUNCH(soc_rhs_kernel, blocks(R, ch, nc, 1), nc, cidx, row_prob, active, alpha, res, c, R,
             ch);
  return hipSuccess;
}

hipError_t copy_rows(int n, const int* row_prob, const int* active, const double* src, double* dst,
                     hipStream_t s, const Rows& R) {
  const int ch = chunks_of(R, 1);
  DRV_LAUNCH(copy_rows_kernel, blocks(R, ch, n, 1), n, row_prob, active, src, dst, R, ch);
  return hipSuccess;
}

hipError_t constraint_scale_vectors(int n, int nc, const long long* cidx, const double* dc,
                                    const double* lbh0, double* sv, double* lbh, hipStream_t s) {
  DRV_LAUNCH(scale_vec_kernel, grid(n), n, sv);
  if (nc > 0) DRV_LAUNCH(scatter_dc_kernel, grid(nc), nc, cidx, dc, lbh0, sv, lbh);
  return hipSuccess;
}

}  // namespace drv
}  // namespace amigo_amd
