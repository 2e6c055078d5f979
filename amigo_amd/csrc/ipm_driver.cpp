// Native interior-point driver for K independent NLPs in lockstep.
//
// The algorithm is the primal-dual interior-point method with a filter line
// search of Waechter & Biegler, Math. Program. 106 (2006) 25-57, in the
// variant the reference's Python driver runs (amigo/algorithm/ipm_driver.py
// and its mixins; the numbered steps below name the files).  The driver is
// organised around phases over sets of problems: every device call of a
// phase covers the problems that take part in it (a mask), the scalars it
// returns are read back together once per phase, and each problem's
// decisions are made on the host from its own numbers only.  A problem is
// therefore computed exactly as it is alone, whichever problems share its
// launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/amigo_assembly.h"
#include "../../include/amigo_driver.h"
#include "ipm_driver.h"

namespace amigo_amd {
namespace drv {

// ---- Regularization ---------------------------------------------------------

bool Regularization::raise_dw() {
  const double before = dw_;
  if (dw_ == 0.0) {
    dw_ = dw_last_ == 0.0 ? 1e-4 : std::max(1e-20, dw_last_ * (1.0 / 3));
  } else {
    // A large first step when the last successful value is small or absent.
    const bool first = dw_last_ == 0.0 || 1e5 * dw_last_ < dw_;
    dw_ *= first ? 100.0 : 8.0;
  }
  if (dw_ > 1e20) {
    dw_ = before;
    dw_last_ = 0.0;
    return false;
  }
  return true;
}

void Regularization::strike(Degeneracy& d) {
  if (d != Degeneracy::Unknown) return;
  if (++strikes_ >= 3) d = Degeneracy::Yes;
}

void Regularization::settle_probe() {
  switch (probe_) {
    case Probe::None:
      return;
    case Probe::Plain:  // the unperturbed system was fine
      if (hess_ == Degeneracy::Unknown) hess_ = Degeneracy::No;
      if (jac_ == Degeneracy::Unknown) jac_ = Degeneracy::No;
      break;
    case Probe::ConstraintsOnly:  // delta_c alone repaired it
      if (hess_ == Degeneracy::Unknown) hess_ = Degeneracy::No;
      strike(jac_);
      break;
    case Probe::HessianOnly:  // delta_w alone repaired it
      if (jac_ == Degeneracy::Unknown) jac_ = Degeneracy::No;
      strike(hess_);
      break;
    case Probe::Both:
      if (++strikes_ >= 3) hess_ = jac_ = Degeneracy::Yes;
      break;
  }
  probe_ = Probe::None;
}

void Regularization::begin_system() {
  settle_probe();
  if (dw_ > 0.0) dw_last_ = dw_;
  if (dc_ > 0.0) dc_last_ = dc_;
  const bool probing = hess_ == Degeneracy::Unknown || jac_ == Degeneracy::Unknown;
  probe_ = probing ? Probe::Plain : Probe::None;
  dc_ = jac_ == Degeneracy::Yes ? dc_value() : 0.0;
  dw_ = 0.0;
  quality_raised_ = false;
  if (hess_ == Degeneracy::Yes) raise_dw();
}

bool Regularization::after_wrong_inertia() {
  settle_probe();
  if (raise_dw()) return true;
  if (dc_ != 0.0) return false;
  // delta_w ran out: switch delta_c on and climb delta_w again from zero.
  dc_ = dc_value();
  dw_ = 0.0;
  if (hess_ == Degeneracy::Yes) hess_ = Degeneracy::Unknown;
  probe_ = Probe::None;
  return raise_dw();
}

bool Regularization::after_singular() {
  const bool probing = hess_ == Degeneracy::Unknown || jac_ == Degeneracy::Unknown;
  if (!probing) {
    if (dc_ > 0.0) return raise_dw();
    dc_ = dc_value();
    return true;
  }
  switch (probe_) {
    case Probe::Plain:
      if (jac_ == Degeneracy::Unknown) {
        dc_ = dc_value();
        probe_ = Probe::ConstraintsOnly;
        return true;
      }
      if (!raise_dw()) return false;
      probe_ = Probe::HessianOnly;
      return true;
    case Probe::ConstraintsOnly:
      dc_ = 0.0;
      if (!raise_dw()) return false;
      probe_ = Probe::HessianOnly;
      return true;
    case Probe::HessianOnly:
      dc_ = dc_value();
      if (!raise_dw()) return false;
      probe_ = Probe::Both;
      return true;
    case Probe::Both:
      return raise_dw();
    case Probe::None:
      return true;
  }
  return true;
}

bool Regularization::next(Outcome why, bool has_constraints) {
  if (why == Outcome::Singular && has_constraints) return after_singular();
  if (why == Outcome::TooFewNegative) {
    // First a better pivot quality of the same matrix (one raise per
    // system, 1e-6 -> 1e-3 -> ... -> 0.1 over the solve); the factor here
    // has a fixed pivot rule, so that retry factors the same matrix again.
    if (!quality_raised_ && pivot_quality_ < 0.1) {
      pivot_quality_ = std::min(0.1, std::pow(pivot_quality_, 0.5));
      quality_raised_ = true;
      return true;
    }
    return after_singular();
  }
  return after_wrong_inertia();
}

}  // namespace drv
}  // namespace amigo_amd

namespace {

using amigo_amd::drv::Filter;
using amigo_amd::drv::Regularization;
using Set = std::vector<char>;

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

struct DriverError {
  int code;
  std::string msg;
};

[[noreturn]] void raise(int code, const std::string& msg) { throw DriverError{code, msg}; }

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) raise(-2, std::string(what) + ": " + hipGetErrorString(e));
}

void ipm_ok(int rc, const char* what) {
  if (rc != 0) raise(rc, std::string(what) + ": " + amigo_ipm_last_error());
}
void asm_ok(int rc, const char* what) {
  if (rc != 0) raise(rc, std::string(what) + ": " + amigo_assembly_last_error());
}
void ldl_ok(int rc, const char* what) {
  if (rc != 0) raise(rc, std::string(what) + ": " + amigo_last_error());
}

bool any(const Set& s) {
  for (char c : s)
    if (c) return true;
  return false;
}
bool all(const Set& s) {
  for (char c : s)
    if (!c) return false;
  return true;
}
Set operator&(const Set& a, const Set& b) {
  Set r(a.size());
  for (size_t i = 0; i < a.size(); i++) r[i] = a[i] && b[i];
  return r;
}

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();
constexpr double kEps = std::numeric_limits<double>::epsilon();

// ---- device memory helpers ----------------------------------------------------

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    n = count;
    if (count) hip_ok(hipMalloc((void**)&p, sizeof(T) * count), "hipMalloc");
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// The device iterate of interior_point_optimizer.h:21-90 over the union:
// x (primals and multipliers), bound duals and bound slacks.
struct State {
  double* x = nullptr;
  double* zl = nullptr;
  double* zu = nullptr;
  double* sl = nullptr;
  double* su = nullptr;
  amigo_ipm_state c() const { return amigo_ipm_state{x, zl, zu, sl, su}; }
};

// Small host-to-device uploads (masks, launch lists, per-problem scalars)
// go through pinned staging chunks that are recycled only after a
// synchronization (fetch()), so a device pointer an upload returns stays
// valid, and its bytes unchanged, until the next fetch().  Uploading bytes
// identical to an earlier upload since that fetch returns the earlier
// copy: the uploads are read-only inputs of the kernels.  Results of
// reductions go to a device arena read back in one copy.
class Exchange {
 public:
  void init(size_t up_bytes, size_t res_doubles) {
    chunk_bytes_ = up_bytes;
    res_cap_ = res_doubles;
    add_chunk(chunk_bytes_);
    hip_ok(hipHostMalloc((void**)&res_host_, res_cap_ * sizeof(double)), "hipHostMalloc");
    hip_ok(hipMalloc((void**)&res_dev_, res_cap_ * sizeof(double)), "hipMalloc");
  }
  ~Exchange() {
    for (Chunk& c : chunks_) {
      if (c.host) (void)hipHostFree(c.host);
      if (c.dev) (void)hipFree(c.dev);
    }
    if (res_host_) (void)hipHostFree(res_host_);
    if (res_dev_) (void)hipFree(res_dev_);
  }
  Exchange() = default;
  Exchange(const Exchange&) = delete;
  Exchange& operator=(const Exchange&) = delete;

  const void* upload(const void* src, size_t bytes, hipStream_t s) {
    const unsigned long long h = digest(src, bytes);
    auto range = seen_.equal_range(h);
    for (auto it = range.first; it != range.second; ++it) {
      const Prior& q = it->second;
      if (q.bytes == bytes && std::memcmp(q.host, src, bytes) == 0) return q.dev;
    }
    const size_t need = (bytes + 255) & ~size_t(255);
    while (chunks_[cur_].used + need > chunks_[cur_].cap) {
      // Later chunks are unused until the next fetch(); add one when none
      // is left (a spare chunk too small for this upload is passed over).
      if (cur_ + 1 == chunks_.size()) add_chunk(std::max(chunk_bytes_, need));
      cur_++;
    }
    Chunk& c = chunks_[cur_];
    char* hp = c.host + c.used;
    char* dp = c.dev + c.used;
    std::memcpy(hp, src, bytes);
    hip_ok(hipMemcpyAsync(dp, hp, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    c.used += need;
    seen_.emplace(h, Prior{bytes, hp, dp});
    return dp;
  }
  // A result region of count doubles; index into host() after fetch().
  size_t slot(size_t count) {
    const size_t off = res_used_;
    res_used_ += (count + 31) & ~size_t(31);
    if (res_used_ > res_cap_) raise(-1, "result arena exhausted");
    return off;
  }
  double* dev(size_t off) { return res_dev_ + off; }
  void fetch(hipStream_t s) {
    if (res_used_)
      hip_ok(hipMemcpyAsync(res_host_, res_dev_, res_used_ * sizeof(double),
                            hipMemcpyDeviceToHost, s),
             "hipMemcpyAsync");
    hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
    for (Chunk& c : chunks_) c.used = 0;
    cur_ = 0;
    seen_.clear();
    generation_++;
    fetched_ = res_used_;
    res_used_ = 0;
  }
  double host(size_t off) const { return res_host_[off]; }
  // Changes at every fetch(): device pointers of earlier uploads are void.
  unsigned long long generation() const { return generation_; }

 private:
  struct Chunk {
    char* host = nullptr;
    char* dev = nullptr;
    size_t cap = 0, used = 0;
  };
  struct Prior {
    size_t bytes;
    const char* host;
    const char* dev;
  };
  static unsigned long long digest(const void* src, size_t bytes) {
    // FNV-1a over 8-byte words, then the tail bytes.
    const unsigned char* p = static_cast<const unsigned char*>(src);
    unsigned long long h = 1469598103934665603ull ^ bytes;
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
      unsigned long long w;
      std::memcpy(&w, p + i, 8);
      h = (h ^ w) * 1099511628211ull;
    }
    for (; i < bytes; i++) h = (h ^ p[i]) * 1099511628211ull;
    return h;
  }
  void alloc_chunk(Chunk& c, size_t bytes) {
    c.cap = bytes;
    hip_ok(hipHostMalloc((void**)&c.host, bytes), "hipHostMalloc");
    hip_ok(hipMalloc((void**)&c.dev, bytes), "hipMalloc");
  }
  void add_chunk(size_t bytes) {
    chunks_.emplace_back();
    alloc_chunk(chunks_.back(), bytes);
  }
  std::vector<Chunk> chunks_;
  size_t cur_ = 0, chunk_bytes_ = 0;
  std::unordered_multimap<unsigned long long, Prior> seen_;
  unsigned long long generation_ = 0;
  double* res_host_ = nullptr;
  double* res_dev_ = nullptr;
  size_t res_cap_ = 0, res_used_ = 0, fetched_ = 0;
};

// A reduction's per-problem results, readable after the next fetch().
struct Red {
  size_t off = 0;
  int nv = 4;
  double at(const Exchange& ex, int k, int q) const { return ex.host(off + (size_t)k * nv + q); }
};

enum Phase { PH_INIT, PH_CONV, PH_FACTOR, PH_SOLVE, PH_SEARCH, PH_RESTORE, PH_FINISH, PH_COUNT };

// Host state of one problem.
struct Prob {
  int n_primal = 0, n_dual = 0;
  double mu = 1.0, mu_before = 1.0, tau = 1.0;
  double obj_scale = 1.0;
  int status = AMIGO_IPM_RUNNING;
  bool running = true;
  double prev_res = kInf;
  int acc_count = 0, floor_count = 0, rejections = 0;
  int filter_rejects = 0, filter_resets = 0;
  // the step of the previous iteration, as recorded
  int line_iters = 0;
  double rec_ax = 0.0, rec_az = 0.0;
  long long rec_xi = -1, rec_zi = -1;
  double theta0 = kNaN;  // filter reference theta (theta_min / theta_max)
  Regularization reg;
  double last_dw = 0.0, last_dc = 0.0;
  Filter filter;
  // watchdog (filter_line_search.py:395-438)
  bool wd_on = false;
  int wd_shortened = 0, wd_trials = 0;
  double wd_theta = 0.0, wd_barr = 0.0, wd_dphi = 0.0;
  // this iteration
  double res_norm = 0.0, ax = 0.0, az = 0.0;
  long long xi = -1, zi = -1;
  double ir_ratio = 0.0;
  double ls_ratio = 0.0;  // step fraction the line search settled on
  int iter = 0;           // main iterations completed
  // feasibility restoration by Newton steps at raised barrier parameters
  // (feasibility_restoration.py:102-172), one restoration iteration per round
  bool restoring = false, restore_ok = false;
  std::vector<double> rlevels;
  int rlevel = 0, riter = 0, rper_level = 0;
  double rsaved_mu = 0.0, rtheta = 0.0, rtheta_start = 0.0, rtarget = 0.0;
  double rax = 0.0, raz = 0.0;
  std::vector<amigo_ipm_iteration> records;
  std::vector<int> inertia, solver_inertia;
  int fp64_fallbacks = 0, ir_steps = 0, restorations = 0;
  double t_end = 0.0;
};

// One problem's filter line search (filter_line_search.py:54-280).
struct Search {
  bool on = false;
  double alpha_x = 1.0, alpha_z = 1.0;
  double ref_theta = 0.0, ref_barr = 0.0, ref_dphi = 0.0;
  double theta_min = 0.0, theta_max = 0.0, alpha_min = 0.0;
  double alpha = 1.0;
  int steps = 0;
  bool last_filter_reject = false;
  // second-order correction
  bool soc = false;
  int soc_count = 0;
  double alpha_soc = 0.0, theta_soc_old = 0.0, a_test = 0.0, soc_ax = 0.0, soc_az = 0.0;
  double trial_theta = 0.0, trial_barr = 0.0;
  // result
  bool accepted = false, filter_reject = false;
  double ratio = 0.0;
  int iters = 0;
};

}  // namespace

struct amigo_ipm_driver {
  amigo_nlp_desc d{};
  std::vector<amigo_nlp_group> groups;
  std::vector<int> xseg, pseg, cseg;
  int K = 0;
  hipStream_t s = nullptr;
  Exchange ex;
  amigo_ipm_options o{};
  std::vector<Prob> P;
  Set ldl_set;  // the set the factorization's mask holds (empty: not known)
  bool scaling = false;
  double phase_s[PH_COUNT] = {0, 0, 0, 0, 0, 0, 0};
  double host_s = 0.0;
  std::chrono::steady_clock::time_point t0;
  amigo_ipm_done_fn on_done = nullptr;
  void* ctx = nullptr;

  // buffers
  State vars, upd, temp, upd_keep, wd_upd;
  DevBuf<double> bz[5 * 4];
  DevBuf<double> grad, res, diag, dbase, reg, px, sol, rhs, corr, kpx, csoc, res_orig, px_orig,
      wd_x, wd_px, xobj, tmp_n, sv;
  DevBuf<double> dzl, dzu, ezl, ezu, tmp_p;
  double* dc = nullptr;  // constraint scales (caller's)
  long long n_factor = 0, n_solve = 0;
  bool captured = false;
  DevBuf<double> ws;
  DevBuf<int> dev_keep;

  // ---- set-up -------------------------------------------------------------
  void setup() {
    K = d.K;
    xseg.assign(d.xseg_host, d.xseg_host + K + 1);
    pseg.assign(d.pseg_host, d.pseg_host + K + 1);
    cseg.assign(d.cseg_host, d.cseg_host + K + 1);
    groups.assign(d.groups, d.groups + d.ngroups);
    const size_t n = d.n, np = d.n_primal, nc = d.n_constraints;
    vars.x = d.x;
    int b = 0;
    auto mk = [&](State& st, bool with_x, bool slacks) {
      if (with_x) {
        bz[b].alloc(n);
        st.x = bz[b++].p;
      }
      bz[b].alloc(np);
      st.zl = bz[b++].p;
      bz[b].alloc(np);
      st.zu = bz[b++].p;
      if (slacks) {
        bz[b].alloc(np);
        st.sl = bz[b++].p;
        bz[b].alloc(np);
        st.su = bz[b++].p;
      }
    };
    vars.zl = d.zl;
    vars.zu = d.zu;
    vars.sl = d.sl;
    vars.su = d.su;
    mk(upd, true, false);
    mk(temp, true, true);
    mk(upd_keep, true, false);
    mk(wd_upd, true, false);
    for (DevBuf<double>* v : {&grad, &res, &diag, &dbase, &reg, &px, &sol, &rhs, &corr, &kpx,
                              &csoc, &res_orig, &px_orig, &wd_x, &wd_px, &xobj, &tmp_n, &sv})
      v->alloc(n);
    for (DevBuf<double>* v : {&dzl, &dzu, &ezl, &ezu, &tmp_p}) v->alloc(std::max<size_t>(np, 1));
    dc = d.dc;
    ws.alloc(4 * (size_t)std::max(d.nblk_total, d.nblk_primal) + 16);
    dev_keep.alloc(K);
    ex.init(size_t(8) << 20, (size_t)K * 4 * 160 + 4096);
    P.assign(K, Prob());
    for (int k = 0; k < K; k++) {
      P[k].n_primal = pseg[k + 1] - pseg[k];
      P[k].n_dual = cseg[k + 1] - cseg[k];
    }
    // Host copies of the block tables and the Hessian segments (the launch
    // lists of small sets are built from them).
    hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    auto down = [&](auto& h, const auto* dv, size_t count) {
      h.resize(count);
      hip_ok(hipMemcpy(h.data(), dv, sizeof(h[0]) * count, hipMemcpyDeviceToHost), "hipMemcpy");
    };
    down(first_t, d.first_total, K);
    down(count_t, d.count_total, K);
    down(first_p, d.first_primal, K);
    down(count_p, d.count_primal, K);
    down(hseg, d.hseg64, (size_t)K + 1);
  }

  // ---- sets, masks, scalars ------------------------------------------------
  // A set of problems as the kernels take it: a mask (null = every problem)
  // and, when the set is at most a quarter of the batch, lists of its
  // problems and of their blocks, so that launches cover only their work
  // (with the results of the masked full launches).  Built once per set
  // between two fetches.
  struct Sel {
    const int* mask = nullptr;
    const int* probs = nullptr;
    int nprobs = 0;
    const int* blk_t = nullptr;
    int nblk_t = 0;
    const int* blk_p = nullptr;
    int nblk_p = 0;
    int max_x = 0, max_c = 0;
    long long max_h = 0;
  };
  struct SelEntry {
    unsigned long long gen;
    Set set;
    Sel sel;
  };
  static constexpr int kListFraction = 4;
  std::vector<SelEntry> sel_cache;
  std::vector<int> first_t, count_t, first_p, count_p;
  std::vector<long long> hseg;

  Sel select(const Set& A) {
    if (all(A)) return Sel{};
    const unsigned long long g = ex.generation();
    for (const SelEntry& e : sel_cache)
      if (e.gen == g && e.set == A) return e.sel;
    std::vector<int> h(K);
    int na = 0;
    for (int k = 0; k < K; k++) {
      h[k] = A[k] ? 1 : 0;
      na += h[k];
    }
    Sel r;
    if (na > 0 && na * kListFraction <= K) {
      // one upload: mask, problems, total-grid blocks, primal-grid blocks
      for (int k = 0; k < K; k++)
        if (A[k]) {
          h.push_back(k);
          r.max_x = std::max(r.max_x, xseg[k + 1] - xseg[k]);
          r.max_c = std::max(r.max_c, cseg[k + 1] - cseg[k]);
          r.max_h = std::max(r.max_h, hseg[k + 1] - hseg[k]);
        }
      for (int k = 0; k < K; k++)
        if (A[k])
          for (int b = first_t[k]; b < first_t[k] + count_t[k]; b++) h.push_back(b);
      const size_t nbt = h.size() - K - na;
      for (int k = 0; k < K; k++)
        if (A[k])
          for (int b = first_p[k]; b < first_p[k] + count_p[k]; b++) h.push_back(b);
      const size_t nbp = h.size() - K - na - nbt;
      const int* dv = (const int*)ex.upload(h.data(), sizeof(int) * h.size(), s);
      r.mask = dv;
      r.probs = dv + K;
      r.nprobs = na;
      r.blk_t = r.probs + na;
      r.nblk_t = int(nbt);
      r.blk_p = r.blk_t + nbt;
      r.nblk_p = int(nbp);
    } else {
      r.mask = (const int*)ex.upload(h.data(), sizeof(int) * K, s);
    }
    sel_cache.erase(std::remove_if(sel_cache.begin(), sel_cache.end(),
                                   [&](const SelEntry& e) { return e.gen != g; }),
                    sel_cache.end());
    if (sel_cache.size() >= 8) sel_cache.erase(sel_cache.begin());
    sel_cache.push_back(SelEntry{g, A, r});
    return r;
  }
  const int* mask(const Set& m) { return select(m).mask; }
  // Rows of the listed problems (x-length and constraint segments).
  amigo_amd::drv::Rows rows_x(const Sel& q) const {
    return q.probs ? amigo_amd::drv::Rows{q.probs, q.nprobs, d.xseg, q.max_x}
                   : amigo_amd::drv::Rows{};
  }
  amigo_amd::drv::Rows rows_c(const Sel& q) const {
    return q.probs ? amigo_amd::drv::Rows{q.probs, q.nprobs, d.cseg, q.max_c}
                   : amigo_amd::drv::Rows{};
  }
  const double* scalars(const std::vector<double>& v) {
    return (const double*)ex.upload(v.data(), sizeof(double) * K, s);
  }
  std::vector<double> per(double Prob::*f) const {
    std::vector<double> v(K);
    for (int k = 0; k < K; k++) v[k] = P[k].*f;
    return v;
  }
  amigo_ipm_batch batch(const int* active) const {
    amigo_ipm_batch B;
    B.num_problems = K;
    B.pseg = d.pseg;
    B.cseg = d.cseg;
    B.xseg = d.xseg;
    B.nblk_total = d.nblk_total;
    B.blk_total = d.blk_total;
    B.first_total = d.first_total;
    B.count_total = d.count_total;
    B.nblk_primal = d.nblk_primal;
    B.blk_primal = d.blk_primal;
    B.first_primal = d.first_primal;
    B.count_primal = d.count_primal;
    B.active = active;
    B.nlaunch_total = B.nlaunch_primal = B.nlaunch_problems = 0;
    B.launch_total = B.launch_primal = B.launch_problems = nullptr;
    return B;
  }
  amigo_ipm_batch batch(const Sel& q) const {
    amigo_ipm_batch B = batch(q.mask);
    if (q.probs) {
      B.nlaunch_total = q.nblk_t;
      B.launch_total = q.blk_t;
      B.nlaunch_primal = q.nblk_p;
      B.launch_primal = q.blk_p;
      B.nlaunch_problems = q.nprobs;
      B.launch_problems = q.probs;
    }
    return B;
  }
  amigo_ipm_batch batch(const Set& A) { return batch(select(A)); }
  amigo_ipm_problem info() const {
    return amigo_ipm_problem{d.n_primal, d.n_constraints, d.pidx, d.cidx, d.lbx, d.ubx, d.lbh};
  }
  Red red_slot(int nv) {
    Red r;
    r.nv = nv;
    r.off = ex.slot((size_t)K * nv);
    return r;
  }
  double at(const Red& r, int k, int q = 0) const { return r.at(ex, k, q); }
  void fetch() { ex.fetch(s); }

  // ---- problem evaluation (optimization_problem.h:718-818) -----------------
  // alpha < 0: each problem's objective scale; otherwise that scale for all.
  void eval(const Set& A, const double* x, double alpha, bool want_g, bool want_h, bool want_l) {
    const int* m = mask(A);
    for (const amigo_nlp_group& g : groups) {
      double* gb = want_g ? d.gbuf + g.goff : nullptr;
      double* hb = want_h ? d.hbuf + g.hoff : nullptr;
      double* lp = want_l ? d.lpart + g.boff : nullptr;
      const double* ab = alpha < 0.0 ? g.alpha_blk : nullptr;
      const double a = alpha < 0.0 ? 1.0 : alpha;
      asm_ok(amigo_elements_eval_masked(g.etype, g.E, g.idx, g.data, g.ndata, x, a, ab,
                                        m ? g.blk_prob : nullptr, m, gb, hb, lp, s),
             "element evaluation");
    }
  }
  // grad = [alpha grad f + J^T lambda; c(x)] of the problems of A, with the
  // constraint rows scaled (newton_direction.py:17-27).
  void gradient(const Set& A, const double* x, double* g) {
    eval(A, x, -1.0, true, false, false);
    const Sel q = select(A);
    if (!q.mask)
      asm_ok(amigo_gather_sum(d.n, d.g_ptr, d.g_src, d.gbuf, g, 0, s), "gradient gather");
    else if (q.probs)
      asm_ok(amigo_gather_sum_listed(d.g_ptr, d.g_src, d.gbuf, g, 0, d.xseg64, q.probs, q.nprobs,
                                     q.max_x, s),
             "gradient gather");
    else
      asm_ok(amigo_gather_sum_masked(d.n, d.g_ptr, d.g_src, d.gbuf, g, 0, d.xseg64, K, q.mask, s),
             "gradient gather");
    if (scaling) {
      const amigo_ipm_batch B = batch(q);
      ipm_ok(amigo_ipm_batch_scale_rows(&B, d.cidx, dc, g, 0, s), "gradient scaling");
    }
  }
  // KKT values of the problems of A, with D H D scaling
  // (optimizer.h:573-584) once the NLP scaling is on.
  void hessian(const Set& A, const double* x, double alpha, bool scale) {
    eval(A, x, alpha, false, true, false);
    const Sel q = select(A);
    const int* m = q.mask;
    if (!m)
      asm_ok(amigo_gather_sum(d.nnz, d.h_ptr, d.h_src, d.hbuf, d.hess, 0, s), "Hessian gather");
    else if (q.probs)
      asm_ok(amigo_gather_sum_listed(d.h_ptr, d.h_src, d.hbuf, d.hess, 0, d.hseg64, q.probs,
                                     q.nprobs, q.max_h, s),
             "Hessian gather");
    else
      asm_ok(amigo_gather_sum_masked(d.nnz, d.h_ptr, d.h_src, d.hbuf, d.hess, 0, d.hseg64, K, m,
                                     s),
             "Hessian gather");
    if (scale && scaling) {
      if (!m)
        ipm_ok(amigo_csr_scale_symmetric(d.n, d.rowp, d.cols, sv.p, d.hess, s), "D H D");
      else
        ipm_ok(amigo_csr_scale_symmetric_masked(d.n, d.rowp, d.cols, sv.p, d.hess, d.row_prob, m,
                                                s),
               "D H D");
    }
  }
  // alpha f(x) of each problem (the Lagrangian with the multipliers
  // zeroed, filter_line_search.py:19-46).
  Red objective(const Set& A, const double* x) {
    const Sel q = select(A);
    hip_ok(amigo_amd::drv::copy_rows(d.n, d.row_prob, q.mask, x, xobj.p, s, rows_x(q)), "copy");
    hip_ok(amigo_amd::drv::zero_rows(d.n_constraints, d.cidx_global, xobj.p, s, rows_c(q)),
           "zero rows");
    eval(A, xobj.p, -1.0, false, false, true);
    Red r = red_slot(1);
    asm_ok(amigo_sum_partials_batch(d.lpart, K, d.lptr, d.lidx, ex.dev(r.off), nullptr, s),
           "Lagrangian sums");
    return r;
  }

  // ---- interior-point kernels ----------------------------------------------
  Red reduce(int kind, const Set& A, const State& v, const double* mu, const double* a,
             const double* b = nullptr, const double* c = nullptr) {
    const amigo_ipm_batch B = batch(A);
    const amigo_ipm_problem I = info();
    const amigo_ipm_state S = v.c();
    Red r = red_slot(4);
    ipm_ok(amigo_ipm_batch_reduce(kind, &B, &I, &S, mu, a, b, c, ws.p, ex.dev(r.off), nullptr, s),
           "reduction");
    return r;
  }
  const double* zeros() {
    std::vector<double> z(K, 0.0);
    return scalars(z);
  }
  // condensed residual (backend.h:202-266); out[0] = ||res||^2
  Red residual(const Set& A, const State& v, const double* g, double* out) {
    const amigo_ipm_batch B = batch(A);
    const amigo_ipm_problem I = info();
    const amigo_ipm_state S = v.c();
    Red r = red_slot(3);
    ipm_ok(amigo_ipm_batch_residual(&B, &I, &S, scalars(per(&Prob::mu)), g, out, ws.p,
                                    ex.dev(r.off), nullptr, s),
           "residual");
    return r;
  }
  void sigma(const Set& A, const State& v, double* out) {
    const amigo_ipm_batch B = batch(A);
    const amigo_ipm_problem I = info();
    const amigo_ipm_state S = v.c();
    ipm_ok(amigo_ipm_batch_diagonal(&B, &I, &S, out, s), "diagonal");
  }
  // update = (px, dzl(px), dzu(px)) (backend.h:302-321)
  void direction(const Set& A, const double* p) {
    const Sel q = select(A);
    hip_ok(amigo_amd::drv::copy_rows(d.n, d.row_prob, q.mask, p, upd.x, s, rows_x(q)), "copy");
    const amigo_ipm_batch B = batch(q);
    const amigo_ipm_problem I = info();
    const amigo_ipm_state S = vars.c();
    ipm_ok(amigo_ipm_batch_bound_dual_step(&B, &I, &S, scalars(per(&Prob::mu)), p, upd.zl, upd.zu,
                                           s),
           "bound dual step");
  }
  Red max_step(const Set& A, const std::vector<double>& tau) {
    const amigo_ipm_batch B = batch(A);
    const amigo_ipm_problem I = info();
    const amigo_ipm_state S = vars.c();
    Red r = red_slot(4);
    ipm_ok(amigo_ipm_batch_max_step(&B, &I, &S, scalars(tau), upd.x, upd.zl, upd.zu, ws.p,
                                    ex.dev(r.off), nullptr, s),
           "max step");
    return r;
  }
  // temp = vars + (ax dx, az dz) (backend.h:375-392)
  void trial(const Set& A, const std::vector<double>& ax, const std::vector<double>& az) {
    const amigo_ipm_batch B = batch(A);
    const amigo_ipm_problem I = info();
    const amigo_ipm_state S = vars.c(), T = temp.c();
    ipm_ok(amigo_ipm_batch_apply_step(&B, &I, &S, scalars(ax), scalars(az), upd.x, upd.zl, upd.zu,
                                      &T, s),
           "apply step");
  }
  void copy_state(const Set& A, const State& dst, const State& src, int what) {
    const amigo_ipm_batch B = batch(A);
    const amigo_ipm_state D = dst.c(), S = src.c();
    ipm_ok(amigo_ipm_batch_copy_state(&B, &D, &S, what, s), "copy state");
  }
  void copy_x(const Set& A, double* dst, const double* src) {
    const Sel q = select(A);
    hip_ok(amigo_amd::drv::copy_rows(d.n, d.row_prob, q.mask, src, dst, s, rows_x(q)), "copy");
  }

  // ---- factor and solves ---------------------------------------------------
  // The factorization's problem mask (it builds launch lists and
  // synchronizes, so it is set only when the set changes).
  void ldl_mask(const Set& A) {
    if (A == ldl_set) return;
    if (all(A)) {
      ldl_ok(amigo_ldl_set_batch_mask(d.ldl, nullptr, s), "set_batch_mask");
    } else {
      std::vector<int> h(K);
      for (int k = 0; k < K; k++) h[k] = A[k] ? 1 : 0;
      ldl_ok(amigo_ldl_set_batch_mask(d.ldl, h.data(), s), "set_batch_mask");
    }
    ldl_set = A;
  }
  // Factor the shifted KKT of the problems of A (diag is the shift);
  // per-problem inertia into pos/neg/nan.
  void factor(const Set& A, int precision, std::vector<int>& pos, std::vector<int>& neg,
              std::vector<int>& nan) {
    ldl_mask(A);
    ldl_ok(amigo_ldl_factor_shifted(d.ldl, d.hess, diag.p, precision, s), "factor");
    n_factor++;
    pos.assign(K, 0);
    neg.assign(K, 0);
    nan.assign(K, 0);
    std::vector<int> zero(K);
    ldl_ok(amigo_ldl_get_batch_inertia(d.ldl, pos.data(), neg.data(), zero.data(), nan.data()),
           "batch inertia");
  }
  // out = K^-1 b for the problems of A (other rows of out: b's).
  void solve(const Set& A, const double* b, double* out) {
    if (out != b)
      hip_ok(hipMemcpyAsync(out, b, sizeof(double) * d.n, hipMemcpyDeviceToDevice, s), "copy");
    ldl_mask(A);
    ldl_ok(amigo_ldl_solve(d.ldl, 1, out, d.n, s), "solve");
    n_solve++;
  }

  void tic(std::chrono::steady_clock::time_point& t) {
    if (o.profile) {
      (void)hipStreamSynchronize(s);
      t = std::chrono::steady_clock::now();
    }
  }
  void toc(int ph, std::chrono::steady_clock::time_point& t) {
    if (!o.profile) return;
    (void)hipStreamSynchronize(s);
    const auto now = std::chrono::steady_clock::now();
    phase_s[ph] += std::chrono::duration<double>(now - t).count();
    t = now;
  }
  double elapsed() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // ---- the solve -------------------------------------------------------------
  void reset();
  void initialize();
  void least_squares_multipliers();
  void affine_multipliers();
  void nlp_scaling();
  std::vector<char> factorize(const Set& A);
  void solve_with_mu(const Set& B);
  std::vector<double> refine(const Set& B);
  void barrier_update(const Set& A);
  void line_search_with_watchdog(const Set& B, std::vector<Search>& out);
  void line_search(const Set& S, std::vector<Search>& ls, const std::vector<double>& phi_cur,
                   const Set& wd_ref);
  void accept_checks(Prob& p, Search& L, double tb, double tt, double a, bool& acc, bool& frej) const;
  void update_filter(Prob& p, const Search& L, double tb, double a) const;
  Set restoration_begin(const Set& R);
  Set restoration_search(const Set& RB);
  void finish_step(const Set& done, const Set& accepted);
  double tau_of(double mu) const {
    return o.adaptive_tau ? std::max(o.tau_min, 1.0 - mu) : o.fraction_to_boundary;
  }
  void increase_barrier(Prob& p) const {
    if (p.rejections >= o.max_consecutive_rejections) {
      const double bumped = std::min(p.mu * o.barrier_increase_factor, o.initial_barrier_param);
      if (bumped > p.mu) p.mu = bumped;
      p.rejections = 0;
    }
  }
  void run();
};

namespace {

// (avg, xi) of optimizer.h:287-300 from the [sum, count, min] reduction.
void complementarity(double sum, double cnt, double mn, double& avg, double& xi) {
  avg = cnt > 0 ? sum / cnt : 0.0;
  if (cnt == 0) mn = std::numeric_limits<double>::max();
  xi = avg > 0 ? std::max(0.0, std::min(1.0, mn / avg)) : 1.0;
}

// (s_d, s_c) of optimality_scaling.py:19-48 from [sum|z|, sum|y|, n_bounds, n_constr].
void optimality_scaling(double z, double y, double nb, double nc, double& s_d, double& s_c) {
  const double s_max = 100.0;
  s_c = nb == 0 ? 1.0 : std::max(s_max, z / nb) / s_max;
  const double n_all = nc + nb;
  s_d = n_all == 0 ? 1.0 : std::max(s_max, (y + z) / n_all) / s_max;
}

}  // namespace

void amigo_ipm_driver::reset() {
  const size_t n = d.n, np = d.n_primal, nc = d.n_constraints;
  hip_ok(hipMemcpyAsync(vars.x, d.x0, sizeof(double) * n, hipMemcpyDeviceToDevice, s), "copy");
  hip_ok(hipMemcpyAsync(d.lbx, d.lbx0, sizeof(double) * np, hipMemcpyDeviceToDevice, s), "copy");
  hip_ok(hipMemcpyAsync(d.ubx, d.ubx0, sizeof(double) * np, hipMemcpyDeviceToDevice, s), "copy");
  hip_ok(hipMemcpyAsync(d.lbh, d.lbh0, sizeof(double) * nc, hipMemcpyDeviceToDevice, s), "copy");
  hip_ok(amigo_amd::drv::fill(np, 0.0, vars.zl, s), "fill");
  hip_ok(amigo_amd::drv::fill(np, 0.0, vars.zu, s), "fill");
  hip_ok(amigo_amd::drv::fill(np, 1.0, vars.sl, s), "fill");
  hip_ok(amigo_amd::drv::fill(np, 1.0, vars.su, s), "fill");
  hip_ok(amigo_amd::drv::fill(nc, 1.0, dc, s), "fill");
  scaling = false;
  for (amigo_nlp_group& g : groups) hip_ok(amigo_amd::drv::fill(g.E / 256, 1.0, g.alpha_blk, s), "fill");
  for (int k = 0; k < K; k++) {
    Prob fresh;
    fresh.n_primal = P[k].n_primal;
    fresh.n_dual = P[k].n_dual;
    P[k] = std::move(fresh);
    P[k].mu = o.initial_barrier_param;
    P[k].filter.configure(o.filter_gamma_phi, o.filter_gamma_theta);
  }
  for (double& t : phase_s) t = 0.0;
  host_s = 0.0;
  n_factor = n_solve = 0;
  captured = false;
}

// NLP scaling (optimizer.h:502-563): objective scale from the primal
// gradient's inf-norm, constraint row scales from the KKT at lambda = 0.
void amigo_ipm_driver::nlp_scaling() {
  const Set A(K, 1);
  const double gmax_allowed = o.nlp_scaling_max_gradient, min_value = 1e-8;
  hip_ok(amigo_amd::drv::gather(d.n_primal, d.pidx_global, grad.p, tmp_p.p, s), "gather");
  Red gm = red_slot(1);
  ipm_ok(amigo_segment_absmax(tmp_p.p, d.pseg, K, ex.dev(gm.off), nullptr, s), "absmax");
  hessian(A, vars.x, 1.0, false);
  ipm_ok(amigo_ipm_constraint_scaling(d.n_constraints, d.cidx_global32, d.rowp, d.hess,
                                      gmax_allowed, min_value, dc, s),
         "constraint scaling");
  hip_ok(amigo_amd::drv::constraint_scale_vectors(d.n, d.n_constraints, d.cidx_global, dc,
                                                  d.lbh0, sv.p, d.lbh, s),
         "scale vectors");
  fetch();
  std::vector<double> alpha_k(K);
  for (int k = 0; k < K; k++) {
    const double gmax = at(gm, k);
    double sc = (P[k].n_primal > 0 && gmax > gmax_allowed) ? gmax_allowed / gmax : 1.0;
    if (sc < min_value) sc = min_value;
    P[k].obj_scale = sc;
    alpha_k[k] = sc;
  }
  for (amigo_nlp_group& g : groups) {
    const int nb = g.E / 256;
    std::vector<double> a(nb);
    for (int b = 0; b < nb; b++) a[b] = alpha_k[g.blk_prob_host[b]];
    if (nb)
      hip_ok(hipMemcpyAsync(g.alpha_blk, ex.upload(a.data(), sizeof(double) * nb, s),
                            sizeof(double) * nb, hipMemcpyDeviceToDevice, s),
             "alpha upload");
  }
  scaling = true;
}

// Least-squares multipliers (multiplier_initialization.py:19-58): solve
// [I J^T; J 0] [dx; lambda] = [-(grad f - z); 0] and keep lambda unless it
// exceeds 1e3 in magnitude.
void amigo_ipm_driver::least_squares_multipliers() {
  const Set A(K, 1);
  const amigo_ipm_batch B = batch(nullptr);
  const amigo_ipm_problem I = info();
  const amigo_ipm_state S = vars.c();
  ipm_ok(amigo_ipm_batch_dual_residual_vector(&B, &I, &S, grad.p, res.p, s), "dual residual");
  hip_ok(amigo_amd::drv::negate(d.n, res.p, s), "negate");
  hip_ok(amigo_amd::drv::unit_primal(d.n, d.is_mult, diag.p, s), "unit primal");
  hessian(A, vars.x, 0.0, true);
  std::vector<int> pos, neg, nan;
  factor(A, o.precision, pos, neg, nan);
  for (int k = 0; k < K; k++)
    if (!nan[k]) {
      P[k].solver_inertia.push_back(pos[k]);
      P[k].solver_inertia.push_back(neg[k]);
    }
  solve(A, res.p, px.p);
  hip_ok(amigo_amd::drv::gather(d.n_constraints, d.cidx_global, px.p, tmp_n.p, s), "gather");
  Red lm = red_slot(1);
  ipm_ok(amigo_segment_absmax(tmp_n.p, d.cseg, K, ex.dev(lm.off), nullptr, s), "absmax");
  fetch();
  std::vector<int> keep(K);
  for (int k = 0; k < K; k++) keep[k] = !(nan[k] || (P[k].n_dual > 0 && at(lm, k) > 1e3));
  const int* kd = (const int*)ex.upload(keep.data(), sizeof(int) * K, s);
  hip_ok(amigo_amd::drv::set_multipliers(d.n_constraints, d.cidx_global, d.row_prob, kd, nullptr,
                                         px.p, vars.x, s),
         "set multipliers");
}

// Multipliers from the affine-scaling step (multiplier_initialization.py:
// 60-93); the starting barrier parameter becomes the average
// complementarity, floored at the initial one.
void amigo_ipm_driver::affine_multipliers() {
  if (K != 1) raise(-1, "affine-step multiplier initialization runs one problem at a time");
  const Set A(K, 1);
  gradient(A, vars.x, grad.p);
  const double* z = zeros();
  {
    const amigo_ipm_batch B = batch(nullptr);
    const amigo_ipm_problem I = info();
    const amigo_ipm_state S = vars.c();
    Red r = red_slot(3);
    ipm_ok(amigo_ipm_batch_residual(&B, &I, &S, z, grad.p, res.p, ws.p, ex.dev(r.off), nullptr, s),
           "residual");
  }
  sigma(A, vars, diag.p);
  hessian(A, vars.x, -1.0, true);
  std::vector<int> pos, neg, nan;
  factor(A, o.precision, pos, neg, nan);
  if (!nan[0]) {
    P[0].solver_inertia.push_back(pos[0]);
    P[0].solver_inertia.push_back(neg[0]);
  }
  solve(A, res.p, px.p);
  hip_ok(amigo_amd::drv::copy_rows(d.n, d.row_prob, nullptr, px.p, upd.x, s), "copy");
  {
    const amigo_ipm_batch B = batch(nullptr);
    const amigo_ipm_problem I = info();
    const amigo_ipm_state S = vars.c();
    ipm_ok(amigo_ipm_batch_bound_dual_step(&B, &I, &S, z, px.p, upd.zl, upd.zu, s), "bound dual step");
  }
  std::vector<int> keep(K, 1);
  const int* kd = (const int*)ex.upload(keep.data(), sizeof(int) * K, s);
  hip_ok(amigo_amd::drv::set_multipliers(d.n_constraints, d.cidx_global, d.row_prob, kd, nullptr,
                                         upd.x, vars.x, s),
         "set multipliers");
  const amigo_ipm_problem I = info();
  const amigo_ipm_state S = vars.c();
  ipm_ok(amigo_ipm_affine_start_point(&I, &S, P[0].mu, upd.zl, upd.zu, vars.zl, vars.zu, s),
         "affine start point");
  Red c = reduce(0, A, vars, z, nullptr);
  fetch();
  double avg, xi;
  complementarity(at(c, 0, 0), at(c, 0, 1), at(c, 0, 2), avg, xi);
  P[0].mu = std::max(avg, P[0].mu);
}

// The starting point (iterate_initialization.py:16-69).
void amigo_ipm_driver::initialize() {
  const Set A(K, 1);
  const amigo_ipm_batch B = batch(nullptr);
  const amigo_ipm_state S = vars.c();
  ipm_ok(amigo_ipm_batch_initialize(&B, d.lbx0, d.ubx0, d.pidx, 1e-8, o.constr_viol_tol, 1e-2,
                                    1e-2, d.lbx, d.ubx, &S, 1, 0, s),
         "relax bounds");
  hip_ok(amigo_amd::drv::zero_rows(d.n_constraints, d.cidx_global, vars.x, s), "zero multipliers");
  ipm_ok(amigo_ipm_batch_initialize(&B, d.lbx0, d.ubx0, d.pidx, 0.0, 0.0, 1e-2, 1e-2, d.lbx, d.ubx,
                                    &S, 0, 1, s),
         "initial point");
  gradient(A, vars.x, grad.p);
  if (o.nlp_scaling_max_gradient > 0) {
    nlp_scaling();
    gradient(A, vars.x, grad.p);
  }
  if (o.init_affine_step_multipliers) {
    least_squares_multipliers();
    affine_multipliers();
  } else if (o.init_least_squares_multipliers) {
    least_squares_multipliers();
  }
  gradient(A, vars.x, grad.p);
}

// Inertia-corrected factorization of the Newton system of the problems of
// A (newton_direction.py:29-131 with inertia_correction.py:261-393): the
// Hessian at the current x, Sigma (in dbase) plus the perturbations of the
// regularization, refactored until the inertia is (n_primal, n_dual, 0).
std::vector<char> amigo_ipm_driver::factorize(const Set& A) {
  hessian(A, vars.x, -1.0, true);
  {
    const Sel q = select(A);
    hip_ok(amigo_amd::drv::regularized_base(d.n, d.is_mult, d.row_prob, q.mask, dbase.p, 1e-12,
                                            reg.p, s, rows_x(q)),
           "regularized diagonal");
  }
  std::vector<char> ok(K, 0);
  Set pending = A;
  for (int k = 0; k < K; k++)
    if (A[k]) P[k].reg.begin_system();
  std::vector<int> pos, neg, nan;
  const int itol = o.inertia_tolerance;
  for (int attempt = 0; attempt <= o.max_inertia_corrections && any(pending); attempt++) {
    std::vector<double> dw(K, 0.0), dcv(K, 0.0);
    for (int k = 0; k < K; k++) {
      dw[k] = P[k].reg.dw();
      dcv[k] = P[k].reg.dc();
    }
    const Sel qp = select(pending);
    hip_ok(amigo_amd::drv::diagonal_shift(d.n, d.is_mult, d.row_prob, qp.mask, reg.p, scalars(dw),
                                          scalars(dcv), diag.p, s, rows_x(qp)),
           "diagonal shift");
    if (!captured && (d.capture_hess || d.capture_shift)) {
      if (d.capture_hess)
        hip_ok(hipMemcpyAsync(d.capture_hess, d.hess, sizeof(double) * d.nnz,
                              hipMemcpyDeviceToDevice, s),
               "capture");
      if (d.capture_shift)
        hip_ok(hipMemcpyAsync(d.capture_shift, diag.p, sizeof(double) * d.n,
                              hipMemcpyDeviceToDevice, s),
               "capture");
      captured = true;
    }
    factor(pending, o.precision, pos, neg, nan);
    for (int k = 0; k < K; k++) {
      if (!pending[k]) continue;
      Prob& p = P[k];
      const bool singular = nan[k] > 0;
      if (!singular) {
        p.inertia.push_back(pos[k]);
        p.inertia.push_back(neg[k]);
        p.solver_inertia.push_back(pos[k]);
        p.solver_inertia.push_back(neg[k]);
      }
      const bool good = !singular && std::abs(pos[k] - p.n_primal) <= itol &&
                        std::abs(neg[k] - p.n_dual) <= itol &&
                        pos[k] + neg[k] >= p.n_primal + p.n_dual - itol;
      if (good) {
        ok[k] = 1;
        pending[k] = 0;
        p.last_dw = p.reg.dw();
        p.last_dc = p.reg.dc();
        continue;
      }
      const Regularization::Outcome why =
          singular ? Regularization::Outcome::Singular
                   : (neg[k] < p.n_dual ? Regularization::Outcome::TooFewNegative
                                        : Regularization::Outcome::WrongInertia);
      if (!p.reg.next(why, p.n_dual > 0)) {
        pending[k] = 0;
        p.last_dw = p.reg.dw();
        p.last_dc = p.reg.dc();
      }
    }
  }
  for (int k = 0; k < K; k++)
    if (pending[k]) {
      P[k].last_dw = P[k].reg.dw();
      P[k].last_dc = P[k].reg.dc();
    }
  return ok;
}

// Iterative refinement on the unreduced system (linear_solver.py:109-264)
// of the problems of B; px holds the first solution, rhs the condensed
// right-hand side.  Returns each problem's last residual ratio.
std::vector<double> amigo_ipm_driver::refine(const Set& B) {
  std::vector<double> final_ratio(K, 0.0);
  const double* mu = scalars(per(&Prob::mu));
  const amigo_ipm_problem I = info();
  const amigo_ipm_state S = vars.c();
  Red rn = red_slot(1);
  {
    const amigo_ipm_batch BB = batch(B);
    ipm_ok(amigo_ipm_batch_ir_init(&BB, &I, &S, mu, grad.p, rhs.p, px.p, dzl.p, dzu.p, ws.p,
                                   ex.dev(rn.off), nullptr, s),
           "refinement init");
  }
  fetch();
  std::vector<double> nrm_rhs(K), ratio_old(K, 1e30);
  for (int k = 0; k < K; k++) nrm_rhs[k] = at(rn, k);
  Set R = B;
  for (int step = 0; step < 10 && any(R); step++) {
    const Sel q = select(R);
    hip_ok(amigo_amd::drv::kkt_matvec(d.n, d.rowp, d.cols, d.hess, diag.p, px.p, kpx.p, d.row_prob,
                                      q.mask, s, rows_x(q)),
           "KKT product");
    const amigo_ipm_batch BB = batch(q);
    Red rr = red_slot(2);
    ipm_ok(amigo_ipm_batch_ir_residual(&BB, &I, &S, scalars(per(&Prob::mu)), grad.p, rhs.p, kpx.p,
                                       px.p, dzl.p, dzu.p, res.p, ezl.p, ezu.p, ws.p,
                                       ex.dev(rr.off), nullptr, s),
           "refinement residual");
    fetch();
    for (int k = 0; k < K; k++) {
      if (!R[k]) continue;
      const double resid = at(rr, k, 0), sol_norm = at(rr, k, 1);
      const double denom = std::min(sol_norm, 1e6 * nrm_rhs[k]) + nrm_rhs[k];
      const double ratio = denom > 0 ? resid / denom : 0.0;
      final_ratio[k] = ratio;
      const bool converged = ratio < 1e-10 && step >= 1;
      const bool stalled = !converged && step >= 1 && ratio > 0.999999999 * ratio_old[k];
      if (converged || stalled)
        R[k] = 0;
      else
        ratio_old[k] = ratio;
    }
    if (!any(R)) break;
    solve(R, res.p, corr.p);
    const amigo_ipm_batch BU = batch(R);
    ipm_ok(amigo_ipm_batch_ir_update(&BU, &I, &S, corr.p, ezl.p, ezu.p, px.p, dzl.p, dzu.p, s),
           "refinement update");
    for (int k = 0; k < K; k++)
      if (R[k]) P[k].ir_steps++;
  }
  return final_ratio;
}

// Newton direction of the problems of B at their barrier parameters
// (newton_direction.py:132-195): condensed residual, solve, refinement
// (with the FP64 factor when mixed-precision refinement stalls above
// 1e-10), bound-dual steps.
void amigo_ipm_driver::solve_with_mu(const Set& B) {
  residual(B, vars, grad.p, res.p);
  copy_x(B, rhs.p, res.p);
  solve(B, res.p, sol.p);
  copy_x(B, px.p, sol.p);
  std::vector<double> ratio = refine(B);
  if (o.precision == AMIGO_PREC_MIXED) {
    Set F(K, 0);
    for (int k = 0; k < K; k++) F[k] = B[k] && !(ratio[k] < 1e-10);
    if (any(F)) {
      std::vector<int> pos, neg, nan;
      factor(F, AMIGO_PREC_FP64, pos, neg, nan);
      for (int k = 0; k < K; k++)
        if (F[k]) P[k].fp64_fallbacks++;
      solve(F, rhs.p, sol.p);
      copy_x(F, px.p, sol.p);
      refine(F);
    }
  }
  direction(B, px.p);
}

// Barrier update of the problems of A past their first iteration: the LOQO rule of
// barrier_strategy/heuristic.py:9-50, or the monotone rule with its
// progress test.
void amigo_ipm_driver::barrier_update(const Set& A) {
  const double mu_floor =
      std::min(o.convergence_tolerance, o.compl_inf_tol) / (o.barrier_tol_factor + 1.0);
  Set M(K, 0);
  for (int k = 0; k < K; k++) M[k] = A[k] && P[k].iter > 0 && P[k].mu > mu_floor;
  if (!any(M)) return;
  if (o.barrier_heuristic) {
    Red c = reduce(0, M, vars, zeros(), nullptr);
    fetch();
    for (int k = 0; k < K; k++) {
      if (!M[k]) continue;
      double avg, xi;
      complementarity(at(c, k, 0), at(c, k, 1), at(c, k, 2), avg, xi);
      double factor;
      if (xi > 1e-10) {
        const double term = (1 - o.heuristic_barrier_r) * (1 - xi) / xi;
        factor = std::pow(std::min(term, 2.0), 3.0);
      } else {
        factor = 8.0;
      }
      P[k].mu = std::max(o.heuristic_barrier_gamma * factor * avg, mu_floor);
    }
    return;
  }
  // monotone: reduce mu while the scaled barrier error is below kappa mu
  Set L(K, 0);
  for (int k = 0; k < K; k++) {
    if (!M[k]) continue;
    const double r = P[k].res_norm, mu = P[k].mu;
    L[k] = o.progress_based_barrier ? (r <= o.barrier_tol_factor * mu) : (r <= 0.1 * mu);
  }
  while (any(L)) {
    Red e = reduce(1, L, vars, scalars(per(&Prob::mu)), grad.p);
    Red z = reduce(9, L, vars, zeros(), nullptr);
    fetch();
    for (int k = 0; k < K; k++) {
      if (!L[k]) continue;
      double s_d, s_c;
      optimality_scaling(at(z, k, 0), at(z, k, 1), at(z, k, 2), at(z, k, 3), s_d, s_c);
      const double err = std::max(std::max(at(e, k, 0) / s_d, at(e, k, 1)), at(e, k, 2) / s_c);
      const double mu = P[k].mu;
      if (err > o.barrier_tol_factor * mu) {
        L[k] = 0;
        continue;
      }
      double next = std::min(o.mu_linear_decrease_factor * mu,
                             std::pow(mu, o.mu_superlinear_decrease_power));
      next = std::max(next, mu_floor);
      if (next >= mu) {
        L[k] = 0;
        continue;
      }
      P[k].mu = next;
    }
  }
}

// ---- filter line search ------------------------------------------------------

void amigo_ipm_driver::accept_checks(Prob& p, Search& L, double tb, double tt, double a, bool& acc,
                                     bool& frej) const {
  (void)p;
  acc = false;
  frej = false;
  const double eps10 = 10.0 * kEps;
  auto switching = [&](double al) {  // f-type step (Eq. 19)
    return L.ref_dphi < 0.0 &&
           al * std::pow(-L.ref_dphi, o.filter_s_phi) > o.filter_delta * std::pow(L.ref_theta, o.filter_s_theta);
  };
  auto armijo = [&](double al) {  // Eq. 20
    return (tb - L.ref_barr) - o.filter_eta_phi * al * L.ref_dphi <= eps10 * std::fabs(L.ref_barr);
  };
  if (tt > L.theta_max) return;
  if (a > 0.0 && switching(a) && L.ref_theta <= L.theta_min) {
    if (!armijo(a)) return;
  } else {
    // sufficient progress in theta or phi against the current iterate (Eq. 18)
    if (tb > L.ref_barr) {
      const double basval = std::fabs(L.ref_barr) > 10.0 ? std::log10(std::fabs(L.ref_barr)) : 1.0;
      if (std::log10(std::max(tb - L.ref_barr, 1e-300)) > 5.0 + basval) return;
    }
    const bool theta_ok = tt - (1.0 - o.filter_gamma_theta) * L.ref_theta <= eps10 * std::fabs(L.ref_theta);
    const bool phi_ok = (tb - L.ref_barr) - (-o.filter_gamma_phi * L.ref_theta) <= eps10 * std::fabs(L.ref_barr);
    if (!(theta_ok || phi_ok)) return;
  }
  acc = p.filter.acceptable(tb, tt);
  frej = !acc;
}

void amigo_ipm_driver::update_filter(Prob& p, const Search& L, double tb, double a) const {
  const double eps10 = 10.0 * kEps;
  const bool ftype = L.ref_dphi < 0.0 && a * std::pow(-L.ref_dphi, o.filter_s_phi) >
                                             o.filter_delta * std::pow(L.ref_theta, o.filter_s_theta);
  const bool armijo =
      (tb - L.ref_barr) - o.filter_eta_phi * a * L.ref_dphi <= eps10 * std::fabs(L.ref_barr);
  if (!(ftype && armijo)) p.filter.add(L.ref_barr, L.ref_theta);
}

// Backtracking filter line search with second-order correction
// (filter_line_search.py:54-280) for the problems of S.  ls[k] carries
// alpha_x / alpha_z in; wd_ref[k] selects the watchdog's reference values
// (ls[k].ref_* preset).  Each round evaluates one trial point per problem;
// problems in a second-order correction first solve for their corrected
// step.
void amigo_ipm_driver::line_search(const Set& S, std::vector<Search>& ls,
                                   const std::vector<double>& phi_cur, const Set& wd_ref) {
  const bool use_soc = o.second_order_correction != 0;
  Set need(K, 0);
  for (int k = 0; k < K; k++) need[k] = S[k] && !wd_ref[k];
  if (any(need)) {
    const double* z = zeros();
    Red th = reduce(4, need, vars, z, grad.p);
    Red dp = reduce(3, need, vars, scalars(per(&Prob::mu)), res.p, px.p, diag.p);
    fetch();
    for (int k = 0; k < K; k++)
      if (need[k]) {
        ls[k].ref_theta = at(th, k);
        ls[k].ref_dphi = at(dp, k);
      }
  }
  for (int k = 0; k < K; k++) {
    if (!S[k]) continue;
    Search& L = ls[k];
    L.on = true;
    if (!wd_ref[k]) L.ref_barr = phi_cur[k];
    const double theta_0 = std::isnan(P[k].theta0) ? L.ref_theta : P[k].theta0;
    L.theta_min = 1e-4 * std::max(1.0, theta_0);
    L.theta_max = 1e4 * std::max(1.0, theta_0);
    if (wd_ref[k]) {
      L.alpha_min = L.alpha_x;
    } else {
      double am = o.filter_gamma_theta;
      if (L.ref_dphi < 0.0) {
        am = std::min(o.filter_gamma_theta, o.filter_gamma_phi * L.ref_theta / (-L.ref_dphi));
        if (L.ref_theta <= L.theta_min)
          am = std::min(am, o.filter_delta * std::pow(L.ref_theta, o.filter_s_theta) /
                                std::pow(-L.ref_dphi, o.filter_s_phi));
      }
      L.alpha_min = am * 0.05;
    }
    L.alpha = L.alpha_x;
    L.steps = 0;
    L.last_filter_reject = false;
    L.soc = false;
    L.accepted = false;
  }
  State keep{upd_keep.x, upd_keep.zl, upd_keep.zu, nullptr, nullptr};
  State u3{upd.x, upd.zl, upd.zu, nullptr, nullptr};
  if (use_soc) {
    residual(S, vars, grad.p, res.p);
    copy_x(S, res_orig.p, res.p);
    copy_state(S, keep, u3, 3);
    copy_x(S, px_orig.p, px.p);
  }
  Set searching = S;
  const double red_factor = o.backtracking_factor;
  while (any(searching)) {
    // Problems whose correction loop has ended go back to backtracking.
    Set restore(K, 0);
    for (int k = 0; k < K; k++) {
      Search& L = ls[k];
      if (!searching[k] || !L.soc) continue;
      const bool stop = L.soc_count >= o.filter_max_soc ||
                        (L.soc_count > 0 && L.trial_theta > o.filter_kappa_soc * L.theta_soc_old);
      if (stop) {
        L.soc = false;
        restore[k] = 1;
        L.alpha *= red_factor;
        L.steps++;
      }
    }
    if (any(restore)) {
      copy_state(restore, u3, keep, 3);
      copy_x(restore, px.p, px_orig.p);
    }
    Set trial_set(K, 0), soc_set(K, 0), failed(K, 0);
    for (int k = 0; k < K; k++) {
      if (!searching[k]) continue;
      Search& L = ls[k];
      if (L.soc) {
        soc_set[k] = 1;
        L.theta_soc_old = L.trial_theta;
      } else if ((L.alpha > L.alpha_min || L.steps == 0) && L.steps < o.max_line_search_iterations) {
        trial_set[k] = 1;
      } else {
        failed[k] = 1;
        searching[k] = 0;
      }
    }
    if (any(failed)) {
      gradient(failed, vars.x, grad.p);
      for (int k = 0; k < K; k++)
        if (failed[k]) {
          Search& L = ls[k];
          L.accepted = false;
          L.ratio = L.alpha / L.alpha_x;
          L.iters = L.steps;
          L.filter_reject = L.last_filter_reject;
        }
    }
    if (any(soc_set)) {
      // corrected right-hand side at the trial point, then the step
      residual(soc_set, temp, grad.p, res.p);
      std::vector<double> asoc(K, 0.0);
      for (int k = 0; k < K; k++) asoc[k] = ls[k].alpha_soc;
      const Sel q = select(soc_set);
      hip_ok(amigo_amd::drv::soc_rhs(d.n_constraints, d.cidx_global, d.row_prob, q.mask,
                                     scalars(asoc), res.p, csoc.p, s, rows_c(q)),
             "correction rhs");
      copy_x(soc_set, res.p, csoc.p);
      gradient(soc_set, vars.x, grad.p);
      solve(soc_set, res.p, sol.p);
      copy_x(soc_set, px.p, sol.p);
      direction(soc_set, px.p);
      std::vector<double> tau = per(&Prob::tau);
      Red ms = max_step(soc_set, tau);
      fetch();
      for (int k = 0; k < K; k++)
        if (soc_set[k]) {
          ls[k].soc_ax = at(ms, k, 0);
          ls[k].soc_az = at(ms, k, 2);
        }
    }
    Set eval_set(K, 0);
    std::vector<double> ax(K, 0.0), az(K, 0.0);
    for (int k = 0; k < K; k++) {
      if (trial_set[k]) {
        eval_set[k] = 1;
        ax[k] = ls[k].alpha;
        az[k] = ls[k].alpha_z;
      } else if (soc_set[k]) {
        eval_set[k] = 1;
        ax[k] = ls[k].soc_ax;
        az[k] = ls[k].soc_az;
      }
    }
    if (!any(eval_set)) continue;
    trial(eval_set, ax, az);
    gradient(eval_set, temp.x, grad.p);
    const double* z = zeros();
    Red th = reduce(4, eval_set, temp, z, grad.p);
    Red ob = objective(eval_set, temp.x);
    Red lg = reduce(2, eval_set, temp, scalars(per(&Prob::mu)), nullptr);
    fetch();
    Set take(K, 0), soc_start(K, 0);
    for (int k = 0; k < K; k++) {
      if (!eval_set[k]) continue;
      Search& L = ls[k];
      const double tt = at(th, k), tb = at(ob, k) + at(lg, k);
      bool acc, frej;
      if (trial_set[k]) {
        accept_checks(P[k], L, tb, tt, L.alpha, acc, frej);
        if (acc) {
          update_filter(P[k], L, tb, L.alpha);
          take[k] = 1;
          searching[k] = 0;
          L.accepted = true;
          L.ratio = L.alpha / L.alpha_x;
          L.iters = L.steps + 1;
          L.filter_reject = L.last_filter_reject;
          continue;
        }
        L.last_filter_reject = frej;
        if (use_soc && L.steps == 0 && tt >= L.ref_theta) {
          L.soc = true;
          L.soc_count = 0;
          L.alpha_soc = L.alpha;
          L.theta_soc_old = 0.0;
          L.a_test = L.alpha;
          L.trial_theta = tt;
          soc_start[k] = 1;
        } else {
          L.alpha *= red_factor;
          L.steps++;
        }
      } else {
        L.trial_theta = tt;
        L.trial_barr = tb;
        accept_checks(P[k], L, tb, tt, L.a_test, acc, frej);
        if (acc) {
          update_filter(P[k], L, tb, L.a_test);
          take[k] = 1;
          searching[k] = 0;
          L.soc = false;
          L.accepted = true;
          L.ratio = 1.0;
          L.iters = L.steps + 1;
          L.filter_reject = false;
        } else {
          L.alpha_soc = L.soc_ax;
          L.soc_count++;
        }
      }
    }
    if (any(take)) copy_state(take, vars, temp, 3);
    if (any(soc_start)) copy_x(soc_start, csoc.p, res_orig.p);
  }
}

// The watchdog around the line search (filter_line_search.py:395-438).
void amigo_ipm_driver::line_search_with_watchdog(const Set& B, std::vector<Search>& ls) {
  State wdu{wd_upd.x, wd_upd.zl, wd_upd.zu, nullptr, nullptr};
  State u3{upd.x, upd.zl, upd.zu, nullptr, nullptr};
  auto restore = [&](const Set& R) {
    copy_x(R, vars.x, wd_x.p);
    copy_state(R, u3, wdu, 3);
    copy_x(R, px.p, wd_px.p);
    gradient(R, vars.x, grad.p);
  };
  Set R1(K, 0), S1(K, 0);
  for (int k = 0; k < K; k++)
    if (B[k] && P[k].wd_on && ls[k].alpha_x < 1e-16) R1[k] = 1;
  if (any(R1)) {
    restore(R1);
    for (int k = 0; k < K; k++)
      if (R1[k]) {
        P[k].wd_on = false;
        P[k].wd_shortened = 0;
      }
  }
  for (int k = 0; k < K; k++)
    if (B[k] && !P[k].wd_on && o.watchdog_shortened_iter_trigger > 0 &&
        P[k].wd_shortened >= o.watchdog_shortened_iter_trigger)
      S1[k] = 1;
  if (any(S1)) {
    copy_x(S1, wd_x.p, vars.x);
    copy_state(S1, wdu, u3, 3);
    copy_x(S1, wd_px.p, px.p);
    Red th = reduce(4, S1, vars, zeros(), grad.p);
    Red ob = objective(S1, vars.x);
    Red lg = reduce(2, S1, vars, scalars(per(&Prob::mu)), nullptr);
    Red dp = reduce(3, S1, vars, scalars(per(&Prob::mu)), res.p, px.p, diag.p);
    fetch();
    for (int k = 0; k < K; k++)
      if (S1[k]) {
        P[k].wd_theta = at(th, k);
        P[k].wd_barr = at(ob, k) + at(lg, k);
        P[k].wd_dphi = at(dp, k);
        P[k].wd_on = true;
        P[k].wd_trials = 0;
      }
  }
  std::vector<double> phi(K, 0.0);
  {
    Red ob = objective(B, vars.x);
    Red lg = reduce(2, B, vars, scalars(per(&Prob::mu)), nullptr);
    fetch();
    for (int k = 0; k < K; k++)
      if (B[k]) phi[k] = at(ob, k) + at(lg, k);
  }
  Set wd_ref(K, 0);
  for (int k = 0; k < K; k++)
    if (B[k] && P[k].wd_on) {
      wd_ref[k] = 1;
      ls[k].ref_theta = P[k].wd_theta;
      ls[k].ref_barr = P[k].wd_barr;
      ls[k].ref_dphi = P[k].wd_dphi;
    }
  line_search(B, ls, phi, wd_ref);
  Set again(K, 0);
  for (int k = 0; k < K; k++) {
    if (!wd_ref[k]) continue;
    Prob& p = P[k];
    if (ls[k].accepted) {
      p.wd_on = false;
      p.wd_shortened = 0;
      continue;
    }
    p.wd_trials++;
    if (p.wd_trials > o.watchdog_trial_iter_max) {
      again[k] = 1;
    } else {
      ls[k].accepted = true;  // the watchdog keeps the iterate for now
    }
  }
  if (!any(again)) return;
  restore(again);
  for (int k = 0; k < K; k++)
    if (again[k]) {
      P[k].wd_on = false;
      P[k].wd_shortened = 0;
    }
  Red ms = max_step(again, per(&Prob::tau));
  Red ob = objective(again, vars.x);
  Red lg = reduce(2, again, vars, scalars(per(&Prob::mu)), nullptr);
  fetch();
  for (int k = 0; k < K; k++)
    if (again[k]) {
      ls[k].alpha_x = at(ms, k, 0);
      ls[k].alpha_z = at(ms, k, 2);
      phi[k] = at(ob, k) + at(lg, k);
    }
  line_search(again, ls, phi, Set(K, 0));
}

// ---- feasibility restoration (feasibility_restoration.py:16-172) -------------
// Near-feasible problems (theta below max(tol, 1e-4 max(1, theta_0)))
// backtrack on the residual norm against the filter at once.  The others
// enter the restoration phase: Newton steps at raised barrier parameters
// until theta falls by 10 %, one step per round, inside the same launches as
// the problems in their main iterations.  Returns the problems whose
// restoration is already complete (all of them succeed).
Set amigo_ipm_driver::restoration_begin(const Set& R) {
  Set success(K, 0);
  Red th = reduce(4, R, vars, zeros(), grad.p);
  fetch();
  std::vector<double> theta_k(K, 0.0);
  Set RA(K, 0), RB(K, 0);
  for (int k = 0; k < K; k++) {
    if (!R[k]) continue;
    theta_k[k] = at(th, k);
    const double t0 = std::isnan(P[k].theta0) ? theta_k[k] : P[k].theta0;
    const double theta_min_restore = 1e-4 * std::max(1.0, t0);
    if (theta_k[k] < std::max(o.convergence_tolerance, theta_min_restore))
      RA[k] = 1;
    else
      RB[k] = 1;
  }
  const int max_ls = o.max_line_search_iterations;
  const double backtrack = o.backtracking_factor;
  if (any(RA)) {
    std::vector<double> tau(K);
    for (int k = 0; k < K; k++) tau[k] = tau_of(P[k].mu);
    Red r0 = residual(RA, vars, grad.p, res.p);
    Red ms = max_step(RA, tau);
    fetch();
    std::vector<double> res_cur(K), ax(K), az(K), alpha(K, 1.0);
    for (int k = 0; k < K; k++)
      if (RA[k]) {
        res_cur[k] = std::sqrt(at(r0, k, 0));
        ax[k] = at(ms, k, 0);
        az[k] = at(ms, k, 2);
      }
    Set act = RA;
    for (int t = 0; t < max_ls && any(act); t++) {
      std::vector<double> a(K, 0.0);
      for (int k = 0; k < K; k++) a[k] = alpha[k] * ax[k];
      trial(act, a, az);
      gradient(act, temp.x, grad.p);
      Red rt = residual(act, temp, grad.p, res.p);
      fetch();
      Set cand(K, 0);
      std::vector<double> rtn(K, 0.0);
      for (int k = 0; k < K; k++)
        if (act[k]) {
          rtn[k] = std::sqrt(at(rt, k, 0));
          cand[k] = rtn[k] < res_cur[k];
        }
      Set take(K, 0);
      if (any(cand)) {
        Red ob = objective(cand, temp.x);
        Red lg = reduce(2, cand, temp, scalars(per(&Prob::mu)), nullptr);
        Red tt = reduce(4, cand, temp, zeros(), grad.p);
        fetch();
        for (int k = 0; k < K; k++) {
          if (!cand[k]) continue;
          const double phi_t = at(ob, k) + at(lg, k), theta_t = at(tt, k);
          if (P[k].filter.acceptable(phi_t, theta_t)) {
            take[k] = 1;
          } else if (rtn[k] < 0.99 * res_cur[k]) {
            take[k] = 1;
            P[k].filter.clear();
            P[k].theta0 = theta_t;
          }
        }
      }
      if (any(take)) copy_state(take, vars, temp, 3);
      for (int k = 0; k < K; k++) {
        if (!act[k]) continue;
        if (take[k]) {
          act[k] = 0;
          success[k] = 1;
        } else {
          alpha[k] *= backtrack;
        }
      }
    }
    if (any(act)) {
      gradient(act, vars.x, grad.p);
      for (int k = 0; k < K; k++)
        if (act[k]) {
          P[k].filter.clear();
          P[k].theta0 = theta_k[k];
          success[k] = 1;
        }
    }
  }
  if (any(RB)) {
    Red ob = objective(RB, vars.x);
    Red lg = reduce(2, RB, vars, scalars(per(&Prob::mu)), nullptr);
    fetch();
    for (int k = 0; k < K; k++) {
      if (!RB[k]) continue;
      Prob& p = P[k];
      p.filter.add(at(ob, k) + at(lg, k), theta_k[k]);
      p.rtheta = p.rtheta_start = theta_k[k];
      p.rtarget = 0.9 * theta_k[k];
      p.rsaved_mu = p.mu;
      std::vector<double> lv;
      for (double m : {std::max(p.mu, 1e-2), 0.1, 1.0})
        if (m >= p.mu) lv.push_back(m);
      std::sort(lv.begin(), lv.end());
      lv.erase(std::unique(lv.begin(), lv.end()), lv.end());
      p.rlevels = lv;
      p.rlevel = 0;
      p.riter = 0;
      p.rper_level = std::max(3, o.filter_restoration_max_iter / int(lv.size()));
      p.restoring = true;
    }
  }
  return success;
}

// The backtracking of one restoration iteration (its Newton step and step
// bounds are computed with the round's other problems): the first trial that
// lowers theta is taken.  Returns the problems whose restoration ended
// (restore_ok tells whether theta fell by 1 %).
Set amigo_ipm_driver::restoration_search(const Set& RB) {
  const int max_ls = o.max_line_search_iterations;
  const double backtrack = o.backtracking_factor;
  std::vector<double> alpha(K, 1.0);
  Set srch = RB, taken(K, 0);
  for (int t = 0; t < max_ls && any(srch); t++) {
    std::vector<double> a(K, 0.0), b(K, 0.0);
    for (int k = 0; k < K; k++) {
      a[k] = alpha[k] * P[k].rax;
      b[k] = alpha[k] * P[k].raz;
    }
    trial(srch, a, b);
    gradient(srch, temp.x, grad.p);
    Red tt = reduce(4, srch, temp, zeros(), grad.p);
    fetch();
    Set take(K, 0);
    for (int k = 0; k < K; k++) {
      if (!srch[k]) continue;
      const double v = at(tt, k);
      if (v < P[k].rtheta) {
        take[k] = 1;
        taken[k] = 1;
        P[k].rtheta = v;
        srch[k] = 0;
      } else {
        alpha[k] *= backtrack;
      }
    }
    if (any(take)) copy_state(take, vars, temp, 3);
  }
  Set ended(K, 0);
  for (int k = 0; k < K; k++) {
    if (!RB[k]) continue;
    Prob& p = P[k];
    bool next_level = false, stop = false;
    if (!taken[k] || p.rtheta < p.rtarget) {
      if (p.rtheta < p.rtarget)
        stop = true;
      else
        next_level = true;
    } else if (++p.riter >= p.rper_level) {
      next_level = true;
    }
    if (next_level) {
      p.riter = 0;
      if (++p.rlevel >= int(p.rlevels.size())) stop = true;
    }
    if (stop) ended[k] = 1;
  }
  if (any(ended)) {
    for (int k = 0; k < K; k++)
      if (ended[k]) {
        Prob& p = P[k];
        p.mu = p.rsaved_mu;
        p.reg.set_mu(p.mu);
        p.restoring = false;
        p.restore_ok = p.rtheta < p.rtheta_start * 0.99;
      }
    gradient(ended, vars.x, grad.p);
    sigma(ended, vars, diag.p);
  }
  return ended;
}

// The end of a main iteration (ipm_driver.py:430-473) for the problems of
// done whose Newton step was found: an accepted step floors the slacks,
// resets the bound multipliers and refreshes the gradient; a rejected one
// restores mu.  The records of the next iteration read the step.
void amigo_ipm_driver::finish_step(const Set& done, const Set& accepted) {
  Set G = done & accepted;
  if (any(G)) {
    std::vector<double> smin(K, 0.0);
    for (int k = 0; k < K; k++) {
      double sm = kEps * std::min(1.0, P[k].mu);
      if (sm == 0.0) sm = std::numeric_limits<double>::min();
      smin[k] = sm;
    }
    const amigo_ipm_batch BG = batch(G);
    const amigo_ipm_problem I = info();
    const amigo_ipm_state S = vars.c();
    Red fl = red_slot(1);
    ipm_ok(amigo_ipm_batch_floor_slacks(&BG, &I, &S, d.lbx0, d.ubx0, scalars(per(&Prob::mu)),
                                        scalars(smin), 1.81898940354586e-12, ws.p, ex.dev(fl.off),
                                        nullptr, s),
           "slack floor");
    ipm_ok(amigo_ipm_batch_reset_bound_multipliers(&BG, &I, &S, scalars(per(&Prob::mu)), 1e10, s),
           "multiplier reset");
    gradient(G, vars.x, grad.p);
  }
  for (int k = 0; k < K; k++) {
    if (!done[k]) continue;
    Prob& p = P[k];
    if (!accepted[k]) {
      p.rec_ax = p.rec_az = 0.0;
      p.rec_xi = p.rec_zi = -1;
      p.mu = p.mu_before;
      increase_barrier(p);
    } else {
      p.rec_ax = p.ls_ratio * p.ax;
      p.rec_az = p.az;
      p.rec_xi = p.xi;
      p.rec_zi = p.zi;
      p.rejections = 0;
      if (!p.wd_on) {
        if (p.ls_ratio == 1.0 || p.line_iters == 1)
          p.wd_shortened = 0;
        else
          p.wd_shortened++;
      }
    }
    p.iter++;
  }
}

// ---- the main loop (ipm_driver.py:106-483) --------------------------------------
// Each round, every running problem takes one step of its own: a main
// iteration (records and convergence test, barrier update, Newton
// direction, line search), or one restoration iteration.  The Newton
// systems of both kinds are factored and solved together.
void amigo_ipm_driver::run() {
  t0 = std::chrono::steady_clock::now();
  auto tp = t0;
  ldl_set.clear();
  reset();
  initialize();
  toc(PH_INIT, tp);
  auto stop = [&](int k, double now) {
    P[k].running = false;
    P[k].t_end = now;
    if (on_done) on_done(ctx, k);
  };
  while (true) {
    Set M(K, 0), RB(K, 0), limit(K, 0);
    for (int k = 0; k < K; k++) {
      if (!P[k].running) continue;
      if (P[k].restoring)
        RB[k] = 1;
      else if (P[k].iter >= o.max_iterations)
        limit[k] = 1;
      else
        M[k] = 1;
    }
    if (any(limit)) {
      // the last step's launches finish before the iterate is handed over
      hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
      const double now = elapsed();
      for (int k = 0; k < K; k++)
        if (limit[k]) stop(k, now);
    }
    if (!any(M) && !any(RB)) break;
    // restoration iterations: the barrier parameter of their level
    for (int k = 0; k < K; k++)
      if (RB[k]) {
        P[k].mu = P[k].rlevels[P[k].rlevel];
        P[k].reg.set_mu(P[k].mu);
      }
    if (any(RB)) gradient(RB, vars.x, grad.p);
    // main iterations: records and the convergence test
    if (any(M)) {
      for (int k = 0; k < K; k++)
        if (M[k]) P[k].reg.set_mu(P[k].mu);
      const double* z0 = zeros();
      Red rr = residual(M, vars, grad.p, res.p);
      Set need(K, 0);
      for (int k = 0; k < K; k++) need[k] = M[k] && std::isnan(P[k].theta0);
      Red th0{};
      if (any(need)) th0 = reduce(4, need, vars, z0, grad.p);
      Red kkt = reduce(1, M, vars, z0, grad.p);
      Red sums = reduce(9, M, vars, z0, nullptr);
      Red xm = red_slot(1);
      ipm_ok(amigo_segment_absmax(vars.x, d.xseg, K, ex.dev(xm.off), nullptr, s), "absmax");
      Red ob{}, lg{};
      if (o.record_objective) {
        ob = objective(M, vars.x);
        lg = reduce(2, M, vars, scalars(per(&Prob::mu)), nullptr);
      }
      fetch();
      const double now = elapsed();
      for (int k = 0; k < K; k++) {
        if (!M[k]) continue;
        Prob& p = P[k];
        p.res_norm = std::sqrt(at(rr, k, 0));
        if (need[k]) p.theta0 = at(th0, k);
        const double du = at(kkt, k, 0), pr = at(kkt, k, 1), co = at(kkt, k, 2);
        double s_d, s_c;
        optimality_scaling(at(sums, k, 0), at(sums, k, 1), at(sums, k, 2), at(sums, k, 3), s_d, s_c);
        const double err = std::max(std::max(du / s_d, pr), co / s_c);
        int status = AMIGO_IPM_RUNNING;
        if (err <= o.convergence_tolerance && du <= o.dual_inf_tol && pr <= o.constr_viol_tol &&
            co <= o.compl_inf_tol) {
          status = AMIGO_IPM_CONVERGED;
        } else if (at(xm, k) > o.diverging_iterates_tol) {
          status = AMIGO_IPM_DIVERGED;
        } else {
          const bool acceptable = err <= o.acceptable_tol && du <= o.acceptable_dual_inf_tol &&
                                  pr <= o.acceptable_constr_viol_tol &&
                                  co <= o.acceptable_compl_inf_tol;
          if (o.acceptable_iter > 0 && acceptable) {
            if (++p.acc_count >= o.acceptable_iter) status = AMIGO_IPM_ACCEPTABLE;
          } else {
            p.acc_count = 0;
          }
          if (status == AMIGO_IPM_RUNNING) {
            const double rel = std::fabs(p.res_norm - p.prev_res) / std::max(p.res_norm, 1e-30);
            p.floor_count = (rel < 1e-14 && p.iter > 0) ? p.floor_count + 1 : 0;
            if (p.floor_count >= 3 && acceptable) status = AMIGO_IPM_PRECISION_FLOOR;
          }
        }
        amigo_ipm_iteration rec{};
        rec.iteration = p.iter;
        rec.time = now;
        rec.residual = p.res_norm;
        rec.barrier_param = p.mu;
        rec.line_iters = p.line_iters;
        rec.alpha_x = p.rec_ax;
        rec.x_index = int(p.rec_xi);
        rec.alpha_z = p.rec_az;
        rec.z_index = int(p.rec_zi);
        rec.inertia_delta = p.last_dw;
        rec.inf_du = du;
        rec.inf_pr = pr;
        rec.compl_inf = co;
        rec.nlp_error = err;
        rec.filter_size = p.filter.size();
        rec.objective = o.record_objective ? at(ob, k) + at(lg, k) : kNaN;
        p.records.push_back(rec);
        if (status != AMIGO_IPM_RUNNING) {
          p.status = status;
          M[k] = 0;
          stop(k, now);
        }
      }
      for (int k = 0; k < K; k++)
        if (M[k]) P[k].prev_res = P[k].res_norm;
    }
    Set N(K, 0);
    for (int k = 0; k < K; k++) N[k] = M[k] || RB[k];
    if (!any(N)) continue;
    sigma(N, vars, diag.p);
    copy_x(N, dbase.p, diag.p);
    for (int k = 0; k < K; k++)
      if (M[k]) P[k].mu_before = P[k].mu;
    if (any(M)) barrier_update(M);
    for (int k = 0; k < K; k++)
      if (M[k]) P[k].reg.set_mu(P[k].mu);
    toc(PH_CONV, tp);
    std::vector<char> ok = factorize(N);
    toc(PH_FACTOR, tp);
    Set B(K, 0);
    for (int k = 0; k < K; k++) B[k] = N[k] && ok[k];
    if (any(B)) solve_with_mu(B);
    toc(PH_SOLVE, tp);
    Set changed(K, 0);
    for (int k = 0; k < K; k++) changed[k] = M[k] && P[k].mu != P[k].mu_before;
    if (any(changed)) {
      Red th = reduce(4, changed, vars, zeros(), grad.p);
      fetch();
      for (int k = 0; k < K; k++) {
        if (!changed[k]) continue;
        Prob& p = P[k];
        p.reg.set_mu(p.mu);
        p.filter.clear();
        p.theta0 = at(th, k);
        p.wd_on = false;
        p.wd_shortened = 0;
        p.wd_trials = 0;
        p.filter_rejects = 0;
        p.filter_resets = 0;
      }
    }
    for (int k = 0; k < K; k++) {
      if (!M[k] || ok[k]) continue;
      Prob& p = P[k];
      p.rejections++;
      p.mu = p.mu_before;
      p.line_iters = 0;
      p.rec_ax = p.rec_az = 0.0;
      p.rec_xi = p.rec_zi = -1;
      increase_barrier(p);
      p.iter++;
    }
    Set BM(K, 0);
    for (int k = 0; k < K; k++) BM[k] = M[k] && ok[k];
    if (any(BM)) residual(BM, vars, grad.p, res.p);
    Set MS(K, 0);
    std::vector<double> tau(K, 1.0);
    for (int k = 0; k < K; k++) {
      MS[k] = BM[k] || RB[k];
      P[k].tau = tau_of(P[k].mu);
      tau[k] = P[k].tau;
    }
    if (any(MS)) {
      Red ms = max_step(MS, tau);
      fetch();
      for (int k = 0; k < K; k++) {
        if (BM[k]) {
          Prob& p = P[k];
          p.ax = at(ms, k, 0);
          p.xi = (long long)at(ms, k, 1);
          p.az = at(ms, k, 2);
          p.zi = (long long)at(ms, k, 3);
          if (o.equal_primal_dual_step) p.ax = p.az = std::min(p.ax, p.az);
        } else if (RB[k]) {
          P[k].rax = at(ms, k, 0);
          P[k].raz = at(ms, k, 2);
        }
      }
    }
    Set done(K, 0), accepted(K, 0);
    if (any(BM)) {
      std::vector<Search> ls(K);
      for (int k = 0; k < K; k++)
        if (BM[k]) {
          ls[k].alpha_x = P[k].ax;
          ls[k].alpha_z = P[k].az;
        }
      line_search_with_watchdog(BM, ls);
      toc(PH_SEARCH, tp);
      Set rest(K, 0);
      for (int k = 0; k < K; k++) {
        if (!BM[k]) continue;
        Prob& p = P[k];
        p.line_iters = ls[k].iters;
        p.ls_ratio = ls[k].ratio;
        if (ls[k].accepted) {
          p.filter_rejects = ls[k].filter_reject ? p.filter_rejects + 1 : 0;
          if (p.filter_rejects >= o.filter_reset_trigger && p.filter_resets < o.max_filter_resets) {
            p.filter.clear();
            p.filter_resets++;
            p.filter_rejects = 0;
          }
          done[k] = accepted[k] = 1;
        } else {
          rest[k] = 1;
        }
      }
      if (any(rest)) {
        Set quick = restoration_begin(rest);
        for (int k = 0; k < K; k++) {
          if (!rest[k]) continue;
          P[k].restorations++;
          if (quick[k]) {
            done[k] = accepted[k] = 1;
            P[k].line_iters = 0;
            P[k].wd_shortened = 0;
          }
        }
      }
    }
    if (any(RB)) {
      Set ended = restoration_search(RB);
      for (int k = 0; k < K; k++) {
        if (!ended[k]) continue;
        done[k] = 1;
        if (P[k].restore_ok) {
          accepted[k] = 1;
          P[k].line_iters = 0;
          P[k].wd_shortened = 0;
        } else {
          P[k].rejections++;
        }
      }
    }
    toc(PH_RESTORE, tp);
    finish_step(done, accepted);
    toc(PH_FINISH, tp);
  }
  hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
}

// ---- C ABI ----------------------------------------------------------------------

extern "C" {

const char* amigo_ipm_driver_last_error(void) { return g_err.c_str(); }

int amigo_ipm_driver_create(const amigo_nlp_desc* desc, amigo_ipm_driver** out) {
  if (!desc || !out) return fail(-1, "amigo_ipm_driver_create: null argument");
  *out = nullptr;
  const amigo_nlp_desc& d = *desc;
  if (d.K < 1 || d.n < 1 || !d.xseg_host || !d.pseg_host || !d.cseg_host || !d.groups || !d.ldl ||
      !d.x || !d.x0 || !d.hess || d.ngroups < 0 || !d.zl || !d.zu || !d.sl || !d.su ||
      (d.n_constraints > 0 && !d.dc))
    return fail(-1, "amigo_ipm_driver_create: incomplete problem description");
  if (d.xseg_host[d.K] != d.n || d.pseg_host[d.K] != d.n_primal ||
      d.cseg_host[d.K] != d.n_constraints || d.n_primal + d.n_constraints != d.n)
    return fail(-1, "amigo_ipm_driver_create: segment sizes do not add up");
  for (int g = 0; g < d.ngroups; g++)
    if (d.groups[g].E % 256 != 0 || !d.groups[g].alpha_blk || !d.groups[g].blk_prob_host)
      return fail(-1, "amigo_ipm_driver_create: element groups must be padded to blocks of 256");
  std::unique_ptr<amigo_ipm_driver> h(new amigo_ipm_driver);
  h->d = d;
  try {
    h->setup();
  } catch (const DriverError& e) {
    return fail(e.code, e.msg);
  } catch (const std::exception& e) {
    return fail(-5, std::string("amigo_ipm_driver_create: ") + e.what());
  }
  *out = h.release();
  return 0;
}

int amigo_ipm_driver_destroy(amigo_ipm_driver* d) {
  delete d;
  return 0;
}

int amigo_ipm_driver_solve(amigo_ipm_driver* d, const amigo_ipm_options* opt,
                           amigo_ipm_done_fn on_done, void* ctx, void* stream) {
  if (!d || !opt) return fail(-1, "amigo_ipm_driver_solve: null argument");
  d->o = *opt;
  d->s = static_cast<hipStream_t>(stream);
  d->on_done = on_done;
  d->ctx = ctx;
  if (d->o.precision != AMIGO_PREC_FP64 && d->o.precision != AMIGO_PREC_MIXED)
    return fail(-1, "amigo_ipm_driver_solve: unknown precision");
  try {
    d->run();
  } catch (const DriverError& e) {
    return fail(e.code, e.msg);
  } catch (const std::exception& e) {
    return fail(-5, std::string("amigo_ipm_driver_solve: ") + e.what());
  }
  return 0;
}

int amigo_ipm_driver_result(const amigo_ipm_driver* d, int k, amigo_ipm_result* out) {
  if (!d || !out || k < 0 || k >= d->K) return fail(-1, "amigo_ipm_driver_result: bad argument");
  const Prob& p = d->P[k];
  out->status = p.status;
  out->num_records = int(p.records.size());
  out->num_inertia = int(p.inertia.size() / 2);
  out->num_solver_inertia = int(p.solver_inertia.size() / 2);
  out->fp64_fallbacks = p.fp64_fallbacks;
  out->ir_steps = p.ir_steps;
  out->restorations = p.restorations;
  out->obj_scale = p.obj_scale;
  out->time = p.t_end;
  return 0;
}

int amigo_ipm_driver_iterations(const amigo_ipm_driver* d, int k, amigo_ipm_iteration* out,
                                int cap) {
  if (!d || k < 0 || k >= d->K) return fail(-1, "amigo_ipm_driver_iterations: bad argument");
  const auto& r = d->P[k].records;
  const int n = std::min<int>(cap, int(r.size()));
  if (out && n > 0) std::memcpy(out, r.data(), sizeof(amigo_ipm_iteration) * n);
  return n;
}

int amigo_ipm_driver_inertia(const amigo_ipm_driver* d, int k, int which, int* out, int cap) {
  if (!d || k < 0 || k >= d->K) return fail(-1, "amigo_ipm_driver_inertia: bad argument");
  const auto& v = which ? d->P[k].solver_inertia : d->P[k].inertia;
  const int n = std::min<int>(cap, int(v.size() / 2));
  if (out && n > 0) std::memcpy(out, v.data(), sizeof(int) * 2 * n);
  return n;
}

int amigo_ipm_driver_counts(const amigo_ipm_driver* d, long long* num_factor,
                            long long* num_solve) {
  if (!d) return fail(-1, "amigo_ipm_driver_counts: null driver");
  if (num_factor) *num_factor = d->n_factor;
  if (num_solve) *num_solve = d->n_solve;
  return 0;
}

int amigo_ipm_driver_copy_vector(const amigo_ipm_driver* d, int which, double* dst, void* stream) {
  if (!d || !dst || which < 0 || which > 3)
    return fail(-1, "amigo_ipm_driver_copy_vector: bad argument");
  const double* src = which == 0 ? d->grad.p : which == 1 ? d->res.p : which == 2 ? d->diag.p : d->px.p;
  const hipError_t e = hipMemcpyAsync(dst, src, sizeof(double) * d->d.n, hipMemcpyDeviceToDevice,
                                      static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(-2, std::string("amigo_ipm_driver_copy_vector: ") + hipGetErrorString(e));
  return 0;
}

int amigo_ipm_driver_phase_times(const amigo_ipm_driver* d, double* out8) {
  if (!d || !out8) return fail(-1, "amigo_ipm_driver_phase_times: null argument");
  for (int i = 0; i < PH_COUNT; i++) out8[i] = d->phase_s[i];
  out8[7] = d->host_s;
  return 0;
}

int amigo_ipm_regularization_trace(const int* outcomes, int n, const double* mu,
                                   int has_constraints, double* trace, int cap) {
  if (!outcomes || n < 0 || !mu || (cap > 0 && !trace))
    return fail(-1, "amigo_ipm_regularization_trace: bad argument");
  Regularization reg;
  int written = 0, system = -1;
  bool live = false;
  for (int i = 0; i < n; i++) {
    const int o = outcomes[i];
    if (o == -1) {
      system++;
      reg.set_mu(mu[system]);
      reg.begin_system();
      live = true;
      continue;
    }
    if (!live) continue;
    if (written < cap) {
      trace[2 * written] = reg.dw();
      trace[2 * written + 1] = reg.dc();
    }
    written++;
    if (o == 0) {
      live = false;
      continue;
    }
    const Regularization::Outcome why = o == 1   ? Regularization::Outcome::Singular
                                        : o == 2 ? Regularization::Outcome::TooFewNegative
                                                 : Regularization::Outcome::WrongInertia;
    if (!reg.next(why, has_constraints != 0)) live = false;
  }
  return std::min(written, cap);
}

}  // extern "C"
