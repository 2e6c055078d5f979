// Native interior-point driver: internal declarations.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

namespace amigo_amd {
namespace drv {

// ---- element-wise helper kernels (driver_kernels.hip) ----------------------
// The rows a launch covers: all n rows (list == nullptr), or only the
// segments seg[list[e]] .. seg[list[e] + 1] of the nlist listed problems
// (device arrays; max_len bounds the listed segments' lengths).  Either way
// a kernel also skips rows of problems whose active entry is zero, and
// computes each row exactly as the full launch does.
struct Rows {
  const int* list = nullptr;
  int nlist = 0;
  const int* seg = nullptr;
  int max_len = 0;
};

hipError_t fill(long long n, double v, double* x, hipStream_t s);
hipError_t regularized_base(int n, const int* is_mult, const int* row_prob, const int* active,
                            const double* base, double eps, double* reg, hipStream_t s,
                            const Rows& rows = Rows{});
hipError_t diagonal_shift(int n, const int* is_mult, const int* row_prob, const int* active,
                          const double* reg, const double* dw, const double* dc, double* shift,
                          hipStream_t s, const Rows& rows = Rows{});
hipError_t unit_primal(int n, const int* is_mult, double* d, hipStream_t s);
hipError_t negate(int n, double* x, hipStream_t s);
hipError_t kkt_matvec(int n, const int* rowp, const int* cols, const double* vals,
                      const double* shift, const double* x, double* y, const int* row_prob,
                      const int* active, hipStream_t s, const Rows& rows = Rows{});
hipError_t gather(int n, const long long* idx, const double* src, double* dst, hipStream_t s);
// The constraint-list kernels take rows over constraint numbers j (cidx[j]).
hipError_t set_multipliers(int nc, const long long* cidx, const int* row_prob, const int* keep,
                           const int* active, const double* src, double* x, hipStream_t s,
                           const Rows& rows = Rows{});
hipError_t zero_rows(int nc, const long long* cidx, double* x, hipStream_t s,
                     const Rows& rows = Rows{});
hipError_t soc_rhs(int nc, const long long* cidx, const int* row_prob, const int* active,
                   const double* alpha, const double* res, double* c, hipStream_t s,
                   const Rows& rows = Rows{});
hipError_t copy_rows(int n, const int* row_prob, const int* active, const double* src, double* dst,
                     hipStream_t s, const Rows& rows = Rows{});
hipError_t constraint_scale_vectors(int n, int nc, const long long* cidx, const double* dc,
                                    const double* lbh0, double* sv, double* lbh, hipStream_t s);

// ---- primal-dual regularization (inertia correction) ------------------------
// Algorithm IC of Waechter & Biegler, Math. Program. 106 (2006), Sec. 3.1,
// with the degeneracy detection of its implementation: the KKT matrix gets
// delta_w on the primal rows and -delta_c on the constraint rows until its
// inertia is (n_primal, n_dual, 0).  delta_w grows from 1e-4 (or a third of
// the last successful value) by 100 on its first increase and by 8 after;
// delta_c = 1e-8 mu^(1/4).  A system whose unperturbed factor is singular is
// probed once with delta_c alone, then delta_w alone, then both; a
// perturbation that is needed in three consecutive systems is switched on
// from the start afterwards (the Jacobian or the Hessian is then taken as
// structurally degenerate).  This is the decision sequence of the
// reference's inertia_correction.py:57-393, written as a state machine over
// those rules.
class Regularization {
 public:
  enum class Degeneracy : unsigned char { Unknown, No, Yes };
  // The trial that the current system is on, by the perturbations it has.
  enum class Probe : unsigned char { None, Plain, ConstraintsOnly, HessianOnly, Both };
  enum class Outcome : unsigned char { Singular, TooFewNegative, WrongInertia };

  void set_mu(double mu) { mu_ = mu; }
  // A new KKT matrix: settle the previous probe, remember the last
  // perturbations that were used, start from the structural ones.
  void begin_system();
  // The next perturbation after a failed factorization; false when every
  // option is used up.  has_constraints: the system has dual rows.
  bool next(Outcome why, bool has_constraints);
  double dw() const { return dw_; }
  double dc() const { return dc_; }

 private:
  double dc_value() const { return 1e-8 * std::pow(mu_, 0.25); }
  bool raise_dw();
  void settle_probe();
  bool after_singular();
  bool after_wrong_inertia();
  void strike(Degeneracy& d);

  double mu_ = 1.0;
  double dw_ = 0.0, dc_ = 0.0;
  double dw_last_ = 0.0, dc_last_ = 0.0;
  Degeneracy hess_ = Degeneracy::Unknown, jac_ = Degeneracy::Unknown;
  int strikes_ = 0;
  Probe probe_ = Probe::None;
  bool quality_raised_ = false;
  double pivot_quality_ = 1e-6;
};

// ---- (phi, theta) filter ----------------------------------------------------
// The filter of Waechter & Biegler (2006), Sec. 2.3: a trial point is
// acceptable when no entry dominates it; an entry is stored with the
// margins gamma_phi theta and gamma_theta theta, and entries it dominates
// are dropped (filter_acceptance.py:13-51).
class Filter {
 public:
  void configure(double gamma_phi, double gamma_theta) {
    gphi_ = gamma_phi;
    gtheta_ = gamma_theta;
  }
  bool acceptable(double phi, double theta) const {
    for (const Entry& e : entries_)
      if (phi > e.phi && theta > e.theta) return false;
    return true;
  }
  void add(double phi, double theta) {
    const Entry n{phi - gphi_ * theta, (1.0 - gtheta_) * theta};
    size_t w = 0;
    for (size_t r = 0; r < entries_.size(); r++)
      if (!(entries_[r].phi >= n.phi && entries_[r].theta >= n.theta)) entries_[w++] = entries_[r];
    entries_.resize(w);
    entries_.push_back(n);
  }
  void clear() { entries_.clear(); }
  int size() const { return int(entries_.size()); }

 private:
  struct Entry {
    double phi, theta;
  };
  std::vector<Entry> entries_;
  double gphi_ = 1e-8, gtheta_ = 1e-5;
};

}  // namespace drv
}  // namespace amigo_amd
