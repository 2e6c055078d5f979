// Numeric multifrontal LDL^T on the GPU (host orchestration).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "symbolic.h"

namespace amigo_amd {

// Device-side copy of the symbolic plan plus derived maps.
struct DevPlan {
  int n = 0, K = 0;
  const int* ns = nullptr;      // [K] pivots of the supernode
  const int* cs = nullptr;      // [K] static contribution rows
  const int* cptr = nullptr;    // [K+1]
  const int* crows = nullptr;   // contribution rows (original numbering)
  const int* cmap = nullptr;    // static position of each row in the parent
  const int* kptr = nullptr;    // [K+2] children pointer
  const int* kids = nullptr;    // [K]
  const int* nptr = nullptr;    // [K+1]
  const int* nvars = nullptr;   // [n] node -> variables (original)
  const int* parent = nullptr;  // [K]
  const int* aptr = nullptr;    // [K+1] original-entry assembly map
  const int* aidx = nullptr;    // CSR value index
  const int* arow = nullptr;    // static row position in the front
  const int* acol = nullptr;    // column (pivot index < ns)
  const int* acstart = nullptr; // [n+1] start of supernode column j's entries at nptr[k] + j
};

// Dynamic factor state.
struct DevState {
  int* D = nullptr;        // [K] delayed pivots received from the children
  int* npiv = nullptr;     // [K]
  int* ndel = nullptr;     // [K] delayed pivots passed to the parent
  int* m = nullptr;        // [K] front size
  int64_t* foff = nullptr; // [K] offset of the front in rpool (doubles)
  int64_t* voff = nullptr; // [K] offset of the front variables in ipool
  int* pstate = nullptr;   // [K*8] blocked path: k, nc, kstart, kw, blocked
  double* rpool = nullptr;
  int* ipool = nullptr;
  int* fpos = nullptr;     // [icap] final front position of each assembly position
  unsigned long long* tops = nullptr;  // [2] bump pointers (rpool, ipool)
  unsigned long long rcap = 0, icap = 0;
  int* flags = nullptr;    // [8]: 0 error, 1..3 inertia pos/neg/zero
  int* ninert = nullptr;   // [K*4] per-node pivot signs: pos, neg, zero, non-finite
  // Batches of independent blocks: block of every node, each block's own
  // pool regions (bump pointers tops[2b], tops[2b+1]) and the mask of the
  // blocks a factorization or solve touches (all ones = every block).  A
  // block left out keeps its fronts and factor untouched.
  int nblk = 0;
  const int* node_blk = nullptr;
  const int* blk_active = nullptr;
  const unsigned long long* blk_rbase = nullptr;
  const unsigned long long* blk_ibase = nullptr;
  const unsigned long long* blk_rcap = nullptr;
  const unsigned long long* blk_icap = nullptr;
};

struct FactorParams {
  double ustab;
  double pivot_tol;
  bool mixed;  // Schur updates of the blocked fronts in FP32 (MFMA)
};

// Kernel classes timed with HIP events when timing is on.
enum TimedKernel {
  TK_LEVEL = 0,
  TK_PANEL = 1,
  TK_TRAILING = 2,
  TK_SOLVE_FWD = 3,
  TK_SOLVE_BWD = 4,
  TK_ASSEMBLE = 5,
  TK_COUNT = 6
};

class LdlNumeric {
 public:
  LdlNumeric() = default;
  ~LdlNumeric();
  // Upload the plan; returns 0 or an error code (message in err).
  int setup(const SymbolicPlan& plan, const int* rowp, const int* cols,
            double delay_growth, std::string& err);
  int factor(const double* vals, const double* shift, double ustab,
             double pivot_tol, bool cholesky, bool mixed, hipStream_t stream,
             std::string& err);
  int solve(int nrhs, double* x, int ldx, hipStream_t stream,
            std::string& err);
  void inertia(int* pos, int* neg, int* zero) const {
    *pos = inertia_[0];
    *neg = inertia_[1];
    *zero = inertia_[2];
  }
  int pivot_counts(int* np, int* nd, std::string& err) const;
  // After a factorization without pivoting: 1-based index of the smallest
  // variable whose pivot has the wrong sign for a quasi-definite matrix
  // split at sqdef (positive below, negative from sqdef on), or 0.
  int sqd_check(int sqdef, int* first_bad, hipStream_t stream, std::string& err);
  // Independent diagonal blocks (e.g. many problems in one block-diagonal
  // KKT): node_seg[b]..node_seg[b+1] are block b's nodes.  After each
  // factorization the inertia is also summed per block, and a non-finite
  // pivot is reported per block instead of failing the factorization.
  int set_batch(int nblocks, const int* node_seg, std::string& err);
  // Blocks that the following factorizations and solves touch (host array
  // of nblocks ints, nonzero = on; null = all).  The others keep their
  // factors.
  int set_batch_mask(const int* mask, hipStream_t stream, std::string& err);
  int batch_blocks() const { return nbatch_; }
  void batch_inertia(int* pos, int* neg, int* zero, int* nan) const {
    for (int b = 0; b < nbatch_; b++) {
      if (pos) pos[b] = binert_[4 * b + 0];
      if (neg) neg[b] = binert_[4 * b + 1];
      if (zero) zero[b] = binert_[4 * b + 2];
      if (nan) nan[b] = binert_[4 * b + 3];
    }
  }
  void stats(int* delayed, int* max_front) const {
    *delayed = delayed_;
    *max_front = max_front_seen_;
  }
  double factor_flops() const { return flops_; }
  double solve_bytes() const { return solve_bytes_; }
  int num_levels() const { return int(level_ptr_.size()) - 1; }
  int max_front() const { return max_front_; }
  bool factored() const { return factored_; }
  // 0: none (the reference's behaviour); 1: symmetric (Ruiz) equilibration
  // of the shifted matrix before the factorization (inertia is unchanged).
  void set_scaling(int mode) { scaling_ = mode != 0; }
  // HIP-event timing of every launch, accumulated per kernel class.
  void set_timing(bool on) { timing_ = on; }
  // Replay captured HIP graphs of the factorization / solve launch
  // sequences (default on).
  void set_graphs(bool on) {
    graphs_ = on;
    if (!on) drop_graphs();
  }
  void get_timing(double* ms, long long* counts) const {
    for (int c = 0; c < TK_COUNT; c++) {
      ms[c] = tk_ms_[c];
      counts[c] = tk_n_[c];
    }
  }
  void reset_timing() {
    for (int c = 0; c < TK_COUNT; c++) {
      tk_ms_[c] = 0.0;
      tk_n_[c] = 0;
    }
  }
  // Algorithmic flops of the fronts factored by the workgroup kernel
  // (static size <= the LDS limit) and by the blocked path.
  double flops_small() const { return flops_small_; }
  double flops_big() const { return flops_big_; }

 private:
  int run_factor(const double* vals, const double* shift, FactorParams prm,
                 bool cholesky, bool sync_mode, hipStream_t stream,
                 std::string& err);
  int enqueue_factor(const double* vals, const double* shift, FactorParams prm,
                     bool cholesky, bool sync_mode, hipStream_t stream,
                     std::string& err);
  int finish_factor(hipStream_t stream, std::string& err);
  int enqueue_solve(int nrhs, double* x, int ldx, hipStream_t stream, std::string& err);
  void capture_factor(const double* vals, const double* shift, FactorParams prm,
                      bool cholesky, hipStream_t stream);
  void capture_solve(int nrhs, double* x, int ldx, hipStream_t stream);
  void drop_graphs();
  bool capture_stream();
  hipStream_t cap_stream_ = nullptr;

  // HIP graphs: the factorization's launch sequence (keyed by its buffers,
  // stream and parameters) and the solve sequences of recent right-hand
  // side buffers.
  struct SolveGraph {
    double* x;
    int nrhs, ldx;
    hipStream_t stream;
    int maxp;
    hipGraphExec_t exec;
    unsigned long long last_use;
  };
  static constexpr size_t kMaxSolveGraphs = 8;
  bool graphs_ = true;
  hipGraphExec_t fgraph_ = nullptr;
  const double* fg_vals_ = nullptr;
  const double* fg_shift_ = nullptr;
  hipStream_t fg_stream_ = nullptr;
  double fg_ustab_ = 0.0, fg_tol_ = 0.0;
  bool fg_chol_ = false, fg_mixed_ = false;
  std::vector<SolveGraph> sgraphs_;
  std::vector<SolveGraph> sseen_;   // eager solves seen once (no graph yet)
  unsigned long long sgraph_clock_ = 0;
  const double* fseen_vals_ = nullptr;
  const double* fseen_shift_ = nullptr;
  hipStream_t fseen_stream_ = nullptr;
  int ensure_pools(unsigned long long rcap, unsigned long long icap,
                   std::string& err);

  int n_ = 0, K_ = 0;
  DevPlan P_;
  DevState S_;
  std::vector<void*> allocs_;
  // Levels: nodes grouped by height in the node tree (leaves = 0).
  std::vector<int> level_ptr_, level_nodes_;
  std::vector<int> level_is_big_;   // per level: has big fronts
  std::vector<int> big_nodes_;      // per level list of big fronts
  std::vector<int> big_ptr_;
  std::vector<int> big_npanels_, big_tiles_;  // static launch plan per level
  std::vector<int> big_mmax_;       // rows of the largest planned blocked front, per level
  std::vector<int> big_wide_;       // level has a front too tall for the LDS panel
  std::vector<int> big_nth_;        // LDS-panel threads (= rows) of the level's launches
  std::vector<int> solve_wide_;     // level solved by 1,024-thread workgroups
  std::vector<int> h_kptr_;         // children pointer (host copy)
  std::vector<int> h_m_;            // front sizes read back in sync mode
  const int* d_big_nodes_ = nullptr;
  const int* d_big_root_ = nullptr;
  const int64_t* d_big_woff_ = nullptr;
  int64_t* d_big_woff_sync_ = nullptr;
  int* d_sync_nodes_ = nullptr;
  int* d_sync_root_ = nullptr;
  int delayed_ = 0, max_front_seen_ = 0;
  std::vector<int> level_lds_;      // LDS doubles of a level's larger fronts
  std::vector<int> level_split_;    // first larger front of each level
  std::vector<int> level_lds_small_;  // LDS doubles of a level's small fronts
  int* d_level_nodes_ = nullptr;
  int* d_big_mask_ = nullptr;       // [K] 1 if front is factored blocked
  std::vector<int> h_ns_, h_cs_;
  int max_front_ = 0;
  double* W_ = nullptr;             // panel workspace (max_m x 64)
  double* ub_ = nullptr;            // forward-solve update vectors (icap x nrhs)
  size_t ub_cap_ = 0;
  size_t W_cap_ = 0;
  double delay_growth_ = 2.0;
  unsigned long long base_rcap_ = 0, base_icap_ = 0;
  int inertia_[3] = {0, 0, 0};
  int nbatch_ = 0;
  int* d_batch_seg_ = nullptr;      // [nbatch+1]
  int* d_batch_inert_ = nullptr;    // [nbatch*4]
  std::vector<int> binert_;
  // Pinned host buffers for the device-to-host results and the mask
  // uploads (asynchronous copies whose host side outlives them).
  int* h_flags_pin_ = nullptr;      // [8]
  int* d_sqd_ = nullptr;            // sign-test result (quasi-definite check)
  int* h_binert_pin_ = nullptr;     // [4 * nbatch]
  char* h_stage_pin_ = nullptr;     // mask / compacted-list upload staging
  size_t h_stage_cap_ = 0;
  int stage(const void* src, size_t bytes, size_t& off, const void** out, std::string& err);
  std::vector<int> h_node_blk_, h_blk_active_;
  std::vector<unsigned long long> blk_rcap_h_, blk_icap_h_;
  int* d_node_blk_ = nullptr;
  int* d_blk_active_ = nullptr;
  unsigned long long* d_blk_layout_ = nullptr;  // rbase, ibase, rcap, icap (4 x nbatch)
  unsigned long long* d_tops_batch_ = nullptr;
  int layout_batch(std::string& err);
  // Launch lists of the blocks a mask leaves on (compacted on the host so
  // that left-out fronts cost no workgroups): level node lists with their
  // LDS-bucket splits, and the blocked-path plan of those fronts.
  bool mask_all_ = true;
  std::vector<int> m_level_ptr_, m_level_split_, m_level_nodes_, m_big_ptr_;
  std::vector<int> m_big_nodes_, m_big_root_;
  std::vector<int64_t> m_big_woff_;
  int* d_m_level_nodes_ = nullptr;
  int* d_m_big_nodes_ = nullptr;
  int* d_m_big_root_ = nullptr;
  int64_t* d_m_big_woff_ = nullptr;
  std::vector<int> h_big_root_;
  std::vector<int64_t> h_big_woff_;
  std::vector<int> h_nptr_, h_nvars_;
  int* d_row_blk_ = nullptr;        // [n] block of every row (batches)
  // Per node: the LDS the front's workgroup would get with its own block
  // alone (its level bucket's largest slot among that block's fronts), so
  // that the choice between the LDS and the global-memory factorization
  // does not depend on the other blocks of a batch.
  int* d_node_cap_ = nullptr;
  bool masked() const { return nbatch_ > 0 && !mask_all_; }
  double flops_ = 0.0, solve_bytes_ = 0.0;
  bool factored_ = false;
  int max_p_ = 0;                   // largest fully summed block seen
  bool scaling_ = false;
  bool timing_ = false;
  double tk_ms_[TK_COUNT] = {0, 0, 0, 0, 0, 0};
  long long tk_n_[TK_COUNT] = {0, 0, 0, 0, 0, 0};
  std::vector<hipEvent_t> ev_;
  std::vector<int> ev_class_;
  int ev_used_ = 0;
  double flops_small_ = 0.0, flops_big_ = 0.0;
  void tk_begin(hipStream_t s);
  void tk_end(hipStream_t s, int cls);
  void tk_collect();
  double* scale_ = nullptr;
  double* scale_tmp_ = nullptr;
  const int* d_rowp_ = nullptr;
  const int* d_cols_ = nullptr;
};

}  // namespace amigo_amd
