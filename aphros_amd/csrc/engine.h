// Host engine: resident systems, launch tables, captured iteration graphs.
//
// One Engine owns a HIP stream and any number of resident systems that share
// one method and precision.  A solve launches the same kernels for every
// active system at once (a single system is a batch of one).  Scalars stay on
// the device; the host only polls the per-system done flags between replays
// of a captured graph of `check_every` iterations.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/aphpois.h"
#include "common.h"
#include "kernels.h"

namespace aph {

struct Tuning {
  int lx = 0;  // threads per row of a stencil tile, 0 = automatic
  int bz = 0;  // planes per stencil workgroup, 0 = automatic
  int vx = 0;  // cells per thread in x (PCG), 0 = automatic
  int sm = 0;  // cells per thread of the streaming kernels, 0 = automatic
};

// Sets the message of aphpois_last_error() on the calling thread; returns code.
int set_api_error(int code, const std::string& msg);

// Launch tiling of one system: the stencil launch of the iteration kernel
// (vx cells per thread in x for PCG, 1 otherwise), the streaming launches and
// the residual launches (vx_ru cells per thread); fp64: the engine's
// vectors are FP64.  sys and blk0 are left to the caller.
void system_tiling(int method, int nx, int ny, int nz, int vx, const Tuning& tu, Rec* st,
                   Rec* sm, Rec* ru, int vx_ru = 1, int fp64 = 1);

struct System {
  int nx = 0, ny = 0, nz = 0;
  long long n = 0;
  double h3 = 1;
  bool has_system = false;
  bool singular = false;
  bool neg = false;
  bool has_m = false;
  bool active = false;  // takes part in the next solve
  bool have_x = false;  // x64/u holds a previous solution (warm start)
  int tune_lx = 0, tune_bz = 0;  // stencil tile of this system (0 = engine default)
  // multigrid levels (host copy; device copy in d_lev)
  std::vector<MgLevel> lev;
  MgLevel* d_lev = nullptr;
  bool done_reported = true;     // the last solve's result has been collected
  std::vector<void*> allocs;
  long long bytes = 0;
  double bnorm2 = -1;        // ||b||^2 of a PCG system (< 0: b changed since)
  double* xref = nullptr;    // reference solution for info.max_diff (or null)
  double* xref_store = nullptr;  // its device storage (kept once allocated)
  void* partials = nullptr;  // reduction scratch, outside the budget
  long long partials_cap = 0;
  SysDev dev{};
  aphpois_info info{};
  double seconds = 0;
};

// per-kernel timing slots of a multigrid PCG engine
constexpr int kProfSlots = 16;

class Engine {
 public:
  Engine() = default;
  ~Engine();
  // returns APHPOIS_* code; message in err()
  // precond (method PCG): 0 = Jacobi scaling, 1 = multigrid cycle; mg_w:
  // levels of the cycle that apply their coarse correction twice
  int init(int device, int method, int precision, int resmode, double tol, int miniter,
           int maxiter, int nullspace, double delta, int check_every, long long budget,
           int capacity, int precond = 0, int mg_w = 0);
  bool multigrid() const { return pc_; }
  // A budget shared with other engines on the same device (the server's
  // size-class lanes): every byte counted against the budget is reserved in
  // *used as well, so the engines' sum stays within budget.  Call before the
  // first allocation.
  void share_budget(std::atomic<long long>* used) { shared_ = used; }
  // A lock shared with other engines whose threads use the same device (the
  // server's lanes): step() holds it while it rebuilds the launch tables and
  // captures the iteration graph, and the other threads hold it around every
  // call that may allocate, free or synchronise outside their own stream, so
  // that no such call runs while a stream of this device is being captured
  // (it would invalidate the capture).  Graph replays run without it.
  void share_api_lock(std::mutex* mu) { api_mu_ = mu; }
  int add_system(int nx, int ny, int nz, double h3, int* slot);
  int set_system(int slot, const double* expr, bool on_device, long long start, long long ly,
                 long long lz);
  int set_rhs(int slot, const double* e7, bool on_device, long long start, long long ly,
              long long lz);
  // The projection's pressure system from density (or face denominators fk;
  // exactly one of the two), face volume fluxes and a volume source (dense
  // fields, x fastest), assembled on the device and made the resident system
  // as by set_system.
  int assemble(int slot, const double* rho, const double* fk, const double* ffv,
               const double* src, double cell, double dt, int sign, bool on_device);
  // Corrected face fluxes of the projection for the solution p.
  int flux_correct(int slot, const double* p, const double* rho, const double* fk,
                   const double* ffv, double* out, double cell, double dt, bool on_device);
  // x0: nullptr (zero guess), or host/device strided field; keep_x: warm start
  // from the resident solution.
  int prepare(int slot, const double* x0, bool on_device, long long start, long long ly,
              long long lz, bool keep_x);
  // Reference solution of a slot (host strided field; nullptr removes it):
  // every later solve reports max |x - xref| (mean removed when singular).
  int set_solution_ref(int slot, const double* xref, long long start, long long ly,
                       long long lz);
  int run();  // solves every prepared system
  int fetch(int slot, double* x, bool on_device, long long start, long long ly, long long lz,
            aphpois_info* info);
  int remove(int slot);
  void set_conf(double tol, int miniter, int maxiter) {
    tol_ = tol;
    miniter_ = miniter;
    maxiter_ = maxiter;
  }
  void set_tuning(const Tuning& t) {
    tuning_ = t;
    dirty_ = true;
  }
  int set_system_tile(int slot, int lx, int bz) {
    System* s = sys(slot);
    if (!s) return fail(APHPOIS_EINVAL, "no such slot");
    s->tune_lx = lx;
    s->tune_bz = bz;
    dirty_ = true;
    return APHPOIS_OK;
  }
  // Continuous batching: runs at most `max_replays` replays of the captured
  // iteration graph over the active systems and returns the slots that
  // finished (converged, hit maxiter or broke down) during this call; they
  // stay in the launch tables, idle, until prepared again or removed.
  int step(int max_replays, std::vector<int>* finished);
  // Next right-hand side of a correlated sequence, on the device (PCG):
  // b += amp * max|b| * (eta - mean eta), eta uniform in [-1, 1] from the
  // splitmix64 stream `key`.
  int perturb_rhs(int slot, double amp, unsigned long long key);
  // z = M r with the multigrid preconditioner of the resident system (dense
  // host arrays, x fastest); leaves any solve state of the slot as it was.
  int precond_apply(int slot, const double* r, double* z);
  long long bytes() const;
  int nslots() const { return (int)slots_.size(); }
  System* sys(int slot) {
    return slot >= 0 && slot < (int)slots_.size() ? slots_[slot].get() : nullptr;
  }
  const std::string& err() const { return err_; }
  hipStream_t stream() const { return st_; }
  int method() const { return method_; }
  int vx() const { return vx_; }
  // launch geometry of the last solve (for measurement)
  int blocks_stencil() const { return L_.nst; }
  int blocks_stream() const { return L_.nsm; }
  // timing of the iteration kernels of the last run (device events), ms
  double last_run_ms() const { return last_run_ms_; }
  int last_run_iters() const { return last_run_iters_; }
  // Per-kernel timing: events recorded between the kernels of every captured
  // iteration; averages over the iterations that ran (not the no-op tail).
  void set_profile(bool on) {
    profile_ = on;
    dirty_ = true;
    prof_reset();
  }
  void prof_reset() {
    for (int j = 0; j < kProfSlots; ++j) {
      prof_ms_[j] = 0;
      prof_n_[j] = 0;
    }
  }
  // iteration kernels, plus (multigrid) [9] level-0 sweep, [10] level-0
  // restriction, [11] coarser levels per cycle, [12] bottom levels,
  // [13] [14] [15] level-0 pre-sweep, first and second post-sweep.  PCG
  // times the update's finalize and the reliable-update kernels after it
  // together, in [3] (no events between them: each event between launches
  // costs the event-timed solve time of its own); [4]..[7] stay empty.
  bool prof_timed(int j) const { return method_ != kPcg || j <= 3 || j >= 8; }
  int prof_kernels() const { return pc_ ? kProfSlots : iter_kernels(method_, 0); }
  double prof_avg_ms(int j) const { return prof_n_[j] ? prof_ms_[j] / prof_n_[j] : 0.0; }
  long long prof_count(int j) const { return prof_n_[j]; }

 private:
  int fail(int code, const std::string& msg);
  int hip(hipError_t e, const char* what);
  // budget accounting: reserve() fails if nbytes more would exceed it
  bool reserve(long long nbytes);
  void release(long long nbytes);
  int alloc(System& s, void** p, long long bytes);
  int build_tables();
  int capture();
  int pull_scalars();
  int run_pending_ru();
  void fill_info(int slot, double seconds);
  // info.max_diff of a finished slot against its reference solution
  int diff_info(int slot);
  bool need_start_ = false;
  int upload_strided(const double* src, bool on_device, long long start, long long ly,
                     long long lz, int nx, int ny, int nz, int comps, const double** dptr);
  void free_system(System& s);
  int ensure_staging(long long bytes);
  // set_system from a dense-strided FP64 Expr already in device memory
  int set_system_staged(int slot, const double* d, long long ly, long long lz);

  int device_ = 0;
  int ncu_ = 256;  // compute units of the device
  int method_ = kConjugate;
  int fp32_ = 0;
  bool pc_ = false;
  int mg_w_ = 0;
  int resmode_ = 0;
  double tol_ = 0;
  int miniter_ = 1, maxiter_ = 100;
  int nullspace_ = 0;
  double delta_ = 0.1;
  int check_every_ = 0;
  long long budget_ = 0;
  std::atomic<long long>* shared_ = nullptr;
  std::mutex* api_mu_ = nullptr;
  int capacity_ = 0;
  int vx_ = 1;
  Tuning tuning_;
  hipStream_t st_ = nullptr;
  std::vector<std::unique_ptr<System>> slots_;
  // device tables
  Scalars* d_sc_ = nullptr;
  Scalars* h_sc_ = nullptr;  // pinned mirror
  unsigned* d_counter_ = nullptr;
  SysDev* d_sds_ = nullptr;
  Rec* d_rec_ = nullptr;
  int* d_b2r_ = nullptr;
  long long rec_cap_ = 0, b2r_cap_ = 0;
  LaunchSet L_{};
  std::vector<int> launch_slots_;
  bool dirty_ = true;
  hipGraphExec_t gexec_ = nullptr;
  hipGraph_t graph_ = nullptr;
  int graph_k_ = 0;
  // event profiling: the cycle below level 0 as one graph (LaunchSet::below0)
  hipGraphExec_t below_exec_ = nullptr;
  hipGraph_t below_graph_ = nullptr;
  int prof_nl_ = 0;  // events of one profiled cycle minus one
  // scratch
  void* staging_ = nullptr;
  long long staging_bytes_ = 0;
  unsigned* d_flags_ = nullptr;
  double* d_red_ = nullptr;  // reduction scratch (2056 doubles)
  unsigned* d_red_counter_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  bool profile_ = false;
  std::vector<hipEvent_t> pev_;
  std::vector<hipEvent_t> pev_mg_;  // V-cycle launches (profiling)
  std::vector<int> mg_kinds_;
  double prof_ms_[kProfSlots] = {};
  long long prof_n_[kProfSlots] = {};
  double last_run_ms_ = 0;
  int last_run_iters_ = 0;
  std::string err_;
};

}  // namespace aph
