// Pressure-Poisson engine: matrix-free fine operator, Krylov / relaxation
// solvers and the aggregation-multigrid V-cycle, all resident on the GPU.
#pragma once

#include <vector>
#include "internal.hpp"
#include "vcycle_tile.hpp"

namespace pcfd {

// Per-cell bits of the fine Poisson operator (pycfd/core/pressure.py:209-251).
enum FineFlag : uint8_t {
  F_SOLV = 1,   // solvable cell: fluid and not enclosed
  F_L = 2,      // face to the x-1 neighbour carries flux
  F_R = 4,      // face to the x+1 neighbour
  F_B = 8,      // face to the y-1 neighbour
  F_T = 16,     // face to the y+1 neighbour
};

// Matrix-free 5-point operator on the interior cells, k = i*ny + j.
struct FineOp {
  const uint8_t* flags;
  int nx, ny;
  double cx, cy;  // 1/dx^2, 1/dy^2
  int px, py;     // periodic axes
  int dir[4];     // Dirichlet (pressure-outlet) walls: left, right, bottom, top
  double pref[4];
  // Stretched mesh (pressure.py:193-207), null on a uniform one: the face
  // coefficient of x face m (0..nx) in row j is hyc[j] / hxv[m], of y face n
  // (0..ny) in column i is hxc[i] / hyv[n], with hxv = hxu * vbar and
  // hyv = hyv_centre * vbar (vbar the mean cell area); row i, j is scaled by
  // hxc[i] * hyc[j] / vbar.
  const double* hxc;  // nx cell widths
  const double* hyc;  // ny cell heights
  const double* hxv;  // nx+1
  const double* hyv;  // ny+1
  double vbar;
};

// A Galerkin coarse level on its own rectangular grid, stored as a row
// dictionary: the level's operator rows take only a few hundred distinct
// values (interior rows repeat by translation invariance; the rest sit near
// walls and bodies), so each cell keeps a 2-byte class index and each class
// keeps only its nonzeros (neighbour offset, coefficient), resident in L2.
// `agg[k]` is the coarse cell that aggregates cell k (-1 if k is inactive),
// i.e. the tentative prolongator of the next level.
template <typename T>
struct ClassLevel {
  int nx, ny, n;
  int px, py;
  const uint16_t* cls;     // per cell
  const uint32_t* cls32;   // per cell, in place of cls past 65535 classes
  const int* row_ptr;      // per class, n_classes + 1
  const int* ent_lin;      // per entry: di*ny + dj (non-periodic neighbour)
  const int8_t* ent_di;    // per entry, for periodic wrap
  const int8_t* ent_dj;
  const T* ent_val;        // per entry coefficient
  const T* cdinv;          // per class 1/diag
  const uint8_t* cactive;  // per class
  // The same resolved per cell, so a row needs no class lookup before its
  // entries: first / one-past-last entry, 1/diag and active flag.
  const int2* rr;
  const T* rdinv;
  const uint8_t* ract;
  int fx, fy, cny;  // aggregation towards the next level (prolongation index)
  int spec_e0, spec_e1;  // entry range of the level's most frequent row
};

// The last solutions of one substep's solve, for the warm start by
// extrapolation in time: kWarmSlots arrays of nx*ny (a ring) and its state
// idx = {newest slot, count, next slot, next count}.  Null slots: no history.
constexpr int kWarmMaxOrder = 4;  // orders 5 and 6 measured no better
// order o reads o + 1; the slot written next is never read
constexpr int kWarmSlots = kWarmMaxOrder + 2;
struct WarmHistory {
  double* slots = nullptr;
  int* idx = nullptr;
};

// Device-resident scalars of one solve.  Kernels read them directly so the
// host never has to synchronise inside an iteration.
struct SolveState {
  double mean_b;     // fluid mean of the assembled right-hand side
  double bnorm;      // ||b||
  double atol;       // CG stopping threshold on the recursive ||r||
  double rr;         // ||r||^2 of the current iterate
  double rho, rho_prev, beta, alpha;
  double mean_r;     // mean of r over the active set
  double mean_x;     // pending constant shift of x (subtracted on active cells)
  double mean_z;     // active-set mean of the last V-cycle output
  double sum_red;    // red-colour sum after a red SOR sweep
  double res, res_prev;
  double res_final;  // relative true residual of the returned iterate
  double mean_final; // fluid mean removed by the normalisation
  double rho_hist[2];  // multigrid-CG rho of iteration i at [i & 1] (read by the next one)
  int iter;          // completed iterations / sweeps / cycles
  int done;
  int converged;
  int stalled;
  int zero_rhs;
  int pad[3];
};

// What the final residual check of one solve records for the host: a
// captured step reads these after it ran (one record per solve).
struct SolveLog {
  int iter;        // iterations / sweeps / cycles completed
  int converged;
  int done;        // 0: the iteration budget ran out before the solve decided
  int pad;
  double res;      // the last check's relative residual (relaxation solvers)
  double res_final;  // the relative true residual of the returned iterate
};

// Partial sums taken over the output of a V-cycle: none, sum(z) for the
// standalone solver's mean removal, or sum(z), sum(b*z), sum(b) for the
// preconditioned-CG inner product.
enum SmoothReduce : int { SR_NONE = 0, SR_MG = 1, SR_MGCG = 2 };

enum SolverKind : int {
  SK_DIRECT = 0,
  SK_CG = 1,
  SK_JACOBI = 2,
  SK_SOR = 3,
  SK_MULTIGRID = 4,
  SK_MGCG = 5,
};

// Host description of one multigrid level handed in by the setup code.
struct MGLevelHost {
  int nx, ny;
  int fx, fy;                    // aggregation factors towards the next level
  std::vector<int8_t> di, dj;    // stencil offsets (coarse levels)
  std::vector<uint16_t> cls;     // per-cell class (coarse levels)
  std::vector<uint32_t> cls32;   // the same past 65535 classes (stretched meshes)
  std::vector<double> table;     // n_classes x S coefficients (dense rows)
  std::vector<double> cdinv;     // per-class 1/diag
  std::vector<uint8_t> cactive;  // per-class active flag
  std::vector<uint8_t> active;   // per-cell active flag (all levels)
  double weight;
  std::vector<double> coarse_inv;  // dense pseudo-inverse (coarsest level)
};

struct PoissonConfig {
  int kind;
  double tol;
  int maxiter;
  double omega;     // SOR over-relaxation / Jacobi damping
  int sweeps;       // multigrid pre/post smoothing sweeps
  int precision;    // 0 = FP64 throughout, 1 = FP32 V-cycle inside FP64 CG
  int fused = 1;    // stored levels through the fused kernels when they fit
  int tile_tx = 0;  // level-0 fused tile (0 = chosen by the cost model)
  int tile_ty = 0;
  int chain = 0;    // cooperative coarse chain: 0 measured, -1 never, k >= 1 from level k
  int stream = 0;   // level-0 streaming legs: 0 measured, -1 never, k > 0 with k rows
  int row_reg = 0;  // stored-level row passes: 0 register-resident where rows fit, -1 looping
};

struct SolveInfo {
  int iterations;
  double residual;
  int converged;
  int status;
};

// Device buffers of one multigrid level (vectors in the cycle's precision).
struct LevelBuf {
  int nx = 0, ny = 0, n = 0, fx = 1, fy = 1;
  void *b = nullptr, *x = nullptr, *y = nullptr, *r = nullptr, *t = nullptr;
  int ncls = 0, nnz = 0, max_row = 0, min_row = 0, lanes = 1;
  uint16_t* cls = nullptr;
  uint32_t* cls32 = nullptr;  // wide class index (never run through the fused kernels)
  int2* rr = nullptr;         // per cell: entry range of its class row
  void* rdinv = nullptr;      // per cell 1/diag
  uint8_t* ract = nullptr;    // per cell active flag
  int *row_ptr = nullptr, *ent_lin = nullptr, *agg = nullptr;
  int8_t *ent_di = nullptr, *ent_dj = nullptr;
  void *ent_val = nullptr, *cdinv = nullptr;
  uint8_t* cactive = nullptr;
  double weight = 1.0;
  double* cinv = nullptr;  // coarsest dense pseudo-inverse (FP64)
  int coarse_n = 0;
  // Fused level kernels (vcycle.hip): tilings and the stencil radius; level 0
  // has one down tiling per start (zero / nonzero iterate).  `fused` is 0 for
  // a stored level whose LDS region does not fit (it then runs pass by pass).
  int fused = 0;
  int rxs = 1, rys = 1;
  FusedTile down0, down1, up;
  // Level 0 only: rows per wavefront of the streaming Krylov-start legs (0:
  // the tiled kernels run them) and the up leg's workgroup count.
  int stream_rows = 0, stream_nb = 0;
  // Stored levels of radius <= 3: the box dictionary of the streaming legs
  // (classes + one zero class) and the rows per wavefront when they run the
  // level (0: fused tiles or pass by pass).
  void* box = nullptr;
  void* box_dinv = nullptr;
  uint8_t* box_act = nullptr;
  int box_radius = 0, box_rows = 0;
  int* ent_off = nullptr;
  std::vector<int8_t> hdi, hdj;  // host copy of the entry offsets (tile planning)
  int hot = -1, hot_n = 0;       // most frequent class and its entry count
  int hot_e0 = 0;                // its first entry
  // Register-resident row passes: lanes per row (0: the looping passes run
  // the level) and the entry range of the level's most frequent row.
  int row_lanes = 0;
  int spec_e0 = 0, spec_e1 = 0;
};

class PoissonEngine {
 public:
  PoissonEngine() = default;
  ~PoissonEngine();
  int init(const FineOp& op, int n_fluid, int singular, hipStream_t stream);
  int configure(const PoissonConfig& cfg, const std::vector<MGLevelHost>& levels);
  // b_raw: assembled right-hand side (fluid rows, Dirichlet terms added, not
  // yet mean-free).  The solution overwrites `x_` (the warm-start buffer).
  int prepare_from_velocity(const double* u, const double* v, int ldu, int ldv, double dx,
                            double dy, const StepScalars& T);
  int prepare_from_rhs(const double* rhs);
  int solve(SolveInfo* info, bool sync_info);
  // ghosted p, ghost ring zeroed; with a history the solution also goes to
  // its ring (the warm start by extrapolation)
  int scatter_pressure(double* p, int ldp, const WarmHistory& wh = WarmHistory{});
  // the guess of the next solve extrapolated from the history (order 1-3)
  int warm_guess(const WarmHistory& wh, int order);
  int write_interior(double* out);            // normalised (nx, ny)
  int reset_guess();
  int set_guess(const double* x0);
  int apply_operator(const double* x, double* y);
  int apply_preconditioner(const double* r, double* z);
  int read_state(SolveState* st);
  // Captured solves: `iters` iterations are enqueued without polling (the
  // kernels past convergence are no-ops); 0 polls as usual.
  void set_budget(int iters) { budget_ = iters; }
  // The final check of the solves that follow writes its record to log
  // slot `slot` (-1: none); read_log copies n records to the host (syncs).
  void set_log_slot(int slot) { log_slot_ = slot; }
  int read_log(SolveLog* out, int n);
  // Whether info.residual reports the final true residual (Krylov solvers)
  // or the last check's (relaxation solvers and standalone multigrid).
  bool reports_final_residual() const {
    return cfg_.kind == SK_CG || cfg_.kind == SK_MGCG || cfg_.kind == SK_DIRECT;
  }
  int maxiter() const { return cfg_.maxiter; }
  bool clock_on() const { return clock_on_; }
  void set_stream(hipStream_t s) { s_ = s; }
  static constexpr int kLogSlots = 8;
  // Times `reps` back-to-back launches of one hot kernel on this engine's
  // buffers with HIP events; reports the mean duration and the kernel's
  // algorithmic bytes per launch (arrays read + written once, at stored size).
  int time_kernel(int which, int reps, double* avg_ms, double* bytes);
  // Live clock: every kernel launch of the solves that follow records
  // start/stop events (up to `max_launches` launches in all), filed under a
  // slot = kind * kClockLevels + level (ClockKind); clock_read stops it and
  // reports per slot the launches, mean duration and algorithmic bytes per
  // launch.  Iterations enqueued past convergence (no-ops) are dropped.
  // The step's other kernels (predictors, boundary passes, right-hand side,
  // start, final check, scatter, correction, stage combination,
  // diagnostics, warm-start guess) are filed under their own kinds at level 0; the solve's
  // once-per-solve folds under CK_FOLD at levels kFoldPrepare.. .
  enum ClockKind : int {
    CK_UPDATE = 0, CK_DIR, CK_FINE_DOWN, CK_FINE_UP, CK_CLASS_DOWN, CK_CLASS_UP,
    CK_ROWS_SMOOTH0, CK_ROWS_SMOOTH, CK_ROWS_RES, CK_ROWS_PROLONG, CK_RESTRICT, CK_DENSE,
    CK_CHAIN, CK_FOLD, CK_PREDICTOR, CK_PREDICTOR_COMB, CK_WALLS, CK_BPASS, CK_PREPARE, CK_INIT,
    CK_CHECK, CK_SCATTER, CK_CORRECT, CK_COMBINE, CK_DIAG, CK_SMALL_CG, CK_WARM, kClockKinds
  };
  enum FoldLevel : int { kFoldPrepare = 3, kFoldInit, kFoldCheck, kFoldDiag };
  static constexpr int kClockLevels = 16, kClockSlots = kClockKinds * kClockLevels;
  int clock_start(int max_launches);
  // A clock for one launch outside the iteration loop (never trimmed); no
  // clock when the live clock is off.
  LaunchClock clock_step(int kind, double bytes, int level = 0);
  // The same as a ClockSource for launchers that issue several kernels:
  // walls and boundary-pass launches (kClockWalls, kClockBoundaryPass).
  ClockSource boundary_clocks() {
    return clock_on_ ? ClockSource{&PoissonEngine::boundary_clock, this} : ClockSource{};
  }
  int clock_read(int32_t* launches, double* avg_ms, double* bytes);
  // Per level: nx, ny, fused, tile x, tile y (see pcfd_mg_plan).
  int plan(int32_t* out, int cap) const {
    const auto& lv = (cfg_.precision == 1) ? lv32_ : lv64_;
    int k = 0, l = 0;
    for (const LevelBuf& L : lv) {
      int mode = (chain_first_ >= 0 && l >= chain_first_) ? 2 : L.fused;
      int sx = 0, sy = 0;  // streaming: rows per wavefront x output columns
      if (l == 0 && L.stream_rows) {
        mode = 3;
        sx = L.stream_rows;
        sy = 60;
      } else if (mode != 2 && L.box_rows) {
        mode = 3;
        sx = L.box_rows;
        sy = 64 - 4 * L.box_radius;
      }
      ++l;
      const int v[5] = {L.nx, L.ny, mode, mode == 1 ? L.down0.tx : sx,
                        mode == 1 ? L.down0.ty : sy};
      for (int q = 0; q < 5; ++q)
        if (k < cap) out[k++] = v[q];
    }
    return (int)lv.size();
  }
  double* guess() { return x_; }
  int last_iterations() const { return last_iters_; }
  hipStream_t stream() const { return s_; }
  int n() const { return n_; }
  int levels() const { return (int)lv64_.size(); }

  double *partial_ = nullptr;  // per-tile partial sums of the two-launch reductions
  // Multigrid-CG: the last update's partials not yet folded (rows > 0), for
  // the next cycle's streaming down leg to fold (UpdateFold), with the
  // iteration count that update completes.
  const double* upd_partial_ = nullptr;
  int upd_rows_ = 0, upd_iter_ = 0;
  int flush_update_fold();
  size_t partial_cap_ = 0;     // capacity in doubles
  int partial_gen_ = 0;        // bumped when partial_ moves (captured graphs hold it)
  unsigned int* counter_ = nullptr;

 private:
  int run_cg();
  // Jacobi-CG on grids of at most kSmallCgCells cells as one workgroup that
  // keeps the whole solve on one CU (0: not eligible; else cells per thread).
  int small_cg_cells_per_thread() const;
  int run_mgcg();
  int run_sor();
  int run_jacobi();
  int run_multigrid();
  int finalize(bool recompute_residual);
  // V-cycle entry modes: from the Krylov residual r_ (zero start, output
  // projected by the caller's sums) or from lv[0].b / lv[0].x (standalone).
  enum CycleMode : int { VC_KRYLOV = 0, VC_STANDALONE = 1 };
  // defer_fold: leave the cycle's partials (cycle_rows() rows of 3) for the
  // next kernel to fold instead of launching the fold.
  int enqueue_vcycle(int mode, int reduce, bool defer_fold = false);
  int cycle_rows() const;
  template <typename T>
  int enqueue_vcycle_t(int mode, int reduce, bool defer_fold);
  template <typename T>
  int plan_level_tiles(std::vector<LevelBuf>& lv);
  template <typename T>
  int level_down(LevelBuf& L, LevelBuf& C, const SolveState* st, int lev = 0);
  template <typename T>
  int level_up(LevelBuf& L, LevelBuf& C, const SolveState* st, int lev = 0);
  template <typename T>
  int tune_level(std::vector<LevelBuf>& lv, int l, SolveState* tst);
  int ensure_partials(size_t n);
  template <typename T>
  int tune_dir(SolveState* tst);
  template <typename T>
  int coarse_legs(std::vector<LevelBuf>& lv, int first, const SolveState* st);
  template <typename T>
  int launch_chain(std::vector<LevelBuf>& lv, int first, const SolveState* st);
  template <typename T>
  int tune_chain(std::vector<LevelBuf>& lv, SolveState* tst);
  int poll(int* iter, int* done, int* converged);
  int start(int mode, double* r);
  void free_levels();

  FineOp op_{};
  hipStream_t s_ = nullptr;
  int n_ = 0, n_fluid_ = 0, singular_ = 0;
  PoissonConfig cfg_{};
  SolveState* st_ = nullptr;
  SolveState* st_host_ = nullptr;  // pinned mirror for polling
  hipEvent_t ev_ = nullptr;
  // live kernel clock
  struct ClockEntry {
    int slot, iter;
    hipEvent_t a, b;
    double bytes;
  };
  bool clock_on_ = false;
  std::vector<hipEvent_t> clock_pool_;
  size_t clock_pool_used_ = 0;
  std::vector<ClockEntry> clock_log_;
  int clock_iter_ = 0;  // iteration tag of the launches being enqueued
  LaunchClock clock_next(int kind, int level, double bytes);
  static LaunchClock boundary_clock(void* self, int kind, double bytes);
  void clock_free();
  // Launches of iterations past the solve's `iters` (tagged after `it0`) ran
  // as no-ops: drop them.
  void clock_trim(int it0, int iters) {
    for (ClockEntry& e : clock_log_)
      if (e.iter > it0 + iters) e.slot = -1;
  }
  double *b_ = nullptr, *x_ = nullptr, *r_ = nullptr, *pa_ = nullptr, *pb_ = nullptr,
         *q_ = nullptr, *z_ = nullptr;
  int last_iters_ = 0;
  int budget_ = 0;     // captured solves: iterations enqueued without polling
  int log_slot_ = -1;
  SolveLog* log_ = nullptr;       // kLogSlots records (device)
  SolveLog* log_host_ = nullptr;  // pinned mirror
  int predicted_ = 0;  // iterations of the last converged solve (first chunk)
  int last_run_ = 0;   // iterations of the last solve

  // multigrid
  std::vector<LevelBuf> lv64_;
  std::vector<LevelBuf> lv32_;
  uint8_t* fine_active_ = nullptr;
  int n_active0_ = 0;
  int n_proj_ = 0;  // cells of the constant mode removed by projections (0: non-singular)
  FusedTile dir_tile_;  // level-0 tiling of the multigrid-CG direction kernel
  int chain_first_ = -1;  // first level of the cooperative coarse chain (-1: none)
  int chain_blocks_ = 0;  // its grid size (co-resident workgroups)
  double weight0_ = 1.0;
};

}  // namespace pcfd
