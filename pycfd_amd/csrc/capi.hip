// C-ABI of libpycfd_hip.so: the simulation context and the time step.
//
// A context owns everything one ProjectionSolver (pycfd/core/solver.py:82)
// needs on the device: the immersed-boundary face lists, the Poisson
// operator flags, the pressure engine and the Runge-Kutta stage buffers.
// Field arrays (u, v, p) belong to the caller and are passed per call.
#include <cstdarg>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/pycfd_hip.h"
#include "poisson.hpp"

namespace pcfd {

static thread_local char g_last_error[1024] = "";

void set_last_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

}  // namespace pcfd

using namespace pcfd;

// Scratch of the allocation-free diagnostics (pcfd_diagnostics_noalloc): the
// per-workgroup partial sums of the reduction live in this fixed array,
// which every GPU the library runs on holds until the process ends.  Host
// code never uses the array's name as an address (on the host it is only a
// stand-in for the symbol): each device's own copy is resolved once with
// hipGetSymbolAddress and that pointer is what kernels receive.
constexpr size_t kNoAllocScratchBytes = PCFD_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, resolved on the first call for
// that device.
int scratch_for_current_device(double** out) {
  int dev = 0;
  PCFD_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_last_error("device %d is beyond the %d devices the scratch table holds", dev,
                   kMaxDevices);
    return kErrUnsupported;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_last_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      return kErrHip;
    }
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

}  // namespace

struct pcfd_ctx {
  int nx = 0, ny = 0;
  int ldu = 0, ldv = 0, ldp = 0;
  double dx = 0, dy = 0, nu = 0, fx = 0, fy = 0;
  int scheme = PCFD_SCHEME_RK3;
  int kind[4] = {0, 0, 0, 0};
  int any_outlet = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  UpdRange upd{};
  BCParams bc{};
  MomParams mom{};
  int n_bodies = 0;
  // device buffers owned by the context
  double* inlet[4] = {nullptr, nullptr, nullptr, nullptr};
  int* uf = nullptr;
  int* vf = nullptr;
  int8_t* uf_body = nullptr;
  int8_t* vf_body = nullptr;
  double* force = nullptr;
  double* nuc = nullptr;  // Smagorinsky nu + nu_t, ghosted cell layout (LES only)
  // Stretched mesh: one allocation holding the metric arrays, the Poisson
  // face denominators and the immersed-boundary face areas.
  double* metric = nullptr;
  Metrics g{};
  const double* poisson_hxv = nullptr;
  const double* poisson_hyw = nullptr;
  LesParams les{};
  bool use_les = false;
  uint8_t* flags = nullptr;
  uint8_t* solid = nullptr;
  double* stage_a[2] = {nullptr, nullptr};
  double* stage_b[2] = {nullptr, nullptr};
  double* diag_out = nullptr;
  double* guess_backup = nullptr;
  // warm start by extrapolation (pcfd_set_warm_start): per substep a ring of
  // kWarmSlots solutions and its state (4 ints); a backup of the state for a
  // captured step that is redone
  int warm_order = 0;
  double* warm_slots = nullptr;
  int* warm_idx = nullptr;
  int* warm_idx_backup = nullptr;
  double* host_out = nullptr;  // pinned
  double* scratch = nullptr;   // this device's copy of the no-allocation scratch
  double* soft_work = nullptr; // (nx*ny) cells: the transposed projection's right-hand side
  bool soft_ib = false;
  PoissonEngine engine;
  bool solver_set = false;
  // Captured steps (pcfd_set_step_graphs): one graph per buffer set, the
  // iteration budget of each substep's solve from the last measured counts.
  struct StepGraph {
    const double *u0 = nullptr, *v0 = nullptr;
    double *u = nullptr, *v = nullptr, *p = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t scalars = nullptr;  // k_set_step_scalars, updated per replay
    hipKernelNodeParams kp{};
    int budget[3] = {0, 0, 0};
    int partial_gen = 0;
    uint64_t used = 0;
    bool stale = false;
    void release() {
      if (exec) hipGraphExecDestroy(exec);
      if (graph) hipGraphDestroy(graph);
      exec = nullptr;
      graph = nullptr;
    }
  };
  std::vector<StepGraph> graphs;
  uint64_t graph_tick = 0;
  bool use_graphs = false;
  int pred_sub[3] = {0, 0, 0};    // iterations each substep's solve took last
  hipStream_t cap_stream = nullptr;
  double* scal = nullptr;         // device [dt, blend] read by captured kernels
  int graph_launches = 0, graph_captures = 0, graph_fallbacks = 0;

  ~pcfd_ctx() {
    for (auto* p : inlet) hipFree(p);
    hipFree(uf);
    hipFree(vf);
    hipFree(uf_body);
    hipFree(vf_body);
    hipFree(force);
    hipFree(flags);
    hipFree(solid);
    for (auto* p : stage_a) hipFree(p);
    for (auto* p : stage_b) hipFree(p);
    hipFree(diag_out);
    hipFree(guess_backup);
    hipFree(warm_slots);
    hipFree(warm_idx);
    hipFree(warm_idx_backup);
    hipFree(nuc);
    hipFree(metric);
    if (host_out) hipHostFree(host_out);
    for (StepGraph& g : graphs) g.release();
    hipFree(scal);
    hipFree(soft_work);
    if (cap_stream) hipStreamDestroy(cap_stream);
  }
};

namespace {

BCOps ops_of(std::initializer_list<int> l) {
  BCOps o{};
  for (int x : l) o.op[o.n++] = x;
  return o;
}

int boundary(pcfd_ctx* c, std::initializer_list<int> l, const StepScalars& S, double* u,
             double* v) {
  return launch_boundary_pass(c->stream, c->bc, ops_of(l), S, u, v, c->engine.boundary_clocks());
}

// The live kernel clock's entry for one step kernel, with its algorithmic
// bytes per cell (each array read or written once at its stored size).
LaunchClock step_clock(pcfd_ctx* c, int kind, double bytes_per_cell) {
  return c->engine.clock_step(kind, bytes_per_cell * (double)c->nx * (double)c->ny);
}

int boundary(pcfd_ctx* c, std::initializer_list<int> l, double dt, double* u, double* v) {
  return boundary(c, l, StepScalars{dt, 0.0, nullptr}, u, v);
}

template <typename T>
int upload(T** dst, const T* src, size_t n) {
  PCFD_HIP_CHECK(hipMalloc(dst, std::max<size_t>(n, 1) * sizeof(T)));
  if (n) PCFD_HIP_CHECK(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return kOk;
}

int check_ctx(pcfd_ctx* c) {
  if (!c) {
    set_last_error("null context");
    return kErrBadArg;
  }
  return kOk;
}

int check_solver(pcfd_ctx* c) {
  if (!c->solver_set) {
    set_last_error("no pressure solver configured (pcfd_set_pressure_solver)");
    return kErrBadArg;
  }
  return kOk;
}

void fill_info(pcfd_solve_info* info, const SolveInfo& s) {
  if (!info) return;
  info->iterations = s.iterations;
  info->residual = s.residual;
  info->converged = s.converged;
  info->status = s.status;
}

// One projected Euler substep (solver.py:612-640).  `ue` non-null selects
// the fused stage combination ca*u + cb*ue as the substep input.
// The warm-start history of substep k (none when the warm start is the
// previous pressure).
WarmHistory warm_of(const pcfd_ctx* c, int k) {
  WarmHistory h;
  if (c->warm_order > 0 && c->warm_slots) {
    h.slots = c->warm_slots + (size_t)k * kWarmSlots * c->nx * c->ny;
    h.idx = c->warm_idx + 4 * k;
  }
  return h;
}

// Forget the extrapolation history (a new start for the guesses).
int warm_reset(pcfd_ctx* c) {
  if (c->warm_idx) PCFD_HIP_CHECK(hipMemsetAsync(c->warm_idx, 0, 12 * sizeof(int), c->stream));
  return kOk;
}

int substep(pcfd_ctx* c, const double* u, const double* v, const double* ue, const double* ve,
            double ca, double cb, const StepScalars& S, double* us, double* vs, double* p,
            SolveInfo* info, bool sync, bool last = true, int sidx = 0) {
  // The reaction force a substep books is overwritten by the next substep's
  // (solver.py:577-583, reset after each predictor), so only the step's last
  // substep sums it; the others just zero the faces.
  const int mask_pred = last ? BOP_MASK_RESET : BOP_MASK_ZERO;
  const int mask_corr = last ? BOP_MASK_ACC : BOP_MASK_ZERO;
  int rc;
  if (c->use_les &&
      (rc = launch_eddy_viscosity(c->stream, c->les, u, ue, v, ve, ca, cb, c->nuc, nullptr)))
    return rc;
  if (ue) {  // u0, ue, v0, ve in; u*, v* out
    rc = launch_predictor_comb(c->stream, u, ue, v, ve, ca, cb, c->mom, c->upd, S, us, vs,
                               step_clock(c, PoissonEngine::CK_PREDICTOR_COMB, 48));
  } else {  // u, v in; u*, v* out
    rc = launch_predictor(c->stream, u, v, c->mom, c->upd, S, us, vs,
                          step_clock(c, PoissonEngine::CK_PREDICTOR, 32));
  }
  if (rc) return rc;
  if (c->any_outlet) {
    rc = boundary(c, {BOP_VEL_PRED, BOP_MASS, mask_pred, BOP_WRAP_FACES}, S, us, vs);
  } else {
    rc = boundary(c, {BOP_VEL_PRED, mask_pred, BOP_WRAP_FACES}, S, us, vs);
  }
  if (rc) return rc;
  const WarmHistory wh = warm_of(c, sidx);
  if (wh.slots && (rc = c->engine.warm_guess(wh, c->warm_order))) return rc;
  if ((rc = c->engine.prepare_from_velocity(us, vs, c->ldu, c->ldv, c->dx, c->dy, S))) return rc;
  if ((rc = c->engine.solve(info, sync))) return rc;
  if ((rc = c->engine.scatter_pressure(p, c->ldp, wh))) return rc;
  // p, u*, v* in; u, v out
  if ((rc = launch_correct_ghosts(c->stream, p, c->ldp, c->bc, c->ldu, c->ldv, c->dx, c->dy,
                                  c->g, c->upd, S, us, vs,
                                  step_clock(c, PoissonEngine::CK_CORRECT, 40))))
    return rc;
  return boundary(c, {mask_corr, BOP_WRAP_FACES, BOP_VEL_CORR}, S, us, vs);
}

int substeps_of(int scheme) {
  return scheme == PCFD_SCHEME_EULER ? 1 : scheme == PCFD_SCHEME_RK2 ? 2 : 3;
}

// ProjectionSolver.step (solver.py:686-731) enqueued on c->stream.  With
// `log`, substep s's solve writes its record to the engine's log slot s.
// `want` fills *s from the last solve (a sync); a captured step passes false
// and the engine's budget is set per substep from `budget` (null: polling).
int step_body(pcfd_ctx* c, const double* u0, const double* v0, const StepScalars& S, double* u,
              double* v, double* p, SolveInfo* s, bool want, bool log, const int* budget) {
  int rc;
  double* A[2] = {c->stage_a[0], c->stage_a[1]};
  double* B[2] = {c->stage_b[0], c->stage_b[1]};
  PoissonEngine& E = c->engine;
  auto sub = [&](int k, const double* a, const double* b, const double* ae, const double* be,
                 double ca, double cb, double* ua, double* va, bool last) -> int {
    E.set_log_slot(log ? k : -1);
    E.set_budget(budget ? budget[k] : 0);
    const int r = substep(c, a, b, ae, be, ca, cb, S, ua, va, p, s, last && want, last, k);
    E.set_log_slot(-1);
    E.set_budget(0);
    return r;
  };
  switch (c->scheme) {
    case PCFD_SCHEME_EULER:
      if ((rc = sub(0, u0, v0, nullptr, nullptr, 0, 0, u, v, true))) return rc;
      break;
    case PCFD_SCHEME_RK2:  // solver.py:702-707
      if ((rc = sub(0, u0, v0, nullptr, nullptr, 0, 0, A[0], A[1], false))) return rc;
      if ((rc = sub(1, A[0], A[1], nullptr, nullptr, 0, 0, B[0], B[1], true))) return rc;
      if ((rc = launch_combine(c->stream, u0, B[0], v0, B[1], 0.5, 0.5, c->nx, c->ny, c->ldu,
                               c->ldv, u, v, step_clock(c, PoissonEngine::CK_COMBINE, 48))))
        return rc;
      break;
    default:  // SSP-RK3, solver.py:709-721
      if ((rc = sub(0, u0, v0, nullptr, nullptr, 0, 0, A[0], A[1], false))) return rc;
      if ((rc = sub(1, A[0], A[1], nullptr, nullptr, 0, 0, B[0], B[1], false))) return rc;
      if ((rc = sub(2, u0, v0, B[0], B[1], 0.75, 0.25, A[0], A[1], true))) return rc;
      if ((rc = launch_combine(c->stream, u0, A[0], v0, A[1], 1.0 / 3.0, 2.0 / 3.0, c->nx, c->ny,
                               c->ldu, c->ldv, u, v,
                               step_clock(c, PoissonEngine::CK_COMBINE, 48))))
        return rc;
      break;
  }
  if (c->scheme != PCFD_SCHEME_EULER) {
    if ((rc = boundary(c, {BOP_VEL_CORR, BOP_WRAP_FACES}, S, u, v))) return rc;
  }
  return kOk;
}

// Budget of a captured solve: the last count plus a margin (the warm-started
// counts drift by an iteration or two from step to step), within maxiter.
int solve_budget(const pcfd_ctx* c, int last) {
  const int b = last + std::max(2, last / 8);
  return std::min(b, std::max(1, c->engine.maxiter()));
}

void set_streams(pcfd_ctx* c, hipStream_t s) {
  c->stream = s;
  c->engine.set_stream(s);
}

// Capture step_body into g (c->stream is the replay stream; the capture runs
// on a private non-blocking stream).
int capture_step(pcfd_ctx* c, pcfd_ctx::StepGraph& g) {
  hipStream_t replay = c->stream;
  set_streams(c, c->cap_stream);
  hipGraph_t graph = nullptr;
  int rc = kOk;
  hipError_t e = hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    set_streams(c, replay);
    set_last_error("hipStreamBeginCapture: %s", hipGetErrorString(e));
    return kErrHip;
  }
  const size_t bytes = (size_t)c->nx * c->ny * sizeof(double);
  rc = launch_set_step_scalars(c->cap_stream, c->scal, 0.0, 0.0);
  if (!rc) {
    // the warm-start pressure the step starts from, kept for a replay that
    // runs out of iterations (its step is then redone from the same start)
    e = hipMemcpyAsync(c->guess_backup, c->engine.guess(), bytes, hipMemcpyDeviceToDevice,
                       c->cap_stream);
    if (e == hipSuccess && c->warm_idx)  // and the extrapolation history's state
      e = hipMemcpyAsync(c->warm_idx_backup, c->warm_idx, 12 * sizeof(int),
                         hipMemcpyDeviceToDevice, c->cap_stream);
    if (e != hipSuccess) {
      set_last_error("hipMemcpyAsync in capture: %s", hipGetErrorString(e));
      rc = kErrHip;
    }
  }
  if (!rc)
    rc = step_body(c, g.u0, g.v0, StepScalars{0.0, 0.0, c->scal}, g.u, g.v, g.p, nullptr, false,
                   true, g.budget);
  e = hipStreamEndCapture(c->cap_stream, &graph);
  set_streams(c, replay);
  if (rc) {
    if (graph) hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess || !graph) {
    set_last_error("hipStreamEndCapture: %s", hipGetErrorString(e));
    return kErrHip;
  }
  size_t nroot = 0;
  hipGraphNode_t root = nullptr;
  e = hipGraphGetRootNodes(graph, nullptr, &nroot);
  if (e == hipSuccess && nroot == 1) e = hipGraphGetRootNodes(graph, &root, &nroot);
  hipGraphNodeType type = hipGraphNodeTypeEmpty;
  if (e == hipSuccess && root) e = hipGraphNodeGetType(root, &type);
  if (e != hipSuccess || nroot != 1 || type != hipGraphNodeTypeKernel) {
    hipGraphDestroy(graph);
    set_last_error("captured step: expected one kernel root node (got %zu)", nroot);
    return kErrHip;
  }
  hipKernelNodeParams kp{};
  if ((e = hipGraphKernelNodeGetParams(root, &kp)) != hipSuccess) {
    hipGraphDestroy(graph);
    set_last_error("hipGraphKernelNodeGetParams: %s", hipGetErrorString(e));
    return kErrHip;
  }
  hipGraphExec_t exec = nullptr;
  if ((e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0)) != hipSuccess) {
    hipGraphDestroy(graph);
    set_last_error("hipGraphInstantiate: %s", hipGetErrorString(e));
    return kErrHip;
  }
  g.release();
  g.graph = graph;
  g.exec = exec;
  g.scalars = root;
  g.kp = kp;
  g.partial_gen = c->engine.partial_gen_;
  g.stale = false;
  ++c->graph_captures;
  return kOk;
}

void info_from_log(const pcfd_ctx* c, const SolveLog& L, SolveInfo* s) {
  s->iterations = L.iter;
  s->residual = c->engine.reports_final_residual() ? L.res_final : L.res;
  s->converged = L.converged;
  s->status = kOk;
}

// The step run eagerly (polling each solve) with every solve logged, so the
// next capture knows each substep's iteration count.
int step_eager_logged(pcfd_ctx* c, const double* u0, const double* v0, double dt, double blend,
                      double* u, double* v, double* p, SolveInfo* s) {
  int rc = step_body(c, u0, v0, StepScalars{dt, blend, nullptr}, u, v, p, s, true, true, nullptr);
  if (rc) return rc;
  SolveLog L[3];
  const int ns = substeps_of(c->scheme);
  if ((rc = c->engine.read_log(L, ns))) return rc;
  for (int k = 0; k < ns; ++k) c->pred_sub[k] = L[k].iter;
  return kOk;
}

// ProjectionSolver.step through a captured graph: replayed with this step's
// dt and blend; a solve that ran out of its budget redoes the step eagerly.
int step_graph(pcfd_ctx* c, const double* u0, const double* v0, double dt, double blend,
               double* u, double* v, double* p, SolveInfo* s) {
  const int ns = substeps_of(c->scheme);
  bool known = true;
  for (int k = 0; k < ns; ++k) known &= c->pred_sub[k] > 0;
  if (!known) return step_eager_logged(c, u0, v0, dt, blend, u, v, p, s);
  pcfd_ctx::StepGraph* g = nullptr;
  for (auto& x : c->graphs)
    if (x.u0 == u0 && x.v0 == v0 && x.u == u && x.v == v && x.p == p) g = &x;
  if (g && (g->stale || g->partial_gen != c->engine.partial_gen_)) g->release();
  if (!g) {
    constexpr size_t kMaxGraphs = 4;
    if (c->graphs.size() < kMaxGraphs) {
      c->graphs.emplace_back();
      g = &c->graphs.back();
    } else {  // replace the least recently used
      g = &c->graphs[0];
      for (auto& x : c->graphs)
        if (x.used < g->used) g = &x;
      g->release();
    }
    g->u0 = u0;
    g->v0 = v0;
    g->u = u;
    g->v = v;
    g->p = p;
  }
  if (!g->exec) {
    for (int k = 0; k < ns; ++k) g->budget[k] = solve_budget(c, c->pred_sub[k]);
    if (int rc = capture_step(c, *g)) return rc;
  }
  g->used = ++c->graph_tick;
  void* args[3] = {(void*)&c->scal, (void*)&dt, (void*)&blend};
  hipKernelNodeParams kp = g->kp;
  kp.kernelParams = args;
  kp.extra = nullptr;
  PCFD_HIP_CHECK(hipGraphExecKernelNodeSetParams(g->exec, g->scalars, &kp));
  PCFD_HIP_CHECK(hipGraphLaunch(g->exec, c->stream));
  ++c->graph_launches;
  SolveLog L[3];
  if (int rc = c->engine.read_log(L, ns)) return rc;
  bool short_budget = false;
  for (int k = 0; k < ns; ++k) short_budget |= !L[k].done;
  if (short_budget) {
    // Redo the step eagerly from the same warm start; the next replay is
    // captured again with the new counts.
    ++c->graph_fallbacks;
    g->stale = true;
    PCFD_HIP_CHECK(hipMemcpyAsync(c->engine.guess(), c->guess_backup,
                                  (size_t)c->nx * c->ny * sizeof(double),
                                  hipMemcpyDeviceToDevice, c->stream));
    // the history's slots the replay wrote are ones the redo's guesses do
    // not read (kWarmSlots leaves one spare); its state goes back
    if (c->warm_idx)
      PCFD_HIP_CHECK(hipMemcpyAsync(c->warm_idx, c->warm_idx_backup, 12 * sizeof(int),
                                    hipMemcpyDeviceToDevice, c->stream));
    return step_eager_logged(c, u0, v0, dt, blend, u, v, p, s);
  }
  for (int k = 0; k < ns; ++k) {
    c->pred_sub[k] = L[k].iter;
    // recapture with a tighter budget once the counts have dropped well
    // below it (the extra iterations are no-op launches)
    if (g->budget[k] > solve_budget(c, L[k].iter) + std::max(2, L[k].iter / 4)) g->stale = true;
  }
  info_from_log(c, L[ns - 1], s);
  return kOk;
}

// Per-cell flag byte of the fine Poisson operator (pressure.py:209-251):
// solvable bit plus one bit per face that carries flux.  Returns the number
// of solvable cells.
int poisson_flags(const pcfd_case_desc& d, const std::vector<uint8_t>& solid, uint8_t* flags) {
  const int nx = d.nx, ny = d.ny;
  const bool px = d.bc_kind[0] == BC_PERIODIC, py = d.bc_kind[2] == BC_PERIODIC;
  bool dir[4];
  for (int w = 0; w < 4; ++w) dir[w] = d.bc_kind[w] == BC_PRESSURE_OUTLET;
  auto sol = [&](int i, int j) -> bool { return solid[(size_t)i * ny + j] != 0; };
  int n_fluid = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      flags[(size_t)i * ny + j] = 0;
      if (sol(i, j)) continue;
      uint8_t f = 0;
      auto nbf = [&](int ii, int jj) -> bool {
        if (px) ii = (ii + nx) % nx;
        if (py) jj = (jj + ny) % ny;
        if (ii < 0 || ii >= nx || jj < 0 || jj >= ny) return false;
        return !sol(ii, jj);
      };
      if (nbf(i - 1, j)) f |= F_L;
      if (nbf(i + 1, j)) f |= F_R;
      if (nbf(i, j - 1)) f |= F_B;
      if (nbf(i, j + 1)) f |= F_T;
      const bool on_dir = (dir[0] && i == 0) || (dir[1] && i == nx - 1) ||
                          (dir[2] && j == 0) || (dir[3] && j == ny - 1);
      if (f != 0 || on_dir) {
        f |= F_SOLV;
        ++n_fluid;
      }
      flags[(size_t)i * ny + j] = f;
    }
  return n_fluid;
}

int project_impl(pcfd_ctx* c, double* u, double* v, double dt, double* p, SolveInfo* info) {
  int rc;
  const StepScalars S{dt, 0.0, nullptr};
  if ((rc = c->engine.prepare_from_velocity(u, v, c->ldu, c->ldv, c->dx, c->dy, S))) return rc;
  if ((rc = c->engine.solve(info, true))) return rc;
  if ((rc = c->engine.scatter_pressure(p, c->ldp))) return rc;
  return launch_correct_ghosts(c->stream, p, c->ldp, c->bc, c->ldu, c->ldv, c->dx, c->dy, c->g,
                               c->upd, S, u, v);
}

}  // namespace

extern "C" {

int pcfd_last_error(char* buf, size_t len) {
  if (buf && len) {
    std::strncpy(buf, g_last_error, len - 1);
    buf[len - 1] = '\0';
  }
  return kOk;
}

int pcfd_device_count(int32_t* count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  if (count) *count = n;
  return n > 0 ? kOk : kErrNoDevice;
}

int pcfd_poisson_flags(const pcfd_case_desc* d, uint8_t* flags_out, int32_t* n_solvable) {
  if (!d || !flags_out || d->nx < 1 || d->ny < 1) {
    set_last_error("pcfd_poisson_flags: bad argument");
    return kErrBadArg;
  }
  std::vector<uint8_t> solid((size_t)d->nx * d->ny, 0);
  if (d->solid)
    for (size_t k = 0; k < solid.size(); ++k) solid[k] = d->solid[k] ? 1 : 0;
  const int n = poisson_flags(*d, solid, flags_out);
  if (n_solvable) *n_solvable = n;
  return kOk;
}

int pcfd_ctx_create(const pcfd_case_desc* d, void* hip_stream, pcfd_ctx** out) {
  if (!d || !out) {
    set_last_error("pcfd_ctx_create: null argument");
    return kErrBadArg;
  }
  *out = nullptr;
  const int nx = d->nx, ny = d->ny;
  if (nx < 4 || ny < 4) {
    set_last_error("grid must be at least 4x4, got %dx%d", nx, ny);
    return kErrBadArg;
  }
  int ndev = 0;
  const hipError_t derr = hipGetDeviceCount(&ndev);
  if (derr != hipSuccess || ndev == 0) {
    set_last_error("no HIP device available (hipGetDeviceCount: %s, %d devices)",
                   hipGetErrorString(derr), ndev);
    return kErrNoDevice;
  }
  std::unique_ptr<pcfd_ctx> c(new (std::nothrow) pcfd_ctx());
  if (!c) return kErrBadArg;
  c->nx = nx;
  c->ny = ny;
  c->ldu = ny + 2;
  c->ldv = ny + 3;
  c->ldp = ny + 2;
  c->dx = d->dx;
  c->dy = d->dy;
  c->nu = d->nu;
  c->fx = d->body_fx;
  c->fy = d->body_fy;
  c->scheme = d->time_scheme;
  for (int w = 0; w < 4; ++w) c->kind[w] = d->bc_kind[w];
  // NULL selects the default stream, which orders this context's work with
  // everything else the caller enqueues there (torch's fills and copies).
  c->stream = (hipStream_t)hip_stream;
  const bool px = d->bc_kind[0] == BC_PERIODIC;
  const bool py = d->bc_kind[2] == BC_PERIODIC;
  if (px != (d->bc_kind[1] == BC_PERIODIC) || py != (d->bc_kind[3] == BC_PERIODIC)) {
    set_last_error("periodicity must be set on both walls of an axis");
    return kErrBadArg;
  }
  const bool stretched = d->dx_cells || d->dy_cells;
  if (stretched) {
    if (!d->dx_cells || !d->dy_cells) {
      set_last_error("a stretched mesh needs both dx_cells and dy_cells");
      return kErrBadArg;
    }
    if (!(d->area_mean > 0.0)) {
      set_last_error("a stretched mesh needs area_mean > 0, got %g", d->area_mean);
      return kErrBadArg;
    }
    for (int i = 0; i < nx; ++i)
      if (!(d->dx_cells[i] > 0.0)) {
        set_last_error("dx_cells[%d] = %g is not positive", i, d->dx_cells[i]);
        return kErrBadArg;
      }
    for (int j = 0; j < ny; ++j)
      if (!(d->dy_cells[j] > 0.0)) {
        set_last_error("dy_cells[%d] = %g is not positive", j, d->dy_cells[j]);
        return kErrBadArg;
      }
    // solver.py:209-230: the first and last cells of a periodic axis must match
    for (int i = 1; i < nx && px; ++i)
      if (d->dx_cells[i] != d->dx_cells[0]) {
        set_last_error("the x axis is periodic and stretched; use a uniform periodic axis");
        return kErrBadArg;
      }
    for (int j = 1; j < ny && py; ++j)
      if (d->dy_cells[j] != d->dy_cells[0]) {
        set_last_error("the y axis is periodic and stretched; use a uniform periodic axis");
        return kErrBadArg;
      }
  }
  const int PO = BC_PRESSURE_OUTLET;
  // solvable face ranges (solver.py:251-263)
  UpdRange R;
  R.u_lo = (px || d->bc_kind[0] == PO) ? 1 : 2;
  R.u_hi = (d->bc_kind[1] == PO && !px) ? nx + 2 : nx + 1;
  R.v_lo = (py || d->bc_kind[2] == PO) ? 1 : 2;
  R.v_hi = (d->bc_kind[3] == PO && !py) ? ny + 2 : ny + 1;
  c->upd = R;

  // immersed-boundary face lists (solver.py:62-79, 276-296)
  std::vector<uint8_t> solid((size_t)nx * ny, 0);
  std::vector<int8_t> body((size_t)nx * ny, -1);
  bool has_obstacle = false;
  if (d->solid) {
    for (size_t k = 0; k < solid.size(); ++k) {
      solid[k] = d->solid[k] ? 1 : 0;
      has_obstacle |= solid[k] != 0;
      if (d->body_id) body[k] = d->body_id[k];
    }
  }
  // a soft body: the solid set shapes the Poisson operator only
  c->soft_ib = d->soft_ib != 0;
  const bool binary_faces = has_obstacle && !c->soft_ib;
  auto sol = [&](int i, int j) -> bool { return solid[(size_t)i * ny + j] != 0; };
  auto bod = [&](int i, int j) -> int8_t { return body[(size_t)i * ny + j]; };
  std::vector<int> uf, vf;
  std::vector<int8_t> ufb, vfb;
  if (binary_faces) {
    for (int m = R.u_lo; m < R.u_hi; ++m)
      for (int j = 0; j < ny; ++j) {
        int a = m - 2, b = m - 1;
        if (px) {
          a = ((a % nx) + nx) % nx;
          b = ((b % nx) + nx) % nx;
        }
        const bool sa = (a >= 0 && a < nx) && sol(a, j);
        const bool sb = (b >= 0 && b < nx) && sol(b, j);
        if (sa || sb) {
          uf.push_back(m * c->ldu + (j + 1));
          ufb.push_back(sa ? bod(a, j) : bod(b, j));
        }
      }
    for (int i = 0; i < nx; ++i)
      for (int n = R.v_lo; n < R.v_hi; ++n) {
        int a = n - 2, b = n - 1;
        if (py) {
          a = ((a % ny) + ny) % ny;
          b = ((b % ny) + ny) % ny;
        }
        const bool sa = (a >= 0 && a < ny) && sol(i, a);
        const bool sb = (b >= 0 && b < ny) && sol(i, b);
        if (sa || sb) {
          vf.push_back((i + 1) * c->ldv + n);
          vfb.push_back(sa ? bod(i, a) : bod(i, b));
        }
      }
  }
  int rc;
  if ((rc = upload(&c->uf, uf.data(), uf.size()))) return rc;
  if ((rc = upload(&c->vf, vf.data(), vf.size()))) return rc;
  if ((rc = upload(&c->uf_body, ufb.data(), ufb.size()))) return rc;
  if ((rc = upload(&c->vf_body, vfb.data(), vfb.size()))) return rc;
  if ((rc = upload(&c->solid, solid.data(), solid.size()))) return rc;
  c->n_bodies = (d->n_bodies > 1 && d->body_id) ? d->n_bodies : 0;
  std::vector<double> zeros(2 + 2 * (size_t)std::max(c->n_bodies, 0), 0.0);
  if ((rc = upload(&c->force, zeros.data(), zeros.size()))) return rc;

  // Stretched-mesh arrays (mesh.py:160-211, pressure.py:193-207,
  // solver.py:197-207, 271-274), each computed with the reference's
  // elementwise arithmetic.
  if (stretched) {
    const double* dxc = d->dx_cells;
    const double* dyc = d->dy_cells;
    std::vector<double> hxe(nx + 2), hye(ny + 2), hxu(nx + 1), hyv(ny + 1), wx(nx + 1),
        wy(ny + 1), hxv(nx + 1), hyw(ny + 1);
    hxe[0] = dxc[0];
    hxe[nx + 1] = dxc[nx - 1];
    for (int i = 0; i < nx; ++i) hxe[i + 1] = dxc[i];
    hye[0] = dyc[0];
    hye[ny + 1] = dyc[ny - 1];
    for (int j = 0; j < ny; ++j) hye[j + 1] = dyc[j];
    for (int m = 0; m <= nx; ++m) {
      hxu[m] = 0.5 * (hxe[m] + hxe[m + 1]);
      wx[m] = 0.5 * hxe[m] / hxu[m];
      hxv[m] = hxu[m] * d->area_mean;
    }
    for (int n = 0; n <= ny; ++n) {
      hyv[n] = 0.5 * (hye[n] + hye[n + 1]);
      wy[n] = 0.5 * hye[n] / hyv[n];
      hyw[n] = hyv[n] * d->area_mean;
    }
    // face control volumes: u face m, row j -> hxu[m-1] * hy[j-1]; v face
    // (i, n) -> hx[i-1] * hyv[n-1]
    std::vector<double> ufa(uf.size()), vfa(vf.size());
    for (size_t k = 0; k < uf.size(); ++k) {
      const int m = uf[k] / c->ldu, j = uf[k] % c->ldu;
      ufa[k] = hxu[m - 1] * dyc[j - 1];
    }
    for (size_t k = 0; k < vf.size(); ++k) {
      const int i = vf[k] / c->ldv, n = vf[k] % c->ldv;
      vfa[k] = dxc[i - 1] * hyv[n - 1];
    }
    std::vector<double> all;
    std::vector<size_t> at;
    for (const std::vector<double>* a :
         {&hxe, &hye, &hxu, &hyv, &wx, &wy, &hxv, &hyw, &ufa, &vfa}) {
      at.push_back(all.size());
      all.insert(all.end(), a->begin(), a->end());
    }
    at.push_back(all.size());
    all.insert(all.end(), dxc, dxc + nx);
    at.push_back(all.size());
    all.insert(all.end(), dyc, dyc + ny);
    if ((rc = upload(&c->metric, all.data(), all.size()))) return rc;
    const double* base = c->metric;
    Metrics& g = c->g;
    g.hxe = base + at[0];
    g.hye = base + at[1];
    g.hxu = base + at[2];
    g.hyv = base + at[3];
    g.wx = base + at[4];
    g.wy = base + at[5];
    g.hx = base + at[10];
    g.hy = base + at[11];
    c->bc.uf_area = uf.empty() ? nullptr : base + at[8];
    c->bc.vf_area = vf.empty() ? nullptr : base + at[9];
    c->bc.hx = g.hx;
    c->bc.hy = g.hy;
    c->poisson_hxv = base + at[6];
    c->poisson_hyw = base + at[7];
  }

  // inlet lines (boundary.py:205-216)
  for (int w = 0; w < 4; ++w) {
    if (d->bc_kind[w] == BC_INLET) {
      if (!d->inlet_line[w]) {
        set_last_error("INLET wall %d without an inlet line", w);
        return kErrBadArg;
      }
      const size_t len = (w < 2) ? (size_t)ny + 2 : (size_t)nx + 2;
      if ((rc = upload(&c->inlet[w], d->inlet_line[w], len))) return rc;
    }
    if (d->bc_kind[w] == BC_OUTLET) c->any_outlet = 1;
  }

  BCParams& B = c->bc;
  B.nx = nx;
  B.ny = ny;
  B.ldu = c->ldu;
  B.ldv = c->ldv;
  for (int w = 0; w < 4; ++w) {
    B.kind[w] = d->bc_kind[w];
    B.vel[w] = d->bc_velocity[w];
    B.pref[w] = d->bc_p_ref[w];
    B.inlet[w] = c->inlet[w];
  }
  B.dx = d->dx;
  B.dy = d->dy;
  B.periodic_x = px;
  B.periodic_y = py;
  B.has_obstacle = binary_faces;
  B.n_uf = (int)uf.size();
  B.n_vf = (int)vf.size();
  B.uf = c->uf;
  B.vf = c->vf;
  B.uf_body = c->uf_body;
  B.vf_body = c->vf_body;
  B.n_bodies = c->n_bodies;
  B.area = d->dx * d->dy;  // mesh.py:372 (_dx * _dy)
  B.force = c->force;
  {
    auto inside = [&](int a, int b, int na, int nb) {  // a in [0, na), b in [0, nb)
      return a >= kMaskApart && a < na - kMaskApart && b >= kMaskApart && b < nb - kMaskApart;
    };
    bool apart = binary_faces;
    for (int f : uf) apart = apart && inside(f / c->ldu, f % c->ldu, nx + 3, ny + 2);
    for (int f : vf) apart = apart && inside(f / c->ldv, f % c->ldv, nx + 2, ny + 3);
    B.mask_apart = apart ? 1 : 0;
  }

  MomParams& M = c->mom;
  M.nx = nx;
  M.ny = ny;
  M.ldu = c->ldu;
  M.ldv = c->ldv;
  M.dx = d->dx;
  M.dy = d->dy;
  M.dx2 = d->dx * d->dx;
  M.dy2 = d->dy * d->dy;
  M.pow2 = !stretched && is_pow2(d->dx) && is_pow2(d->dy);
  M.idx = 1.0 / d->dx;
  M.idy = 1.0 / d->dy;
  M.idx2 = 1.0 / M.dx2;
  M.idy2 = 1.0 / M.dy2;
  M.nu = d->nu;
  M.fx = d->body_fx;
  M.fy = d->body_fy;
  M.blend = 0.0;
  M.nuc = nullptr;
  M.g = c->g;
  if (d->use_les) {
    if (!(d->les_coeff >= 0.0) || (stretched && !(d->les_cs >= 0.0))) {
      set_last_error("les_coeff and les_cs must be >= 0, got %g, %g", d->les_coeff, d->les_cs);
      return kErrBadArg;
    }
    c->use_les = true;
    c->les = LesParams{nx, ny, c->ldu, c->ldv, d->dx, d->dy, d->nu, d->les_coeff, d->les_cs, c->g};
    PCFD_HIP_CHECK(hipMalloc(&c->nuc, (size_t)(nx + 2) * (ny + 2) * sizeof(double)));
    PCFD_HIP_CHECK(hipMemsetAsync(c->nuc, 0, (size_t)(nx + 2) * (ny + 2) * sizeof(double),
                                  c->stream));
    M.nuc = c->nuc;
  }

  // Poisson operator flags (pressure.py:209-251)
  bool dir[4];
  for (int w = 0; w < 4; ++w) dir[w] = d->bc_kind[w] == PO;
  std::vector<uint8_t> flags((size_t)nx * ny, 0);
  const int n_fluid = poisson_flags(*d, solid, flags.data());
  if ((rc = upload(&c->flags, flags.data(), flags.size()))) return rc;
  const bool singular = !(dir[0] || dir[1] || dir[2] || dir[3]);

  FineOp op{};
  op.flags = c->flags;
  op.nx = nx;
  op.ny = ny;
  op.cx = 1.0 / (d->dx * d->dx);
  op.cy = 1.0 / (d->dy * d->dy);
  op.px = px;
  op.py = py;
  for (int w = 0; w < 4; ++w) {
    op.dir[w] = dir[w];
    op.pref[w] = d->bc_p_ref[w];
  }
  if (stretched) {
    op.hxc = c->g.hx;
    op.hyc = c->g.hy;
    op.hxv = c->poisson_hxv;
    op.hyv = c->poisson_hyw;
    op.vbar = d->area_mean;
  }
  if ((rc = c->engine.init(op, n_fluid, singular, c->stream))) return rc;

  const size_t usz = (size_t)(nx + 3) * c->ldu, vsz = (size_t)(nx + 2) * c->ldv;
  PCFD_HIP_CHECK(hipMalloc(&c->stage_a[0], usz * sizeof(double)));
  PCFD_HIP_CHECK(hipMalloc(&c->stage_a[1], vsz * sizeof(double)));
  PCFD_HIP_CHECK(hipMalloc(&c->stage_b[0], usz * sizeof(double)));
  PCFD_HIP_CHECK(hipMalloc(&c->stage_b[1], vsz * sizeof(double)));
  PCFD_HIP_CHECK(hipMalloc(&c->diag_out, 16 * sizeof(double)));
  PCFD_HIP_CHECK(hipMalloc(&c->guess_backup, (size_t)nx * ny * sizeof(double)));
  PCFD_HIP_CHECK(hipHostMalloc(&c->host_out, 64 * sizeof(double), hipHostMallocDefault));
  if (c->soft_ib) PCFD_HIP_CHECK(hipMalloc(&c->soft_work, (size_t)nx * ny * sizeof(double)));
  if ((rc = scratch_for_current_device(&c->scratch))) return rc;
  PCFD_HIP_CHECK(hipStreamSynchronize(c->stream));
  *out = c.release();
  return kOk;
}

int pcfd_ctx_destroy(pcfd_ctx* c) {
  if (!c) return kOk;
  hipStreamSynchronize(c->stream);
  const bool own = c->own_stream;
  hipStream_t s = c->stream;
  delete c;
  if (own) hipStreamDestroy(s);
  return kOk;
}

int pcfd_synchronize(pcfd_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  PCFD_HIP_CHECK(hipStreamSynchronize(c->stream));
  return kOk;
}

int pcfd_set_pressure_solver(pcfd_ctx* c, const pcfd_solver_desc* sd,
                             const pcfd_mg_level_desc* levels, int32_t n_levels) {
  if (int rc = check_ctx(c)) return rc;
  if (!sd) {
    set_last_error("null solver description");
    return kErrBadArg;
  }
  PoissonConfig cfg;
  cfg.kind = sd->kind;
  cfg.tol = sd->tol;
  cfg.maxiter = sd->maxiter;
  cfg.omega = sd->omega;
  cfg.sweeps = sd->sweeps;
  cfg.precision = sd->precision;
  cfg.fused = sd->fused_levels;
  cfg.tile_tx = sd->tile_tx;
  cfg.tile_ty = sd->tile_ty;
  cfg.chain = sd->coarse_chain;
  cfg.stream = sd->stream_rows;
  cfg.row_reg = sd->row_passes;
  if (cfg.stream > 0 && cfg.stream % 4) {
    set_last_error("stream_rows must be a multiple of 4, got %d", cfg.stream);
    return kErrBadArg;
  }
  if (cfg.tile_tx < 0 || cfg.tile_ty < 0) {
    set_last_error("tile sizes must be >= 0, got %dx%d", cfg.tile_tx, cfg.tile_ty);
    return kErrBadArg;
  }
  if (cfg.sweeps < 1) {
    set_last_error("mg sweeps must be at least 1, got %d", cfg.sweeps);
    return kErrBadArg;
  }
  std::vector<MGLevelHost> lv;
  for (int l = 0; l < n_levels; ++l) {
    const pcfd_mg_level_desc& L = levels[l];
    MGLevelHost h;
    h.nx = L.nx;
    h.ny = L.ny;
    h.fx = L.fx;
    h.fy = L.fy;
    h.weight = L.weight;
    const size_t n = (size_t)L.nx * L.ny;
    if (l == 0 && (L.nx != c->nx || L.ny != c->ny)) {
      set_last_error("level 0 of the hierarchy is %dx%d, grid is %dx%d", L.nx, L.ny, c->nx,
                     c->ny);
      return kErrBadArg;
    }
    if (L.active) h.active.assign(L.active, L.active + n);
    else h.active.assign(n, 1);
    if (l > 0) {
      if ((!L.cls && !L.cls32) || !L.table || !L.class_dinv || !L.class_active ||
          L.n_offsets <= 0 || L.n_classes <= 0 || !L.di || !L.dj) {
        set_last_error("coarse level %d lacks its stencil dictionary", l);
        return kErrBadArg;
      }
      if (!L.cls32 && L.n_classes > 65535) {
        set_last_error("level %d: %d classes need the 32-bit class index", l, L.n_classes);
        return kErrBadArg;
      }
      h.di.assign(L.di, L.di + L.n_offsets);
      h.dj.assign(L.dj, L.dj + L.n_offsets);
      bool in_range = true;
      if (L.cls32) {
        h.cls32.assign(L.cls32, L.cls32 + n);
        for (size_t k = 0; k < n; ++k) in_range &= h.cls32[k] < (uint32_t)L.n_classes;
      } else {
        h.cls.assign(L.cls, L.cls + n);
        for (size_t k = 0; k < n; ++k) in_range &= h.cls[k] < (uint32_t)L.n_classes;
      }
      if (!in_range) {
        set_last_error("level %d: class index out of range", l);
        return kErrBadArg;
      }
      h.table.assign(L.table, L.table + (size_t)L.n_classes * L.n_offsets);
      h.cdinv.assign(L.class_dinv, L.class_dinv + L.n_classes);
      h.cactive.assign(L.class_active, L.class_active + L.n_classes);
    }
    if (l == n_levels - 1) {
      if (!L.coarse_inv || L.coarse_n != (int)n) {
        set_last_error("coarsest level needs its %zux%zu pseudo-inverse", n, n);
        return kErrBadArg;
      }
      if (n > 8192) {
        set_last_error("coarsest level has %zu unknowns (max 8192)", n);
        return kErrUnsupported;
      }
      h.coarse_inv.assign(L.coarse_inv, L.coarse_inv + n * n);
    }
    lv.push_back(std::move(h));
  }
  // captured steps hold the old engine's buffers and iteration counts
  for (auto& g : c->graphs) g.release();
  c->graphs.clear();
  for (int& k : c->pred_sub) k = 0;
  int rc = c->engine.configure(cfg, lv);
  if (rc) return rc;
  c->solver_set = true;
  return kOk;
}

int pcfd_momentum_rhs(pcfd_ctx* c, const double* u, const double* v, double blend,
                      double* rhs_u, double* rhs_v) {
  if (int rc = check_ctx(c)) return rc;
  MomParams P = c->mom;
  P.blend = blend;
  if (c->use_les) {
    if (int rc = launch_eddy_viscosity(c->stream, c->les, u, nullptr, v, nullptr, 1.0, 0.0, c->nuc,
                                       nullptr))
      return rc;
  }
  return launch_momentum_rhs(c->stream, u, v, P, rhs_u, rhs_v);
}

int pcfd_eddy_viscosity(pcfd_ctx* c, const double* u, const double* v, double* nu_c,
                        double* nu_corner) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->use_les) {
    set_last_error("pcfd_eddy_viscosity: the context was created without use_les");
    return kErrBadArg;
  }
  if (!nu_c) {
    set_last_error("pcfd_eddy_viscosity: null nu_c");
    return kErrBadArg;
  }
  return launch_eddy_viscosity(c->stream, c->les, u, nullptr, v, nullptr, 1.0, 0.0, nu_c,
                               nu_corner);
}

int pcfd_momentum_rhs_variable(pcfd_ctx* c, const double* u, const double* v, double blend,
                               const double* nu_c, double* rhs_u, double* rhs_v) {
  if (int rc = check_ctx(c)) return rc;
  if (!nu_c) {
    set_last_error("pcfd_momentum_rhs_variable: null nu_c");
    return kErrBadArg;
  }
  MomParams P = c->mom;
  P.blend = blend;
  P.nuc = nu_c;
  return launch_momentum_rhs(c->stream, u, v, P, rhs_u, rhs_v);
}

int pcfd_divergence(pcfd_ctx* c, const double* u, const double* v, double* out) {
  if (int rc = check_ctx(c)) return rc;
  return launch_divergence(c->stream, u, v, c->nx, c->ny, c->ldu, c->ldv, c->dx, c->dy, c->g,
                           out);
}

int pcfd_apply_velocity_bc(pcfd_ctx* c, double* u, double* v, int32_t predictor) {
  if (int rc = check_ctx(c)) return rc;
  return boundary(c, {predictor ? BOP_VEL_PRED : BOP_VEL_CORR}, 1.0, u, v);
}

int pcfd_mass_balance(pcfd_ctx* c, double* u, double* v) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->any_outlet) return kOk;
  return boundary(c, {BOP_MASS}, 1.0, u, v);
}

int pcfd_mask_obstacle(pcfd_ctx* c, double* u, double* v, double dt, int32_t mode) {
  if (int rc = check_ctx(c)) return rc;
  const int op = mode == PCFD_MASK_RESET      ? BOP_MASK_RESET
                 : mode == PCFD_MASK_ACCUMULATE ? BOP_MASK_ACC
                                                : BOP_MASK_ZERO;
  if (op != BOP_MASK_ZERO && !(dt > 0.0)) {
    set_last_error("mask with force accumulation needs dt > 0");
    return kErrBadArg;
  }
  return boundary(c, {op}, dt, u, v);
}

int pcfd_wrap_faces(pcfd_ctx* c, double* u, double* v) {
  if (int rc = check_ctx(c)) return rc;
  return boundary(c, {BOP_WRAP_FACES}, 1.0, u, v);
}

int pcfd_apply_pressure_bc(pcfd_ctx* c, double* p) {
  if (int rc = check_ctx(c)) return rc;
  return launch_pressure_ghosts(c->stream, c->bc, p, c->ldp);
}

int pcfd_pressure_solve(pcfd_ctx* c, const double* rhs, const double* p0, double* p_out,
                        pcfd_solve_info* info) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_solver(c)) return rc;
  const size_t bytes = (size_t)c->nx * c->ny * sizeof(double);
  PoissonEngine& E = c->engine;
  PCFD_HIP_CHECK(hipMemcpyAsync(c->guess_backup, E.guess(), bytes, hipMemcpyDeviceToDevice,
                                c->stream));
  int rc = p0 ? E.set_guess(p0) : E.reset_guess();
  if (rc) return rc;
  SolveInfo s{};
  if ((rc = E.prepare_from_rhs(rhs))) return rc;
  if ((rc = E.solve(&s, true))) return rc;
  if ((rc = E.write_interior(p_out))) return rc;
  PCFD_HIP_CHECK(hipMemcpyAsync(E.guess(), c->guess_backup, bytes, hipMemcpyDeviceToDevice,
                                c->stream));
  PCFD_HIP_CHECK(hipStreamSynchronize(c->stream));
  fill_info(info, s);
  return kOk;
}

int pcfd_project(pcfd_ctx* c, double* u, double* v, double dt, double* p, pcfd_solve_info* info) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_solver(c)) return rc;
  SolveInfo s{};
  int rc = project_impl(c, u, v, dt, p, &s);
  if (rc) return rc;
  PCFD_HIP_CHECK(hipStreamSynchronize(c->stream));
  fill_info(info, s);
  return kOk;
}

int pcfd_initialize(pcfd_ctx* c, double* u, double* v, double* p, pcfd_solve_info* info) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_solver(c)) return rc;
  int rc;
  if (c->any_outlet) {
    rc = boundary(c, {BOP_VEL_PRED, BOP_MASS, BOP_MASK_ZERO, BOP_WRAP_FACES}, 1.0, u, v);
  } else {
    rc = boundary(c, {BOP_VEL_PRED, BOP_MASK_ZERO, BOP_WRAP_FACES}, 1.0, u, v);
  }
  if (rc) return rc;
  SolveInfo s{};
  if ((rc = warm_reset(c))) return rc;
  if ((rc = project_impl(c, u, v, 1.0, p, &s))) return rc;
  if ((rc = boundary(c, {BOP_MASK_ZERO, BOP_WRAP_FACES, BOP_VEL_CORR}, 1.0, u, v))) return rc;
  PCFD_HIP_CHECK(hipStreamSynchronize(c->stream));
  fill_info(info, s);
  return kOk;
}

int pcfd_step(pcfd_ctx* c, const double* u0, const double* v0, double dt, double blend,
              double* u, double* v, double* p, pcfd_solve_info* info) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_solver(c)) return rc;
  c->mom.blend = blend;
  SolveInfo s{};
  const bool want = info != nullptr;
  // A captured step needs distinct input and output buffers (the replay
  // reads u0 while it writes u) and no live kernel clock.
  const bool graph = c->use_graphs && !c->engine.clock_on() && u0 != u && v0 != v;
  int rc = graph ? step_graph(c, u0, v0, dt, blend, u, v, p, &s)
                 : step_body(c, u0, v0, StepScalars{dt, blend, nullptr}, u, v, p, &s, want,
                             false, nullptr);
  if (rc) return rc;
  if (want) fill_info(info, s);
  return kOk;
}

}  // extern "C"

// What the differentiable step (adjoint.hip) needs of a context.
struct SoftView {
  hipStream_t stream;
  int nx, ny, ldu, ldv, ldp;
  double dx, dy, nu;
  UpdRange upd;
  const BCParams* bc;
  const MomParams* mom;
  const Metrics* g;
  PoissonEngine* engine;
  double* force;
  double* scratch_a[2];
  double* scratch_b[2];
  double* work;
  const uint8_t* flags;
  double* guess_backup;
  int kind[4];
  int scheme;
  const uint8_t* solid;
};

int pcfd_soft_view(pcfd_ctx* c, SoftView* out) {
  if (!c->soft_ib) {
    set_last_error("the context was not created with soft_ib = 1");
    return kErrBadArg;
  }
  if (!c->solver_set) {
    set_last_error("no pressure solver configured (pcfd_set_pressure_solver)");
    return kErrBadArg;
  }
  for (int w = 0; w < 4; ++w)
    if (c->kind[w] == PCFD_BC_PRESSURE_OUTLET && c->bc.pref[w] != 0.0) {
      set_last_error("the differentiable step needs p_ref = 0 on pressure outlets");
      return kErrUnsupported;
    }
  SoftView& V = *out;
  V.stream = c->stream;
  V.nx = c->nx;
  V.ny = c->ny;
  V.ldu = c->ldu;
  V.ldv = c->ldv;
  V.ldp = c->ldp;
  V.dx = c->dx;
  V.dy = c->dy;
  V.nu = c->nu;
  V.upd = c->upd;
  V.bc = &c->bc;
  V.mom = &c->mom;
  V.g = &c->g;
  V.engine = &c->engine;
  V.force = c->force;
  V.scratch_a[0] = c->stage_a[0];
  V.scratch_a[1] = c->stage_a[1];
  V.scratch_b[0] = c->stage_b[0];
  V.scratch_b[1] = c->stage_b[1];
  V.work = c->soft_work;
  V.flags = c->flags;
  V.guess_backup = c->guess_backup;
  for (int w = 0; w < 4; ++w) V.kind[w] = c->kind[w];
  V.scheme = c->scheme;
  V.solid = c->solid;
  return kOk;
}

extern "C" {

int pcfd_set_step_graphs(pcfd_ctx* c, int32_t enable) {
  if (int rc = check_ctx(c)) return rc;
  if (enable && !c->cap_stream) {
    PCFD_HIP_CHECK(hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
    PCFD_HIP_CHECK(hipMalloc(&c->scal, 4 * sizeof(double)));
  }
  if (!enable) {
    for (auto& g : c->graphs) g.release();
    c->graphs.clear();
  }
  c->use_graphs = enable != 0;
  return kOk;
}

int pcfd_step_graph_stats(pcfd_ctx* c, int32_t* out3) {
  if (int rc = check_ctx(c)) return rc;
  if (!out3) {
    set_last_error("pcfd_step_graph_stats: null output");
    return kErrBadArg;
  }
  out3[0] = c->graph_launches;
  out3[1] = c->graph_captures;
  out3[2] = c->graph_fallbacks;
  return kOk;
}

int pcfd_reset_guess(pcfd_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = warm_reset(c)) return rc;
  return c->engine.reset_guess();
}

int pcfd_set_warm_start(pcfd_ctx* c, int32_t order) {
  if (int rc = check_ctx(c)) return rc;
  if (order < 0 || order > kWarmMaxOrder) {
    set_last_error("warm-start extrapolation order must be 0-%d, got %d", kWarmMaxOrder, order);
    return kErrBadArg;
  }
  if (order > 0 && !c->warm_slots) {
    const size_t n = (size_t)c->nx * c->ny;
    PCFD_HIP_CHECK(hipMalloc(&c->warm_slots, 3 * kWarmSlots * n * sizeof(double)));
    PCFD_HIP_CHECK(hipMalloc(&c->warm_idx, 12 * sizeof(int)));
    PCFD_HIP_CHECK(hipMalloc(&c->warm_idx_backup, 12 * sizeof(int)));
  }
  c->warm_order = order;
  for (auto& g : c->graphs) g.stale = true;  // the captured steps change
  return warm_reset(c);
}

int pcfd_pressure_guess(pcfd_ctx* c, double* buf, int32_t store) {
  if (int rc = check_ctx(c)) return rc;
  if (!buf) {
    set_last_error("pcfd_pressure_guess: null buffer");
    return kErrBadArg;
  }
  const size_t bytes = (size_t)c->nx * c->ny * sizeof(double);
  if (store) {
    if (int rc = warm_reset(c)) return rc;
    PCFD_HIP_CHECK(
        hipMemcpyAsync(c->engine.guess(), buf, bytes, hipMemcpyDeviceToDevice, c->stream));
  } else {
    PCFD_HIP_CHECK(
        hipMemcpyAsync(buf, c->engine.guess(), bytes, hipMemcpyDeviceToDevice, c->stream));
  }
  PCFD_HIP_CHECK(hipStreamSynchronize(c->stream));
  return kOk;
}

int pcfd_diagnostics(pcfd_ctx* c, const double* u, const double* v, const double* p,
                     const double* u_prev, const double* v_prev, double* out_host) {
  if (int rc = check_ctx(c)) return rc;
  // u, v, p, u_prev, v_prev, the solid byte
  int rc = launch_diagnostics(c->stream, u, v, p, u_prev, v_prev, c->solid, c->nx, c->ny, c->ldu,
                              c->ldv, c->ldp, c->dx, c->dy, c->g, c->engine.partial_,
                              c->engine.counter_, c->diag_out,
                              step_clock(c, PoissonEngine::CK_DIAG, 41));
  if (rc) return rc;
  PCFD_HIP_CHECK(
      hipMemcpyAsync(c->host_out, c->diag_out, 7 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PCFD_HIP_CHECK(hipMemcpyAsync(c->host_out + 7, c->force, 2 * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream));
  PCFD_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (out_host) std::memcpy(out_host, c->host_out, 9 * sizeof(double));
  return kOk;
}

int pcfd_diagnostics_noalloc(pcfd_ctx* c, const double* u, const double* v, const double* p,
                             const double* u_prev, const double* v_prev, double* out_dev) {
  if (int rc = check_ctx(c)) return rc;
  if (!u || !v || !out_dev) {
    set_last_error("pcfd_diagnostics_noalloc: u, v and out_dev must be device pointers");
    return kErrBadArg;
  }
  const size_t need = diagnostics_partials(c->nx, c->ny) * sizeof(double);
  if (need > kNoAllocScratchBytes) {
    set_last_error(
        "pcfd_diagnostics_noalloc: a %dx%d grid needs %zu bytes of reduction scratch, more than "
        "the %zu-byte maximum; use pcfd_diagnostics",
        c->nx, c->ny, need, kNoAllocScratchBytes);
    return kErrBadArg;
  }
  // Every partial the fold reads is written by this call's first kernel, so
  // the scratch needs no clearing.
  int rc = launch_diagnostics(c->stream, u, v, p, u_prev, v_prev, c->solid, c->nx, c->ny, c->ldu,
                              c->ldv, c->ldp, c->dx, c->dy, c->g, c->scratch, nullptr, out_dev);
  if (rc) return rc;
  PCFD_HIP_CHECK(hipMemcpyAsync(out_dev + 7, c->force, 2 * sizeof(double),
                                hipMemcpyDeviceToDevice, c->stream));
  return kOk;
}

int pcfd_get_force(pcfd_ctx* c, double* out_host) {
  if (int rc = check_ctx(c)) return rc;
  const size_t n = 2 + 2 * (size_t)c->n_bodies;
  PCFD_HIP_CHECK(
      hipMemcpyAsync(c->host_out, c->force, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PCFD_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (out_host) std::memcpy(out_host, c->host_out, n * sizeof(double));
  return kOk;
}

int pcfd_poisson_apply(pcfd_ctx* c, const double* x, double* y) {
  if (int rc = check_ctx(c)) return rc;
  return c->engine.apply_operator(x, y);
}

int pcfd_time_kernel(pcfd_ctx* c, int32_t which, int32_t reps, double* avg_ms,
                     double* bytes_per_launch) {
  if (int rc = check_ctx(c)) return rc;
  if (reps < 1 || !avg_ms || !bytes_per_launch) {
    set_last_error("pcfd_time_kernel: bad argument");
    return kErrBadArg;
  }
  return c->engine.time_kernel(which, reps, avg_ms, bytes_per_launch);
}

int pcfd_kernel_clock_start(pcfd_ctx* c, int32_t max_launches) {
  if (int rc = check_ctx(c)) return rc;
  return c->engine.clock_start(max_launches);
}

int pcfd_kernel_clock_read(pcfd_ctx* c, int32_t* launches, double* avg_ms,
                           double* bytes_per_launch) {
  if (int rc = check_ctx(c)) return rc;
  if (!launches || !avg_ms || !bytes_per_launch) {
    set_last_error("pcfd_kernel_clock_read: null output");
    return kErrBadArg;
  }
  return c->engine.clock_read(launches, avg_ms, bytes_per_launch);
}

int pcfd_mg_plan(pcfd_ctx* c, int32_t* out, int32_t cap) {
  if (int rc = check_ctx(c)) return -rc;
  if (!out || cap < 0) {
    set_last_error("pcfd_mg_plan: bad argument");
    return -kErrBadArg;
  }
  return c->engine.plan(out, cap);
}

int pcfd_mg_apply(pcfd_ctx* c, const double* r, double* z) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_solver(c)) return rc;
  return c->engine.apply_preconditioner(r, z);
}

}  // extern "C"
