/*
 * pycfd_hip.h -- C-ABI of libpycfd_hip.so, the gfx950 (MI355X) engine behind
 * the projection time step of pycfd.
 *
 * The reference (kiliancmd/pycfd) is pure Python, so there is no existing FFI
 * to replace; each entry point below stands in for the Python method cited
 * next to it, and the Python package `pycfd_amd` binds them with ctypes
 * (pycfd_amd/_lib.py) behind the reference's own classes.
 *
 * Conventions
 *   - All field pointers are DEVICE pointers into FP64 arrays in the
 *     reference's ghosted C-order layout (pycfd/core/mesh.py:14-31):
 *       u (nx+3, ny+2), v (nx+2, ny+3), p (nx+2, ny+2), contiguous.
 *     Poisson vectors are the nx*ny interior cells, k = i*ny + j.
 *   - The library never frees caller memory.  It owns its scratch buffers.
 *   - Work is enqueued on the HIP stream given at creation; functions that
 *     return host values synchronise that stream, the others do not.
 *   - Every function returns a status code; pcfd_last_error() describes the
 *     most recent failure of the calling thread.
 *   - A context is not thread-safe: one context = one simulation = one stream.
 */
#ifndef PYCFD_HIP_H
#define PYCFD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PCFD_OK 0
#define PCFD_ERR_HIP 1
#define PCFD_ERR_NOT_CONVERGED 2
#define PCFD_ERR_STALLED 3
#define PCFD_ERR_NONFINITE 4
#define PCFD_ERR_BAD_ARG 5
#define PCFD_ERR_UNSUPPORTED 6
#define PCFD_ERR_NO_DEVICE 7

/* wall kinds, pycfd/config.py:85-94 (BCKind) */
#define PCFD_BC_NO_SLIP 0
#define PCFD_BC_MOVING_WALL 1
#define PCFD_BC_INLET 2
#define PCFD_BC_OUTLET 3
#define PCFD_BC_PRESSURE_OUTLET 4
#define PCFD_BC_PERIODIC 5
#define PCFD_BC_SYMMETRY 6

/* pressure solvers, pycfd/config.py:74-82 (PressureSolver) */
#define PCFD_SOLVER_DIRECT 0
#define PCFD_SOLVER_CG 1
#define PCFD_SOLVER_JACOBI 2
#define PCFD_SOLVER_SOR 3
#define PCFD_SOLVER_MULTIGRID 4
#define PCFD_SOLVER_MGCG 5

/* time schemes, pycfd/config.py:52-64 (TimeScheme) */
#define PCFD_SCHEME_EULER 0
#define PCFD_SCHEME_RK2 1
#define PCFD_SCHEME_RK3 2

/* boundary-pass operation codes for pcfd_mask_obstacle */
#define PCFD_MASK_ZERO 0
#define PCFD_MASK_RESET 1
#define PCFD_MASK_ACCUMULATE 2

typedef struct pcfd_ctx pcfd_ctx;

/* One case: grid, physics, walls, immersed body.  Host pointers are copied. */
typedef struct pcfd_case_desc {
  int32_t nx, ny;
  double dx, dy;                /* uniform spacing (mesh.py:313-529) */
  double nu;                    /* config.py:248-251 */
  double body_fx, body_fy;      /* config.py:209 */
  int32_t time_scheme;          /* PCFD_SCHEME_* */
  int32_t bc_kind[4];           /* PCFD_BC_*, walls left, right, bottom, top */
  double bc_velocity[4];
  double bc_p_ref[4];
  const double* inlet_line[4];  /* INLET walls: full normal-velocity line, length
                                   ny+2 (left/right) or nx+2 (bottom/top),
                                   boundary.py:205-216; NULL otherwise */
  const uint8_t* solid;         /* (nx*ny) 1 inside a body, or NULL */
  const int8_t* body_id;        /* (nx*ny) body of each solid cell, or NULL */
  int32_t n_bodies;             /* > 1 enables per-body forces (solver.py:136-146) */
  int32_t use_les;              /* Smagorinsky closure (config.py use_les, solver.py:164-167) */
  double les_coeff;             /* (Cs * sqrt(dx dy))^2 (turbulence.py:45-52) */
  /* Stretched mesh (mesh.py:122-233, 269-311): the nx cell widths and ny
     cell heights, both NULL for the uniform mesh of dx, dy.  With them every
     operator takes the reference's array-metric form.  area_mean is the mean
     cell area as numpy computes it (pressure.py:197); les_cs the Smagorinsky
     constant, whose filter width is then per cell (turbulence.py:47-58).  A
     stretched axis may not be periodic (solver.py:209-230). */
  const double* dx_cells;
  const double* dy_cells;
  double area_mean;
  double les_cs;
  /* 1: the body is forced by a soft mask passed to the pcfd_soft_* calls;
     `solid` then only fixes the Poisson operator's topology (no binary face
     forcing).  0: the reference's binary forcing (solver.py:553-599). */
  int32_t soft_ib;
} pcfd_case_desc;

typedef struct pcfd_solver_desc {
  int32_t kind;       /* PCFD_SOLVER_* */
  double tol;         /* relative residual (config.py:217) */
  int32_t maxiter;    /* config.py:218 */
  double omega;       /* SOR over-relaxation (config.py:219) or Jacobi damping */
  int32_t sweeps;     /* V-cycle pre/post sweeps (config.py:220) */
  int32_t precision;  /* 0: FP64 throughout; 1: FP32 V-cycle, FP64 Krylov */
  /* Multigrid level kernels (no counterpart in the reference: its V-cycle is
   * SciPy sparse products).  fused_levels = 0 runs every stored level pass by
   * pass; tile_tx / tile_ty fix the level-0 tile (0: measured choice);
   * coarse_chain picks the cooperative launch of the deepest levels (0:
   * measured choice, -1: never, k >= 1: from level k down).  A zeroed tail
   * keeps the defaults. */
  int32_t fused_levels;
  int32_t tile_tx, tile_ty;
  int32_t coarse_chain;
  /* Level 0's Krylov-start legs in streaming form (uniform mesh, one sweep):
   * 0 measured choice, -1 never, k > 0 always, k rows per wavefront (a
   * multiple of 4). */
  int32_t stream_rows;
  /* Stored-level passes (when a level runs pass by pass): 0 register-resident
   * rows wherever a row fits 8 entries per lane on 64 lanes, -1 always the
   * looping row kernels. */
  int32_t row_passes;
} pcfd_solver_desc;

/* One level of the aggregation hierarchy (multigrid.py:293-322).  Level 0 is
 * the fine operator, applied matrix-free: only nx, ny, fx, fy, active and
 * weight are read for it.  A coarser level carries its Galerkin operator as a
 * stencil dictionary: row k of the operator is row cls[k] of `table`
 * (n_classes x n_offsets, row-major), coefficient s acting on the cell at
 * offset (di[s], dj[s]); per class its 1/diag and active flag.  The coarsest
 * level carries the dense pseudo-inverse (coarse_n x coarse_n, row-major) of
 * its operator, projections included (multigrid.py:524-543). */
typedef struct pcfd_mg_level_desc {
  int32_t nx, ny;
  int32_t fx, fy;              /* aggregation factors towards the next level */
  int32_t n_offsets;
  const int8_t* di;
  const int8_t* dj;
  int32_t n_classes;
  const uint16_t* cls;         /* (nx*ny) */
  const double* table;         /* (n_classes*n_offsets) */
  const double* class_dinv;    /* (n_classes) */
  const uint8_t* class_active; /* (n_classes) */
  const uint8_t* active;       /* (nx*ny) */
  double weight;               /* damped-Jacobi weight (multigrid.py:354-362) */
  int32_t coarse_n;
  const double* coarse_inv;
  const uint32_t* cls32;       /* (nx*ny) in place of cls when n_classes > 65535
                                  (stretched meshes); NULL otherwise */
} pcfd_mg_level_desc;

typedef struct pcfd_solve_info {
  int32_t iterations;  /* last_iterations (pressure.py:296) */
  double residual;     /* last_residual, relative 2-norm (pressure.py:318-322) */
  int32_t converged;
  int32_t status;
} pcfd_solve_info;

/* The fine Poisson operator to assemble (pressure.py:110-270): the grid, the
 * body, the face coefficients (1/dx^2 and 1/dy^2 on a uniform mesh; per face
 * on a stretched one: face area / centre distance / mean cell area) and the
 * walls with a pressure value.  Host pointers, read during the call only. */
typedef struct pcfd_poisson_assembly {
  int32_t nx, ny;
  int32_t periodic_x, periodic_y;
  const uint8_t* solid;    /* (nx*ny) nonzero in a body, or NULL */
  double cx, cy;           /* uniform face coefficients */
  const double* cx_face;   /* ((nx+1)*ny) x-face coefficients of a stretched mesh, or NULL */
  const double* cy_face;   /* (nx*(ny+1)) y-face coefficients, or NULL */
  int32_t dirichlet[4];    /* walls left, right, bottom, top with a pressure value */
  double p_ref[4];         /* their values */
} pcfd_poisson_assembly;

/* The fine operator handed to the hierarchy build: the assembled Poisson
 * matrix (pressure.py:110-270) in compressed rows, the per-cell active mask,
 * and the power iteration's seeded start vector (numpy
 * default_rng(0).standard_normal(nx*ny); each level uses its first n
 * entries, which is what a fresh generator of that seed draws for n,
 * multigrid.py:147-149).  Host pointers, read during the call only. */
typedef struct pcfd_mg_operator {
  int32_t nx, ny;
  const int64_t* indptr;   /* (nx*ny + 1) */
  const int32_t* indices;  /* (indptr[nx*ny]) */
  const double* data;      /* (indptr[nx*ny]) */
  const uint8_t* active;   /* (nx*ny) */
  const double* x0;        /* (nx*ny) */
  int32_t singular;
  int32_t periodic_x, periodic_y;
  int32_t threads;         /* host threads for the build */
  int32_t coarse_max;      /* stop at the first level with at most this many
                              unknowns (at most 96: the whole hierarchy) */
  int32_t every_row;       /* 1: form every coarse row; 0: each distinct one
                              once, translated to the equal rows (the same
                              operator bit for bit, less work) */
} pcfd_mg_operator;

typedef struct pcfd_mg_setup pcfd_mg_setup;

/* One level of a built hierarchy. */
typedef struct pcfd_mg_built_level {
  int32_t n_levels;
  int32_t nx, ny;
  int32_t fx, fy;          /* aggregation factors towards the next level */
  int32_t pin;             /* first active cell (the coarsest level's pin) */
  int32_t n_offsets;       /* stencil offsets of the row dictionary (levels >= 1) */
  int64_t nnz;             /* stored entries of the level's operator */
  int64_t n_classes;       /* distinct operator rows (levels >= 1) */
  double weight;           /* damped-Jacobi weight (multigrid.py:354-362) */
} pcfd_mg_built_level;

/* --- lifetime --------------------------------------------------------- */
/* ProjectionSolver.__init__ (solver.py:101-195): device state, IB face lists,
 * Poisson operator flags (pressure.py:110-270).  hip_stream NULL = the
 * default stream. */
int pcfd_ctx_create(const pcfd_case_desc* desc, void* hip_stream, pcfd_ctx** out);
int pcfd_ctx_destroy(pcfd_ctx* ctx);
int pcfd_last_error(char* buf, size_t len);
int pcfd_device_count(int32_t* count);
int pcfd_synchronize(pcfd_ctx* ctx);
/* Host-only: the per-cell flag byte of the fine Poisson operator the device
 * applies (bit0 solvable, bits 1-4 faces to x-1, x+1, y-1, y+1 carry flux),
 * i.e. the sparsity of assemble_poisson_matrix (pressure.py:209-251).  Needs
 * no device; returns the solvable-cell count in *n_solvable. */
int pcfd_poisson_flags(const pcfd_case_desc* desc, uint8_t* flags_out, int32_t* n_solvable);

/* Host-only: assemble_poisson_matrix (pressure.py:110-270) in compressed
 * rows, columns ascending.  With indices == NULL only indptr (nx*ny + 1, the
 * row starts) is written; a second call with indptr from the first fills
 * indices and data (indptr[nx*ny] entries), the boundary right-hand side bc
 * (nx*ny) and the solvable mask (nx*ny; the other cells get identity rows). */
int pcfd_poisson_assemble(const pcfd_poisson_assembly* a, int64_t* indptr, int32_t* indices,
                          double* data, double* bc, uint8_t* solvable);
/* Host-only: MultigridHierarchy.__init__ (multigrid.py:324-474) -- the
 * aggregation hierarchy of an operator, every coarse level also turned into
 * its row dictionary (pcfd_mg_level_desc).  Built natively on `threads` host
 * threads; the result does not depend on the thread count. */
int pcfd_mg_build(const pcfd_mg_operator* op, pcfd_mg_setup** out);
int pcfd_mg_free(pcfd_mg_setup* h);
int pcfd_mg_level_info(const pcfd_mg_setup* h, int32_t level, pcfd_mg_built_level* info);
/* Copies one level out (any pointer may be NULL): its operator in compressed
 * rows (nnz entries, columns ascending; not for level 0, the caller's own
 * operator, which the build does not keep), 1/diag and the active mask (nx*ny),
 * and for levels >= 1 the row dictionary: offsets (n_offsets), the class of
 * every cell (nx*ny), the class table (n_classes*n_offsets, row-major), and per
 * class 1/diag and active flag.  Classes are numbered in the lexicographic
 * order of their rows. */
int pcfd_mg_level_copy(const pcfd_mg_setup* h, int32_t level, int64_t* indptr,
                       int32_t* indices, double* data, double* dinv, uint8_t* active,
                       int8_t* di, int8_t* dj, uint32_t* cls, double* table,
                       double* class_dinv, uint8_t* class_active);
/* make_pressure_solver (pressure.py:645-660) with a hierarchy built by the
 * host setup (MultigridHierarchy._build, multigrid.py:364-395). */
int pcfd_set_pressure_solver(pcfd_ctx* ctx, const pcfd_solver_desc* desc,
                             const pcfd_mg_level_desc* levels, int32_t n_levels);

/* --- operators of one substep ------------------------------------------- */
/* ProjectionSolver.momentum_rhs (solver.py:501-529): rhs_u (nx+1, ny),
 * rhs_v (nx, ny+1), contiguous. */
int pcfd_momentum_rhs(pcfd_ctx* ctx, const double* u, const double* v, double blend,
                      double* rhs_u, double* rhs_v);
/* SmagorinskyModel.eddy_viscosity (turbulence.py:101-113) of a context
 * created with use_les: nu_c (nx+2, ny+2) and, unless NULL, nu_corner
 * (nx+1, ny+1). */
int pcfd_eddy_viscosity(pcfd_ctx* ctx, const double* u, const double* v, double* nu_c,
                        double* nu_corner);
/* The momentum RHS with the stress-divergence viscous term of a given
 * cell-centred viscosity nu_c (nx+2, ny+2) (ProjectionSolver._diffusion_variable,
 * solver.py:467-499, inside momentum_rhs, solver.py:523-529). */
int pcfd_momentum_rhs_variable(pcfd_ctx* ctx, const double* u, const double* v, double blend,
                               const double* nu_c, double* rhs_u, double* rhs_v);
/* ProjectionSolver.divergence (solver.py:532-543): out (nx, ny). */
int pcfd_divergence(pcfd_ctx* ctx, const double* u, const double* v, double* out);
/* BoundaryManager.apply_velocity (boundary.py:375-379). */
int pcfd_apply_velocity_bc(pcfd_ctx* ctx, double* u, double* v, int32_t predictor);
/* BoundaryManager.enforce_global_mass_balance (boundary.py:423-487). */
int pcfd_mass_balance(pcfd_ctx* ctx, double* u, double* v);
/* ProjectionSolver._mask_obstacle (solver.py:553-599); mode PCFD_MASK_*. */
int pcfd_mask_obstacle(pcfd_ctx* ctx, double* u, double* v, double dt, int32_t mode);
/* ProjectionSolver._wrap_periodic_faces (solver.py:601-607). */
int pcfd_wrap_faces(pcfd_ctx* ctx, double* u, double* v);
/* BoundaryManager.apply_pressure (boundary.py:381-385) on a ghosted p. */
int pcfd_apply_pressure_bc(pcfd_ctx* ctx, double* p);

/* PoissonSolverBase.solve (pressure.py:325-327): rhs (nx, ny) -> p_out
 * (nx, ny), normalised; p0 (nx, ny) or NULL for a zero start.  Leaves the
 * warm-start guess of pcfd_project untouched. */
int pcfd_pressure_solve(pcfd_ctx* ctx, const double* rhs, const double* p0, double* p_out,
                        pcfd_solve_info* info);
/* ProjectionSolver.project (solver.py:643-683): u, v corrected in place, p
 * (ghosted) written; warm-started from the previous normalised pressure. */
int pcfd_project(pcfd_ctx* ctx, double* u, double* v, double dt, double* p,
                 pcfd_solve_info* info);
/* ProjectionSolver.initialize (solver.py:301-329) after the caller has written
 * the initial physical faces into u, v (ghosts zero). */
int pcfd_initialize(pcfd_ctx* ctx, double* u, double* v, double* p, pcfd_solve_info* info);
/* ProjectionSolver.step (solver.py:686-731): (u0, v0) -> (u, v, p). */
int pcfd_step(pcfd_ctx* ctx, const double* u0, const double* v0, double dt, double blend,
              double* u, double* v, double* p, pcfd_solve_info* info);
/* Whole steps as HIP graphs (no counterpart in the reference, whose loop is
 * Python).  With enable = 1, pcfd_step captures each (u0, v0) -> (u, v, p)
 * buffer set once (up to four sets are kept) and replays it: dt and the
 * blend are written by the graph's first node, and each solve runs a fixed
 * iteration budget (the last measured count plus a margin; iterations past
 * convergence are no-op kernels), so a step is one launch and one
 * synchronisation instead of a launch per kernel and a poll per solve.  The
 * first step after enabling (or after pcfd_set_pressure_solver) runs
 * eagerly to measure the counts; a replay whose solve ran out of budget is
 * redone eagerly from the same warm start and recaptured, so results are
 * those of the eager path.  Steps with u0 == u, and steps while the kernel
 * clock runs, are never captured. */
int pcfd_set_step_graphs(pcfd_ctx* ctx, int32_t enable);
/* Warm start of each substep's pressure solve (no counterpart in the
 * reference, whose ProjectionSolver._p_guess (solver.py:171, 662-663) is the
 * last pressure solved, order 0 here and the default): order 1-4 starts the
 * solve of RK substep s from the polynomial extrapolation in time of substep
 * s's last order+1 solutions (equal steps assumed), fewer while fewer are
 * stored.  The solve's tolerance is unchanged; only its iteration count
 * drops.  pcfd_reset_guess, pcfd_pressure_guess (store) and pcfd_initialize
 * restart the history.  Holds 18 nx*ny FP64 arrays once enabled. */
int pcfd_set_warm_start(pcfd_ctx* ctx, int32_t order);
/* Graph replays, captures and eager redos since the context was created. */
int pcfd_step_graph_stats(pcfd_ctx* ctx, int32_t* out3);
/* Forget the warm-start pressure (ProjectionSolver._p_guess = None). */
int pcfd_reset_guess(pcfd_ctx* ctx);
/* The warm-start pressure (ProjectionSolver._p_guess, solver.py:171,662-663)
 * as an (nx, ny) device array: store = 0 copies it out to buf, store = 1
 * replaces it with buf.  The reference keeps no such state in a checkpoint;
 * persisting it makes an iterative-solver resume continue bit for bit. */
int pcfd_pressure_guess(pcfd_ctx* ctx, double* buf, int32_t store);

/* --- reductions for TimeStepper (timestepper.py:94-142, 208-266) ----------
 * out[0] max|u_phys| out[1] max|v_phys| out[2] non-finite count in u, v, p
 * out[3] max|u - u_prev| out[4] max|v - v_prev| out[5] max|div| over fluid
 * out[6] sum over cells of uc^2 + vc^2, out[7..8] body_force_reaction (fx, fy).
 * out_host holds 9 doubles.  u_prev/v_prev/p may be NULL. */
int pcfd_diagnostics(pcfd_ctx* ctx, const double* u, const double* v, const double* p,
                     const double* u_prev, const double* v_prev, double* out_host);
/* No counterpart in the reference.  The reductions of pcfd_diagnostics for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the same nine values
 * (out_dev[0..6] as out_host above, out_dev[7..8] the total force), written
 * to the caller's DEVICE buffer out_dev (>= 9 doubles) on the context's
 * stream, without synchronising.  The reduction's partial sums live in a
 * fixed device array of PCFD_NOALLOC_SCRATCH_BYTES per GPU (held until the
 * process ends), so this call allocates nothing; grids whose partials exceed
 * it (about 19 M cells) are refused with PCFD_ERR_BAD_ARG.  All calls on one
 * GPU share that array: calls that overlap in time on one GPU (other
 * streams, other threads) race. */
#define PCFD_NOALLOC_SCRATCH_BYTES (1u << 20)
int pcfd_diagnostics_noalloc(pcfd_ctx* ctx, const double* u, const double* v, const double* p,
                             const double* u_prev, const double* v_prev, double* out_dev);
/* body_force_reaction then body_force_reactions (solver.py:173-178):
 * out_host has 2 + 2*n_bodies doubles. */
int pcfd_get_force(pcfd_ctx* ctx, double* out_host);

/* --- differentiable step with a soft immersed-boundary mask -------------
 * No counterpart in the reference (its mask is binary, solver.py:553-599).
 * A context created with soft_ib = 1: the body is a per-cell field m in
 * [0, 1]; a solvable face's weight is m_f = 1 - (1 - m_a)(1 - m_b) of its
 * two cells (solver.py:62-79 with the OR made soft), the forcing u_f <-
 * (1 - m_f) u_f after the predictor and after the projection, the reaction
 * force sum m_f u_f A / dt.  m in {0, 1} gives the binary step exactly.
 * The Poisson operator keeps the topology of the context's obstacle (the
 * binary body); faces touching an obstacle cell keep weight 1, so the soft
 * weights act on the faces between two fluid cells around the body.
 * SSP-RK3, uniform mesh, constant viscosity, central advection, walls of
 * kind inlet / no-slip / moving wall / symmetry / pressure outlet (p_ref 0).
 * All arrays are device arrays; mf_u, mf_v, lam_*, g_mf_* have the ghosted
 * shapes of u and v, m and g_theta (nx, ny). */
/* Face weights from the cell mask. */
int pcfd_soft_face_masks(pcfd_ctx* ctx, const double* m, double* mf_u, double* mf_v);
/* Doubles of one step's tape (the states the adjoint needs). */
int pcfd_soft_tape_size(pcfd_ctx* ctx, int64_t* doubles);
/* One step (u0, v0) -> (u, v, p) with fixed dt; the tape (or NULL) records
 * it; force_dev (2 doubles, or NULL) receives the last substep's force. */
int pcfd_soft_step(pcfd_ctx* ctx, const double* u0, const double* v0, double dt,
                   const double* mf_u, const double* mf_v, double* u, double* v, double* p,
                   double* tape, double* force_dev);
/* Transpose of a recorded step: lam_u, lam_v enter as the adjoint of its
 * output (u, v) and leave as that of its input (u0, v0); dL_dFx weights the
 * step's x-force; g_mf_u, g_mf_v accumulate the adjoint of the face
 * weights.  The transposed projection solves with the configured solver
 * from a zero start (the forward's warm start is kept). */
int pcfd_soft_step_adjoint(pcfd_ctx* ctx, const double* tape, double dt, const double* mf_u,
                           const double* mf_v, double dL_dFx, double* lam_u, double* lam_v,
                           double* g_mf_u, double* g_mf_v);
/* g_theta += (d m_f / d m)^T g_mf * m (1 - m): the face adjoints on the
 * cells, through m = sigmoid(theta). */
int pcfd_soft_cell_gradient(pcfd_ctx* ctx, const double* m, const double* g_mf_u,
                            const double* g_mf_v, double* g_theta);

/* --- Poisson building blocks (tests, autotuning) ----------------------- */
/* y = A x with the fine operator (pressure.py:110-270); both (nx, ny). */
int pcfd_poisson_apply(pcfd_ctx* ctx, const double* x, double* y);
/* z = M r, one V-cycle from zero with the projections
 * (MultigridHierarchy.apply, multigrid.py:460-474). */
int pcfd_mg_apply(pcfd_ctx* ctx, const double* r, double* z);

/* --- measurement -------------------------------------------------------
 * Times `reps` back-to-back launches of one hot fine-level kernel on the
 * context's stream with HIP events (which: 0 CG update, 1 multigrid-CG
 * direction + SpMV, 2 level-0 V-cycle down leg, 3 level-0 up leg) and reports
 * the mean duration and the kernel's algorithmic bytes per launch.  No
 * reference counterpart: measurement only. */
int pcfd_time_kernel(pcfd_ctx* ctx, int32_t which, int32_t reps, double* avg_ms,
                     double* bytes_per_launch);
/* Every kernel of the time steps and pressure solves that follow, timed
 * live: each launch records events at the kernel's own start and end
 * (hipExtLaunchKernelGGL), up to max_launches launches in all; launches of
 * iterations enqueued past convergence (no-ops) are dropped.  _read stops the
 * clock and fills, per slot (arrays of PCFD_CLOCK_SLOTS), launches timed, mean
 * duration in ms and algorithmic bytes per launch.  slot = kind *
 * PCFD_CLOCK_LEVELS + level, kinds: 0 CG update, 1 multigrid-CG direction,
 * 2 / 3 level-0 down / up leg, 4 / 5 fused stored-level down / up leg, 6-9
 * stored-level row passes (zero-start smooth + residual, smooth, residual,
 * prolongation), 10 restriction, 11 dense coarsest solve, 12 cooperative
 * chain, 13 reductions (level 0: cycle (standalone multigrid; multigrid-CG's
 * direction kernel folds the cycle's sums itself), 1: alpha (Jacobi-CG only;
 * multigrid-CG's update kernel folds alpha itself), 2: update, 3: right-hand
 * side mean, 4: solve start, 5: final check); the step's other kernels at
 * level 0: 14 predictor, 15 stage-combining predictor, 16 wall conditions,
 * 17 one-workgroup boundary pass, 18 right-hand side, 19 solve start, 20
 * final residual check, 21 pressure scatter, 22 gradient correction, 23
 * stage combination, 24 diagnostics, 25 one-workgroup Jacobi-CG solve, 26
 * warm-start extrapolation.  No reference counterpart: measurement only. */
#define PCFD_CLOCK_KINDS 27
#define PCFD_CLOCK_LEVELS 16
#define PCFD_CLOCK_SLOTS (PCFD_CLOCK_KINDS * PCFD_CLOCK_LEVELS)
int pcfd_kernel_clock_start(pcfd_ctx* ctx, int32_t max_launches);
int pcfd_kernel_clock_read(pcfd_ctx* ctx, int32_t* launches, double* avg_ms,
                           double* bytes_per_launch);

/* The multigrid plan of the configured solver, for reporting: per level (all
 * levels of the cycle precision, finest first) five integers -- nx, ny,
 * mode (1: one fused launch per leg, 0: pass by pass, 2: inside the
 * cooperative launch of the deepest levels, 3: streaming legs, tile = rows
 * per wavefront x output columns per wavefront; the coarsest level is the
 * dense solve), tile x, tile y -- written to out (capacity
 * `cap` integers).  Returns the number of levels, or a negative status.  No
 * reference counterpart. */
int pcfd_mg_plan(pcfd_ctx* ctx, int32_t* out, int32_t cap);

#ifdef __cplusplus
}
#endif

#endif /* PYCFD_HIP_H */
