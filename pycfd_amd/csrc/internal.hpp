// Declarations shared by the translation units of libpycfd_hip.so.
#pragma once

#include <cmath>

#include "common.hpp"

namespace pcfd {

// Wall-condition kinds (pycfd/config.py:85-94), same numbering as the C-ABI.
enum BCKind : int {
  BC_NO_SLIP = 0,
  BC_MOVING_WALL = 1,
  BC_INLET = 2,
  BC_OUTLET = 3,
  BC_PRESSURE_OUTLET = 4,
  BC_PERIODIC = 5,
  BC_SYMMETRY = 6,
};

// Operations of a boundary pass, executed in order by one workgroup.
enum BoundaryOp : int {
  BOP_VEL_PRED = 1,   // apply_velocity(predictor=True)  (boundary.py:375-379)
  BOP_VEL_CORR = 2,   // apply_velocity(predictor=False)
  BOP_MASS = 3,       // enforce_global_mass_balance      (boundary.py:423-487)
  BOP_MASK_RESET = 4, // _mask_obstacle(dt, reset=True)   (solver.py:553-599)
  BOP_MASK_ACC = 5,   // _mask_obstacle(dt, reset=False)
  BOP_MASK_ZERO = 6,  // _mask_obstacle(dt=None)
  BOP_WRAP_FACES = 7, // _wrap_periodic_faces             (solver.py:601-607)
};

struct BCOps {
  int n;
  int op[8];
};

// Spacings of a stretched mesh (pycfd/core/mesh.py:122-233) as device
// arrays.  All null on a uniform mesh, whose kernels divide by the scalars
// dx, dy with the reference's uniform arithmetic instead.  The reference
// switches every operator to the array form as soon as either axis is
// stretched, so a stretched mesh uses these on both axes.
struct Metrics {
  const double* hx;   // nx    cell widths
  const double* hy;   // ny    cell heights
  const double* hxu;  // nx+1  centre-to-centre distance across x face m
  const double* hyv;  // ny+1
  const double* hxe;  // nx+2  widths with one mirrored ghost at each end
  const double* hye;  // ny+2
  const double* wx;   // nx+1  corner weight of the upper node
  const double* wy;   // ny+1
};

struct BCParams {
  int nx, ny, ldu, ldv;
  int kind[4];             // left, right, bottom, top
  double vel[4];
  double pref[4];
  const double* inlet[4];  // device: full normal-velocity line of an INLET wall
  double dx, dy;
  int periodic_x, periodic_y;
  int has_obstacle;
  int n_uf, n_vf;          // immersed-boundary faces (flat indices into u / v)
  const int* uf;
  const int* vf;
  const int8_t* uf_body;
  const int8_t* vf_body;
  int n_bodies;
  double area;             // uniform cell area
  double* force;           // device [fx, fy, (fx, fy) per body]
  // Stretched mesh: control-volume area of each immersed-boundary face
  // (solver.py:197-207, 271-274) and the wall widths weighting the outlet
  // fluxes (boundary.py:438-459); null on a uniform mesh.
  const double* uf_area;
  const double* vf_area;
  const double* hx;        // nx cell widths
  const double* hy;        // ny cell heights
  // Every immersed-boundary face lies at least kMaskApart lines inside the
  // arrays, away from all that the wall conditions and the mass balance read
  // or write: masking then commutes with both and runs beside the walls.
  int mask_apart;
};
constexpr int kMaskApart = 4;

struct MomParams {
  int nx, ny, ldu, ldv;
  double dx, dy, dx2, dy2, nu, fx, fy, blend;
  // 1/dx, 1/dy, 1/dx^2, 1/dy^2: used in place of the divisions when dx and
  // dy are powers of two (pow2), where both give the same bits
  double idx, idy, idx2, idy2;
  int pow2;
  // Variable viscosity (Smagorinsky LES): cell-centred nu + nu_t on the
  // ghosted (nx+2, ny+2) layout, row pitch ny + 2; nullptr: constant nu.
  const double* nuc;
  Metrics g;  // stretched mesh, or all null
};

// Smagorinsky constant term (Cs * sqrt(dx dy))^2 of the eddy viscosity; on a
// stretched mesh the filter width is per cell (turbulence.py:47-58) and the
// kernel forms (cs * sqrt(hx hy))^2 itself.
struct LesParams {
  int nx, ny, ldu, ldv;
  double dx, dy, nu, coeff;
  double cs;
  Metrics g;
};

// The time step and upwind blend of a substep: by value, or, when `dev` is
// set, read from device memory when the kernel runs (a captured step graph
// is replayed with new values written there by its first node).
struct StepScalars {
  double dt = 0.0, blend = 0.0;
  const double* dev = nullptr;  // [dt, blend] or null
  __device__ __forceinline__ double get_dt() const { return dev ? dev[0] : dt; }
  __device__ __forceinline__ double get_blend() const { return dev ? dev[1] : blend; }
};

// Solvable face ranges u_upd / v_upd (solver.py:235-263): u faces i in
// [u_lo, u_hi) x rows 1..ny; v faces columns 1..nx x j in [v_lo, v_hi).
struct UpdRange {
  int u_lo, u_hi, v_lo, v_hi;
};

// stencil.hip
// nu_c = nu + coeff |S| on the ghosted cell layout (turbulence.py:63-113),
// from (u, v) or from the stage combination ca*(u0, v0) + cb*(ue, ve) when ue
// is given; optionally the corner average nu_corner (nx+1, ny+1).
int launch_eddy_viscosity(hipStream_t s, const LesParams& P, const double* u0, const double* ue,
                          const double* v0, const double* ve, double ca, double cb, double* nuc,
                          double* nu_corner);
int launch_momentum_rhs(hipStream_t s, const double* u, const double* v, const MomParams& P,
                        double* rhs_u, double* rhs_v);
// The predictors take dt and the blend from `S` (P.blend is ignored).  The
// step's launchers take an optional LaunchClock (the live kernel clock).
int launch_predictor(hipStream_t s, const double* u, const double* v, const MomParams& P,
                     const UpdRange& R, const StepScalars& S, double* us, double* vs,
                     const LaunchClock& clk = LaunchClock{});
int launch_predictor_comb(hipStream_t s, const double* u0, const double* ue, const double* v0,
                          const double* ve, double ca, double cb, const MomParams& P,
                          const UpdRange& R, const StepScalars& S, double* us, double* vs,
                          const LaunchClock& clk = LaunchClock{});
int launch_combine(hipStream_t s, const double* u0, const double* ue, const double* v0,
                   const double* ve, double ca, double cb, int nx, int ny, int ldu, int ldv,
                   double* uo, double* vo, const LaunchClock& clk = LaunchClock{});
// The launchers below take the mesh's Metrics: null arrays select the
// uniform arithmetic on dx, dy.
int launch_correct(hipStream_t s, const double* p, int ldp, int nx, int ny, int ldu, int ldv,
                   double dx, double dy, const Metrics& g, const UpdRange& R,
                   const StepScalars& S, double* u, double* v);
int launch_divergence(hipStream_t s, const double* u, const double* v, int nx, int ny, int ldu,
                      int ldv, double dx, double dy, const Metrics& g, double* out);
// The pressure ghost layer of the normalised interior p (apply_pressure)
// and the gradient correction of u, v, in one pass.
int launch_correct_ghosts(hipStream_t s, double* p, int ldp, const BCParams& B, int ldu, int ldv,
                          double dx, double dy, const Metrics& g, const UpdRange& R,
                          const StepScalars& T, double* u, double* v,
                          const LaunchClock& clk = LaunchClock{});
// A spacing whose reciprocal is exact (a power of two): division by it and
// multiplication by its reciprocal give the same bits, so kernels multiply.
inline bool is_pow2(double h) {
  int e = 0;
  return h > 0.0 && std::frexp(h, &e) == 0.5 && std::ldexp(1.0, -e + 1) * h == 1.0;
}
// Doubles of partial sums launch_diagnostics writes for an nx x ny grid.
size_t diagnostics_partials(int nx, int ny);
int launch_diagnostics(hipStream_t s, const double* u, const double* v, const double* p,
                       const double* up, const double* vp, const uint8_t* solid, int nx, int ny,
                       int ldu, int ldv, int ldp, double dx, double dy, const Metrics& g,
                       double* partial, unsigned int* counter, double* out,
                       const LaunchClock& clk = LaunchClock{});

// Clocks for launchers that issue several kernels: next(ctx, kind, bytes)
// hands out one LaunchClock per launch (none without a source).
struct ClockSource {
  LaunchClock (*next)(void* ctx, int kind, double bytes) = nullptr;
  void* ctx = nullptr;
  LaunchClock get(int kind, double bytes) const {
    return next ? next(ctx, kind, bytes) : LaunchClock{};
  }
};

// boundary.hip: the ClockSource kinds of its two kernels
constexpr int kClockWalls = 0, kClockBoundaryPass = 1;
int launch_boundary_pass(hipStream_t s, const BCParams& P, const BCOps& ops,
                         const StepScalars& S, double* u, double* v,
                         const ClockSource& clk = ClockSource{});
int launch_pressure_ghosts(hipStream_t s, const BCParams& P, double* p, int ldp);
// Transpose of apply_velocity(predictor) on adjoint fields lu, lv (walls
// with no periodic axis and no OUTLET).
int launch_walls_adjoint(hipStream_t s, const BCParams& P, bool predictor, double* lu,
                         double* lv);
// dst[0] = dt, dst[1] = blend: the first node of a captured step.
int launch_set_step_scalars(hipStream_t s, double* dst, double dt, double blend);

}  // namespace pcfd
