// Field stencils of the projection step on the staggered MAC grid.
//
// Every expression below keeps the operation order of the reference's NumPy
// array path (pycfd/core/solver.py:334-465, 532-543, 677-682 and the fused
// kernel pycfd/core/kernels.py:39-110), and the library is compiled with
// -ffp-contract=off, so each output is bit-identical to the reference for the
// same inputs: no fused multiply-add, IEEE division by dx, dy, dx^2, dy^2.
// A stretched mesh has its own kernel instances (template flag S) with the
// reference's array-metric arithmetic; the uniform ones are unchanged.
#include "internal.hpp"

#include <type_traits>

namespace pcfd {

// ----------------------------------------------------------------------------
// Variable-viscosity (stress-divergence) viscous terms, solver.py:467-499 on a
// uniform mesh.  nu_c is cell-centred on the ghosted layout; the corner value
// is the four-cell average of turbulence.py:107-110; tau = nu_k (du/dy +
// dv/dx) at corners; the normal stresses 2 nu_c du/dx, 2 nu_c dv/dy at centres.
// Index conventions below are the ghosted arrays': u face m, v face n, cell
// (i, j) of nu_c, corner (a, b) between cells a, a+1 and b, b+1.
// ----------------------------------------------------------------------------
// Spacings by ghosted index (mesh.py:122-233): the scalar dx / dy on a
// uniform mesh (S = false), the metric arrays on a stretched one.  HX(i):
// width of ghosted cell i (1..nx); HXU(m): span across ghosted u face m
// (1..nx+1); HXE(q): width of ghosted cell q including the ghosts (0..nx+1).
template <bool S> struct Sp;
template <> struct Sp<false> {
  template <class P> static __device__ __forceinline__ double HX(const P& p, int) { return p.dx; }
  template <class P> static __device__ __forceinline__ double HY(const P& p, int) { return p.dy; }
  template <class P> static __device__ __forceinline__ double HXU(const P& p, int) { return p.dx; }
  template <class P> static __device__ __forceinline__ double HYV(const P& p, int) { return p.dy; }
  template <class P> static __device__ __forceinline__ double HXE(const P& p, int) { return p.dx; }
  template <class P> static __device__ __forceinline__ double HYE(const P& p, int) { return p.dy; }
};
template <> struct Sp<true> {
  template <class P> static __device__ __forceinline__ double HX(const P& p, int i) { return p.g.hx[i - 1]; }
  template <class P> static __device__ __forceinline__ double HY(const P& p, int j) { return p.g.hy[j - 1]; }
  template <class P> static __device__ __forceinline__ double HXU(const P& p, int m) { return p.g.hxu[m - 1]; }
  template <class P> static __device__ __forceinline__ double HYV(const P& p, int n) { return p.g.hyv[n - 1]; }
  template <class P> static __device__ __forceinline__ double HXE(const P& p, int q) { return p.g.hxe[q]; }
  template <class P> static __device__ __forceinline__ double HYE(const P& p, int q) { return p.g.hye[q]; }
};

// tau at corner (a, b): the shear derivatives span a face, hyv[b] and
// hxu[a] on a stretched mesh (solver.py:482-486).
template <bool S, class FU, class FV>
__device__ __forceinline__ double tau_corner(const FU& U, const FV& V, int a, int b,
                                             const MomParams& P) {
  const double* nc = P.nuc;
  const int ld = P.ny + 2;
  const double nuk = 0.25 * (nc[(size_t)a * ld + b] + nc[(size_t)(a + 1) * ld + b] +
                             nc[(size_t)a * ld + b + 1] + nc[(size_t)(a + 1) * ld + b + 1]);
  const double dudy = (U(a + 1, b + 1) - U(a + 1, b)) / Sp<S>::HYV(P, b + 1);
  const double dvdx = (V(a + 1, b + 1) - V(a, b + 1)) / Sp<S>::HXU(P, a + 1);
  return nuk * (dudy + dvdx);
}

template <bool S, class FU, class FV>
__device__ __forceinline__ double visc_u_face(const FU& U, const FV& V, int m, int j,
                                              const MomParams& P) {
  const double* nc = P.nuc;
  const int ld = P.ny + 2;
  const double gx_hi = (U(m + 1, j) - U(m, j)) / Sp<S>::HXE(P, m);      // du/dx in cell m
  const double gx_lo = (U(m, j) - U(m - 1, j)) / Sp<S>::HXE(P, m - 1);  // du/dx in cell m-1
  const double sxx = 2.0 * (nc[(size_t)m * ld + j] * gx_hi - nc[(size_t)(m - 1) * ld + j] * gx_lo) /
                     Sp<S>::HXU(P, m);
  const double sxy = (tau_corner<S>(U, V, m - 1, j, P) - tau_corner<S>(U, V, m - 1, j - 1, P)) /
                     Sp<S>::HY(P, j);
  return sxx + sxy;
}

template <bool S, class FU, class FV>
__device__ __forceinline__ double visc_v_face(const FU& U, const FV& V, int i, int n,
                                              const MomParams& P) {
  const double* nc = P.nuc;
  const int ld = P.ny + 2;
  const double syx = (tau_corner<S>(U, V, i, n - 1, P) - tau_corner<S>(U, V, i - 1, n - 1, P)) /
                     Sp<S>::HX(P, i);
  const double gy_hi = (V(i, n + 1) - V(i, n)) / Sp<S>::HYE(P, n);
  const double gy_lo = (V(i, n) - V(i, n - 1)) / Sp<S>::HYE(P, n - 1);
  const double syy = 2.0 * (nc[(size_t)i * ld + n] * gy_hi - nc[(size_t)i * ld + n - 1] * gy_lo) /
                     Sp<S>::HYV(P, n);
  return syx + syy;
}

// ----------------------------------------------------------------------------
// Momentum right-hand side, -A(u) + nu*lap(u) + f, evaluated at one face.
// `U(i, j)` / `V(i, j)` return the ghosted velocity arrays' entries; they are
// functors so a Runge-Kutta stage combination can be folded into the loads.
// ----------------------------------------------------------------------------
// a / h on a uniform mesh; a * (1 / h), the same bits, when h is a power of
// two (P2)
template <bool P2>
__device__ __forceinline__ double dvd(double a, double h, double inv_h) {
  if constexpr (P2) return a * inv_h;
  else return a / h;
}

template <bool P2, class FU, class FV>
__device__ __forceinline__ double rhs_u_uniform(const FU& U, const FV& V, int m, int j,
                                                const MomParams& P) {
  const double u0 = U(m, j);
  const double um = U(m - 1, j), up = U(m + 1, j);
  const double ud = U(m, j - 1), uu = U(m, j + 1);
  // cell-centred interpolants right/left of the face (solver.py:376, 382-383)
  const double uu_r = 0.5 * (u0 + up);
  const double uu_l = 0.5 * (um + u0);
  double duudx = dvd<P2>(uu_r * uu_r - uu_l * uu_l, P.dx, P.idx);
  // corner products above/below the face (solver.py:351-357, 385)
  const double u_top = 0.5 * (u0 + uu);
  const double v_top = 0.5 * (V(m - 1, j + 1) + V(m, j + 1));
  const double u_bot = 0.5 * (ud + u0);
  const double v_bot = 0.5 * (V(m - 1, j) + V(m, j));
  double duvdy = dvd<P2>(u_top * v_top - u_bot * v_bot, P.dy, P.idy);
  if (P.blend > 0.0) {  // donor-cell correction (solver.py:393-407)
    duudx += dvd<P2>(P.blend, P.dx, P.idx) *
             (fabs(uu_r) * (u0 - up) / 2.0 - fabs(uu_l) * (um - u0) / 2.0);
    duvdy += dvd<P2>(P.blend, P.dy, P.idy) *
             (fabs(v_top) * (u0 - uu) / 2.0 - fabs(v_bot) * (ud - u0) / 2.0);
  }
  if (P.nuc) return -(duudx + duvdy) + visc_u_face<false>(U, V, m, j, P) + P.fx;
  const double d2u =
      dvd<P2>(um - 2.0 * u0 + up, P.dx2, P.idx2) + dvd<P2>(ud - 2.0 * u0 + uu, P.dy2, P.idy2);
  return -(duudx + duvdy) + P.nu * d2u + P.fx;
}

template <bool P2, class FU, class FV>
__device__ __forceinline__ double rhs_v_uniform(const FU& U, const FV& V, int i, int n,
                                                const MomParams& P) {
  const double v0 = V(i, n);
  const double vm = V(i, n - 1), vp = V(i, n + 1);
  const double vl = V(i - 1, n), vr = V(i + 1, n);
  const double vv_t = 0.5 * (v0 + vp);
  const double vv_b = 0.5 * (vm + v0);
  double dvvdy = dvd<P2>(vv_t * vv_t - vv_b * vv_b, P.dy, P.idy);
  const double u_right = 0.5 * (U(i + 1, n - 1) + U(i + 1, n));
  const double v_right = 0.5 * (v0 + vr);
  const double u_left = 0.5 * (U(i, n - 1) + U(i, n));
  const double v_left = 0.5 * (vl + v0);
  double duvdx = dvd<P2>(u_right * v_right - u_left * v_left, P.dx, P.idx);
  if (P.blend > 0.0) {  // solver.py:409-417
    dvvdy += dvd<P2>(P.blend, P.dy, P.idy) *
             (fabs(vv_t) * (v0 - vp) / 2.0 - fabs(vv_b) * (vm - v0) / 2.0);
    duvdx += dvd<P2>(P.blend, P.dx, P.idx) *
             (fabs(u_right) * (v0 - vr) / 2.0 - fabs(u_left) * (vl - v0) / 2.0);
  }
  if (P.nuc) return -(duvdx + dvvdy) + visc_v_face<false>(U, V, i, n, P) + P.fy;
  const double d2v =
      dvd<P2>(vl - 2.0 * v0 + vr, P.dx2, P.idx2) + dvd<P2>(vm - 2.0 * v0 + vp, P.dy2, P.idy2);
  return -(duvdx + dvvdy) + P.nu * d2v + P.fy;
}

// The same on a stretched mesh (solver.py:334-465 with array metrics):
// corner values interpolated with the upper-node weights wy / wx, each flux
// differenced over its face's own control volume (hxu, hy for u; hx, hyv for
// v), the viscous term as a difference of face gradients, each over the
// distance it spans.  u face m, row j: rhs row m-1, column j-1.
template <class FU, class FV>
__device__ __forceinline__ double rhs_u_stretched(const FU& U, const FV& V, int m, int j,
                                                  const MomParams& P) {
  const Metrics& g = P.g;
  const double u0 = U(m, j);
  const double um = U(m - 1, j), up = U(m + 1, j);
  const double ud = U(m, j - 1), uu = U(m, j + 1);
  const double uu_r = 0.5 * (u0 + up);
  const double uu_l = 0.5 * (um + u0);
  const double hxu = g.hxu[m - 1], hy = g.hy[j - 1];
  double duudx = (uu_r * uu_r - uu_l * uu_l) / hxu;
  // corners (m-1, j) above and (m-1, j-1) below: u_c = u_lo + wy (u_hi - u_lo),
  // v_c = v_lo + wx (v_hi - v_lo) (solver.py:353-358)
  const double u_top = u0 + g.wy[j] * (uu - u0);
  const double v_lo_t = V(m - 1, j + 1);
  const double v_top = v_lo_t + g.wx[m - 1] * (V(m, j + 1) - v_lo_t);
  const double u_bot = ud + g.wy[j - 1] * (u0 - ud);
  const double v_lo_b = V(m - 1, j);
  const double v_bot = v_lo_b + g.wx[m - 1] * (V(m, j) - v_lo_b);
  double duvdy = (u_top * v_top - u_bot * v_bot) / hy;
  if (P.blend > 0.0) {
    duudx += P.blend / hxu * (fabs(uu_r) * (u0 - up) / 2.0 - fabs(uu_l) * (um - u0) / 2.0);
    duvdy += P.blend / hy * (fabs(v_top) * (u0 - uu) / 2.0 - fabs(v_bot) * (ud - u0) / 2.0);
  }
  if (P.nuc) return -(duudx + duvdy) + visc_u_face<true>(U, V, m, j, P) + P.fx;
  const double gx_hi = (up - u0) / g.hxe[m];
  const double gx_lo = (u0 - um) / g.hxe[m - 1];
  const double gy_hi = (uu - u0) / g.hyv[j];
  const double gy_lo = (u0 - ud) / g.hyv[j - 1];
  const double d2u = (gx_hi - gx_lo) / hxu + (gy_hi - gy_lo) / hy;
  return -(duudx + duvdy) + P.nu * d2u + P.fx;
}

// v face (i, n): rhs row i-1, column n-1.
template <class FU, class FV>
__device__ __forceinline__ double rhs_v_stretched(const FU& U, const FV& V, int i, int n,
                                                  const MomParams& P) {
  const Metrics& g = P.g;
  const double v0 = V(i, n);
  const double vm = V(i, n - 1), vp = V(i, n + 1);
  const double vl = V(i - 1, n), vr = V(i + 1, n);
  const double vv_t = 0.5 * (v0 + vp);
  const double vv_b = 0.5 * (vm + v0);
  const double hx = g.hx[i - 1], hyv = g.hyv[n - 1];
  double dvvdy = (vv_t * vv_t - vv_b * vv_b) / hyv;
  // corners (i, n-1) right and (i-1, n-1) left
  const double u_lo_r = U(i + 1, n - 1);
  const double u_right = u_lo_r + g.wy[n - 1] * (U(i + 1, n) - u_lo_r);
  const double v_right = v0 + g.wx[i] * (vr - v0);
  const double u_lo_l = U(i, n - 1);
  const double u_left = u_lo_l + g.wy[n - 1] * (U(i, n) - u_lo_l);
  const double v_left = vl + g.wx[i - 1] * (v0 - vl);
  double duvdx = (u_right * v_right - u_left * v_left) / hx;
  if (P.blend > 0.0) {
    dvvdy += P.blend / hyv * (fabs(vv_t) * (v0 - vp) / 2.0 - fabs(vv_b) * (vm - v0) / 2.0);
    duvdx += P.blend / hx * (fabs(u_right) * (v0 - vr) / 2.0 - fabs(u_left) * (vl - v0) / 2.0);
  }
  if (P.nuc) return -(duvdx + dvvdy) + visc_v_face<true>(U, V, i, n, P) + P.fy;
  const double hx_hi = (vr - v0) / g.hxu[i];
  const double hx_lo = (v0 - vl) / g.hxu[i - 1];
  const double hy_hi = (vp - v0) / g.hye[n];
  const double hy_lo = (v0 - vm) / g.hye[n - 1];
  const double d2v = (hx_hi - hx_lo) / hx + (hy_hi - hy_lo) / hyv;
  return -(duvdx + dvvdy) + P.nu * d2v + P.fy;
}

template <bool S, bool P2, class FU, class FV>
__device__ __forceinline__ double rhs_u_face(const FU& U, const FV& V, int m, int j,
                                             const MomParams& P) {
  if constexpr (S) return rhs_u_stretched(U, V, m, j, P);
  else return rhs_u_uniform<P2>(U, V, m, j, P);
}

template <bool S, bool P2, class FU, class FV>
__device__ __forceinline__ double rhs_v_face(const FU& U, const FV& V, int i, int n,
                                             const MomParams& P) {
  if constexpr (S) return rhs_v_stretched(U, V, i, n, P);
  else return rhs_v_uniform<P2>(U, V, i, n, P);
}

struct PlainLoad {
  const double* a;
  int ld;
  __device__ __forceinline__ double operator()(int i, int j) const { return a[(size_t)i * ld + j]; }
};

// ca*a + cb*b, the SSP-RK3 stage combination (solver.py:716-717), fused into
// the loads of the next substep's stencil.
struct CombLoad {
  const double* a;
  const double* b;
  int ld;
  double ca, cb;
  __device__ __forceinline__ double operator()(int i, int j) const {
    const size_t k = (size_t)i * ld + j;
    return ca * a[k] + cb * b[k];
  }
};

// Smagorinsky eddy viscosity (turbulence.py:63-113, uniform mesh), one cell of
// the ghosted (nx+2, ny+2) layout per thread.  The corner shear rate is
// averaged to interior cells; ghost cells take the nearest interior value
// (the reference copies rows, then columns, into the ghost ring).
template <bool S, class FU, class FV>
__device__ __forceinline__ double s12_corner(const FU& U, const FV& V, int a, int b,
                                             const LesParams& P) {
  const double dudy = (U(a + 1, b + 1) - U(a + 1, b)) / Sp<S>::HYV(P, b + 1);
  const double dvdx = (V(a + 1, b + 1) - V(a, b + 1)) / Sp<S>::HXU(P, a + 1);
  return 0.5 * (dudy + dvdx);
}

template <bool S, class FU, class FV>
__device__ __forceinline__ void eddy_viscosity_body(const FU& U, const FV& V, const LesParams& P,
                                                    double* __restrict__ nuc,
                                                    double* __restrict__ nuk) {
  const Tile2D t = tile_coords(P.ny + 2);
  const int i = t.i, j = t.j;
  if (i > P.nx + 1 || j > P.ny + 1) return;
  const double s11 = (U(i + 1, j) - U(i, j)) / Sp<S>::HXE(P, i);
  const double s22 = (V(i, j + 1) - V(i, j)) / Sp<S>::HYE(P, j);
  const int ic = min(max(i, 1), P.nx), jc = min(max(j, 1), P.ny);
  const double s12 = 0.25 * (s12_corner<S>(U, V, ic - 1, jc - 1, P) +
                             s12_corner<S>(U, V, ic, jc - 1, P) +
                             s12_corner<S>(U, V, ic - 1, jc, P) + s12_corner<S>(U, V, ic, jc, P));
  const double s_mag = sqrt(2.0 * (s11 * s11) + 2.0 * (s22 * s22) + 4.0 * (s12 * s12));
  double coeff = P.coeff;
  if constexpr (S) {  // (cs * sqrt(hx hy))^2, ghost ring repeating its neighbour
    const double cd = P.cs * sqrt(P.g.hx[ic - 1] * P.g.hy[jc - 1]);
    coeff = cd * cd;
  }
  nuc[(size_t)i * (P.ny + 2) + j] = P.nu + coeff * s_mag;
  (void)nuk;
}

template <bool S, class FU, class FV>
__global__ void __launch_bounds__(kBlock)
k_eddy_viscosity(FU U, FV V, LesParams P, double* __restrict__ nuc, double* __restrict__ nuk) {
  eddy_viscosity_body<S>(U, V, P, nuc, nuk);
}

// Corner average of nu_c (turbulence.py:107-110), (nx+1, ny+1).
__global__ void __launch_bounds__(kBlock)
k_corner_viscosity(const double* __restrict__ nuc, int nx, int ny, double* __restrict__ nuk) {
  const Tile2D t = tile_coords(ny + 1);
  const int a = t.i, b = t.j;
  if (a > nx || b > ny) return;
  const int ld = ny + 2;
  nuk[(size_t)a * (ny + 1) + b] =
      0.25 * (nuc[(size_t)a * ld + b] + nuc[(size_t)(a + 1) * ld + b] + nuc[(size_t)a * ld + b + 1] +
              nuc[(size_t)(a + 1) * ld + b + 1]);
}

// rhs_u (nx+1, ny) and rhs_v (nx, ny+1), contiguous (solver.py:501-529).
template <bool S, bool P2>
__global__ void __launch_bounds__(kBlock)
k_momentum_rhs(const double* __restrict__ u, const double* __restrict__ v, MomParams P,
               double* __restrict__ rhs_u, double* __restrict__ rhs_v) {
  const Tile2D t = tile_coords(P.ny + 1);
  const int i = t.i, j = t.j;  // i: 0..nx, j: 0..ny
  PlainLoad U{u, P.ldu}, V{v, P.ldv};
  if (i <= P.nx && j < P.ny)
    rhs_u[(size_t)i * P.ny + j] = rhs_u_face<S, P2>(U, V, i + 1, j + 1, P);
  if (i < P.nx && j <= P.ny)
    rhs_v[(size_t)i * (P.ny + 1) + j] = rhs_v_face<S, P2>(U, V, i + 1, j + 1, P);
}

// Predictor us = u; us[u_upd] += dt*rhs_u (solver.py:617-622), both
// components, over the full ghosted arrays.
template <bool S, bool P2, class FU, class FV>
__device__ __forceinline__ void predictor_body(const FU& U, const FV& V, const MomParams& P,
                                               const UpdRange& R, double dt,
                                               double* __restrict__ us, double* __restrict__ vs) {
  const Tile2D t = tile_coords(P.ny + 3);
  const int i = t.i, j = t.j;
  if (i <= P.nx + 2 && j <= P.ny + 1) {
    double val = U(i, j);
    if (i >= R.u_lo && i < R.u_hi && j >= 1 && j <= P.ny)
      val += dt * rhs_u_face<S, P2>(U, V, i, j, P);
    us[(size_t)i * P.ldu + j] = val;
  }
  if (i <= P.nx + 1 && j <= P.ny + 2) {
    double val = V(i, j);
    if (i >= 1 && i <= P.nx && j >= R.v_lo && j < R.v_hi)
      val += dt * rhs_v_face<S, P2>(U, V, i, j, P);
    vs[(size_t)i * P.ldv + j] = val;
  }
}

template <bool S, bool P2>
__global__ void __launch_bounds__(kBlock)
k_predictor(const double* __restrict__ u, const double* __restrict__ v, MomParams P,
            UpdRange R, StepScalars T, double* __restrict__ us, double* __restrict__ vs) {
  P.blend = T.get_blend();
  predictor_body<S, P2>(PlainLoad{u, P.ldu}, PlainLoad{v, P.ldv}, P, R, T.get_dt(), us, vs);
}

template <bool S, bool P2>
__global__ void __launch_bounds__(kBlock)
k_predictor_comb(const double* __restrict__ u0, const double* __restrict__ ue,
                 const double* __restrict__ v0, const double* __restrict__ ve, double ca,
                 double cb, MomParams P, UpdRange R, StepScalars T, double* __restrict__ us,
                 double* __restrict__ vs) {
  P.blend = T.get_blend();
  predictor_body<S, P2>(CombLoad{u0, ue, P.ldu, ca, cb}, CombLoad{v0, ve, P.ldv, ca, cb}, P, R,
                        T.get_dt(), us, vs);
}

// Row-blocked predictor on a uniform mesh: a thread evaluates the u and v
// faces of kPredRows consecutive x rows at one y index.  The values its faces
// read -- rows i0-1 .. i0+kPredRows, columns j-1 .. j+1 of u and of v, the
// stage combination folded in -- are loaded once, at indices clamped into
// each array (a value at a clamped index is only read by faces outside the
// update ranges, which copy their own, in-range value), and the faces take
// them from registers: about 3.75 loads a face instead of 9.  The arithmetic
// of every face is rhs_u_uniform / rhs_v_uniform's, so the results are the
// one-face-per-thread kernel's bit for bit.
constexpr int kPredRows = 4;      // rows per thread, plain loads
constexpr int kPredRowsComb = 4;  // with the stage combination (2 rows measured slower)

// A view of a register cache c[a][b] = value (i0 - 1 + a, j0 - 1 + b).  The
// offsets live outside the array, so once the rows are unrolled every index
// folds to a constant and the array is promoted to registers.
struct RowCacheView {
  const double (*c)[3];
  int i0, j0;
  __device__ __forceinline__ double operator()(int i, int j) const {
    return c[i - i0 + 1][j - j0 + 1];
  }
};

template <bool P2, bool LES, int NR, class Src>
__global__ void __launch_bounds__(kBlock)
k_predictor_rows(Src Us, Src Vs, MomParams P, UpdRange R, StepScalars T, double* __restrict__ us,
                 double* __restrict__ vs) {
  P.blend = T.get_blend();
  if constexpr (!LES) P.nuc = nullptr;  // the constant-viscosity form, folded at compile time
  const double dt = T.get_dt();
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int ty = tiles_y(P.ny + 3);
  const int i0 = (tile / ty) * (kTileX * NR) + ((int)threadIdx.x / kTileY) * NR;
  const int j = (tile % ty) * kTileY + (int)threadIdx.x % kTileY;
  double cu[NR + 2][3], cv[NR + 2][3];
#pragma unroll
  for (int a = 0; a < NR + 2; ++a) {
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const int ii = i0 - 1 + a, jj = j - 1 + b;
      cu[a][b] = Us(min(max(ii, 0), P.nx + 2), min(max(jj, 0), P.ny + 1));
      cv[a][b] = Vs(min(max(ii, 0), P.nx + 1), min(max(jj, 0), P.ny + 2));
    }
  }
  const RowCacheView U{cu, i0, j}, V{cv, i0, j};
  // one row's two faces; r is a compile-time constant, so every cache index
  // is one and the cache stays in registers
  auto row = [&](auto rc) {
    constexpr int r = decltype(rc)::value;
    const int i = i0 + r;
    if (i <= P.nx + 2 && j <= P.ny + 1) {
      double val = cu[r + 1][1];
      if (i >= R.u_lo && i < R.u_hi && j >= 1 && j <= P.ny)
        val += dt * rhs_u_uniform<P2>(U, V, i, j, P);
      us[(size_t)i * P.ldu + j] = val;
    }
    if (i <= P.nx + 1 && j <= P.ny + 2) {
      double val = cv[r + 1][1];
      if (i >= 1 && i <= P.nx && j >= R.v_lo && j < R.v_hi)
        val += dt * rhs_v_uniform<P2>(U, V, i, j, P);
      vs[(size_t)i * P.ldv + j] = val;
    }
  };
  static_assert(NR == 2 || NR == 4, "the rows below");
  row(std::integral_constant<int, 0>{});
  row(std::integral_constant<int, 1>{});
  if constexpr (NR == 4) {
    row(std::integral_constant<int, 2>{});
    row(std::integral_constant<int, 3>{});
  }
}

static dim3 grid_rows(int nxr, int nyr, int nr) {
  return dim3((unsigned)(((nxr + kTileX * nr - 1) / (kTileX * nr)) * tiles_y(nyr)));
}

// out = ca*a + cb*b over both ghosted velocity arrays (solver.py:706-721).
__global__ void __launch_bounds__(kBlock)
k_combine(const double* __restrict__ u0, const double* __restrict__ ue,
          const double* __restrict__ v0, const double* __restrict__ ve, double ca, double cb,
          int nx, int ny, int ldu, int ldv, double* __restrict__ uo, double* __restrict__ vo) {
  const Tile2D t = tile_coords(ny + 3);
  const int i = t.i, j = t.j;
  if (i <= nx + 2 && j <= ny + 1) {
    const size_t k = (size_t)i * ldu + j;
    uo[k] = ca * u0[k] + cb * ue[k];
  }
  if (i <= nx + 1 && j <= ny + 2) {
    const size_t k = (size_t)i * ldv + j;
    vo[k] = ca * v0[k] + cb * ve[k];
  }
}

// Gradient correction u[u_upd] -= dt*(p[i]-p[i-1])/dx, same for v
// (solver.py:677-682), in place; divided by hxu / hyv of the face on a
// stretched mesh.
__global__ void __launch_bounds__(kBlock)
k_correct(const double* __restrict__ p, int ldp, int nx, int ny, int ldu, int ldv, double dx,
          double dy, Metrics M, UpdRange R, StepScalars T, double* __restrict__ u,
          double* __restrict__ v) {
  const double dt = T.get_dt();
  const Tile2D t = tile_coords(ny + 3);
  const int i = t.i, j = t.j;
  if (i >= R.u_lo && i < R.u_hi && j >= 1 && j <= ny) {
    const double h = M.hxu ? M.hxu[i - 1] : dx;
    const double g = dt * (p[(size_t)i * ldp + j] - p[(size_t)(i - 1) * ldp + j]) / h;
    u[(size_t)i * ldu + j] -= g;
  }
  if (i >= 1 && i <= nx && j >= R.v_lo && j < R.v_hi) {
    const double h = M.hyv ? M.hyv[j - 1] : dy;
    const double g = dt * (p[(size_t)i * ldp + j] - p[(size_t)i * ldp + j - 1]) / h;
    v[(size_t)i * ldv + j] -= g;
  }
}

// Gradient correction with the pressure ghost layer filled in the same pass
// (apply_pressure then the correction, solver.py:664-682, boundary.py:166-169,
// 293-303, 381-385, 412-420).  p holds the normalised interior; every ghost
// the correction or the output needs is formed from it: a wall's ghost is
// the adjacent interior value (Neumann) or 2 p_ref - p_in (pressure outlet),
// a periodic axis wraps, and the corner values follow the reference's wall
// order (left, right, bottom, top, then the x and y wraps), so the ghost
// ring written equals what the separate ghost pass writes.  The interior of
// p is only read, the ring only written: no thread reads what another
// writes.
struct PGhost {
  int nx, ny, ldp;
  int kind[4];
  double pref[4];
  int px, py;
};

__device__ __forceinline__ double p_wall(const PGhost& G, int w, double in) {
  return G.kind[w] == BC_PRESSURE_OUTLET ? 2.0 * G.pref[w] - in : in;
}

// Rows 1..ny of ghosted column i (0..nx+1).
__device__ __forceinline__ double p_col(const PGhost& G, const double* __restrict__ p, int i,
                                        int j) {
  if (i >= 1 && i <= G.nx) return p[(size_t)i * G.ldp + j];
  if (G.px) return p[(size_t)(i == 0 ? G.nx : 1) * G.ldp + j];
  const int src = (i == 0) ? 1 : G.nx;
  return p_wall(G, i == 0 ? 0 : 1, p[(size_t)src * G.ldp + j]);
}

// Any ghosted cell (0..nx+1, 0..ny+1).
__device__ __forceinline__ double p_full(const PGhost& G, const double* __restrict__ p, int i,
                                         int j) {
  if (j >= 1 && j <= G.ny) return p_col(G, p, i, j);
  if (G.py) return p_col(G, p, i, j == 0 ? G.ny : 1);
  return p_wall(G, j == 0 ? 2 : 3, p_col(G, p, i, j == 0 ? 1 : G.ny));
}

template <bool P2>
__global__ void __launch_bounds__(kBlock)
k_correct_ghosts(double* __restrict__ p, PGhost G, int ldu, int ldv, double dx, double dy,
                 double idx, double idy, Metrics M, UpdRange R, StepScalars T,
                 double* __restrict__ u, double* __restrict__ v) {
  const double dt = T.get_dt();
  const int nx = G.nx, ny = G.ny;
  const Tile2D t = tile_coords(ny + 3);
  const int i = t.i, j = t.j;
  if (i >= R.u_lo && i < R.u_hi && j >= 1 && j <= ny) {
    const double d = dt * (p_col(G, p, i, j) - p_col(G, p, i - 1, j));
    double g;
    if (M.hxu) g = d / M.hxu[i - 1];
    else if constexpr (P2) g = d * idx;
    else g = d / dx;
    u[(size_t)i * ldu + j] -= g;
  }
  if (i >= 1 && i <= nx && j >= R.v_lo && j < R.v_hi) {
    const double d = dt * (p_full(G, p, i, j) - p_full(G, p, i, j - 1));
    double g;
    if (M.hyv) g = d / M.hyv[j - 1];
    else if constexpr (P2) g = d * idy;
    else g = d / dy;
    v[(size_t)i * ldv + j] -= g;
  }
  if (i <= nx + 1 && j <= ny + 1 && (i == 0 || i == nx + 1 || j == 0 || j == ny + 1))
    p[(size_t)i * G.ldp + j] = p_full(G, p, i, j);
}

// Raw divergence (nx, ny) (solver.py:532-543), over each cell's own width
// and height on a stretched mesh.
__global__ void __launch_bounds__(kBlock)
k_divergence(const double* __restrict__ u, const double* __restrict__ v, int nx, int ny, int ldu,
             int ldv, double dx, double dy, Metrics M, double* __restrict__ out) {
  const Tile2D t = tile_coords(ny);
  const int i = t.i, j = t.j;
  if (i < nx && j < ny) {
    const int a = i + 1, b = j + 1;
    const double hx = M.hx ? M.hx[i] : dx, hy = M.hy ? M.hy[j] : dy;
    out[(size_t)i * ny + j] = (u[(size_t)(a + 1) * ldu + b] - u[(size_t)a * ldu + b]) / hx +
                              (v[(size_t)a * ldv + b + 1] - v[(size_t)a * ldv + b]) / hy;
  }
}

// Per-step diagnostics in one pass (timestepper.py:94-142, 208-212, 251-266;
// fields.py:70-96; solver.py:545-550):
//   out[0] max|u_phys|  out[1] max|v_phys|  out[2] non-finite count (u,v,p)
//   out[3] max|u-u_prev| out[4] max|v-v_prev| out[5] max|div| over fluid
//   out[6] sum(uc^2+vc^2) over cells, each weighted by its area on a
//          stretched mesh (mesh.py:213-223)
// A workgroup covers kDiagTiles tiles of the (nx+3) x (ny+3) cell range:
// every load of its cells is issued first, at indices clamped into each
// array's range (masked out afterwards), so no load waits on a branch; one
// row of partials serves the workgroup.  Optional inputs come as `have`
// bits (kHave*) with the launcher standing in a readable array of at least
// the same extent for an absent one, so the loads need no tests.
constexpr int kDiagTiles = 4;
constexpr int kHaveP = 1, kHavePrev = 2, kHaveSolid = 4, kHaveHx = 8, kHaveHy = 16;

static inline int diag_blocks(int nx, int ny) {
  return (num_tiles(nx + 3, ny + 3) + kDiagTiles - 1) / kDiagTiles;
}

__global__ void __launch_bounds__(kBlock)
k_diagnostics(const double* __restrict__ u, const double* __restrict__ v,
              const double* __restrict__ p, const double* __restrict__ up,
              const double* __restrict__ vp, const uint8_t* __restrict__ solid,
              const double* __restrict__ hx, const double* __restrict__ hy, int have, int nx,
              int ny, int ldu, int ldv, int ldp, double dx, double dy, double* partial) {
  const bool has_p = have & kHaveP, has_prev = have & kHavePrev, has_s = have & kHaveSolid;
  const bool has_hx = have & kHaveHx, has_hy = have & kHaveHy;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int ty = tiles_y(ny + 3), ntile = tiles_x(nx + 3) * ty;
  int ci[kDiagTiles], cj[kDiagTiles];
  bool in[kDiagTiles];
  double au[kDiagTiles], av[kDiagTiles], aup[kDiagTiles], avp[kDiagTiles], ap[kDiagTiles];
  double ua[kDiagTiles], ub[kDiagTiles], va[kDiagTiles], vb[kDiagTiles];
  double hxv[kDiagTiles], hyv[kDiagTiles];
  uint8_t so[kDiagTiles];
#pragma unroll
  for (int q = 0; q < kDiagTiles; ++q) {
    const int t = blk * kDiagTiles + q;
    const int i = (t / ty) * kTileX + (int)threadIdx.x / kTileY;
    const int j = (t % ty) * kTileY + (int)threadIdx.x % kTileY;
    ci[q] = i;
    cj[q] = j;
    in[q] = t < ntile;
    const size_t ku = (size_t)min(i, nx + 2) * ldu + min(j, ny + 1);
    const size_t kv = (size_t)min(i, nx + 1) * ldv + min(j, ny + 2);
    const int ii = max(1, min(i, nx)), jj = max(1, min(j, ny));
    au[q] = u[ku];
    av[q] = v[kv];
    ua[q] = u[(size_t)ii * ldu + jj];
    ub[q] = u[(size_t)(ii + 1) * ldu + jj];
    va[q] = v[(size_t)ii * ldv + jj];
    vb[q] = v[(size_t)ii * ldv + jj + 1];
    aup[q] = up[ku];
    avp[q] = vp[kv];
    ap[q] = p[(size_t)min(i, nx + 1) * ldp + min(j, ny + 1)];
    so[q] = solid[(size_t)(ii - 1) * ny + (jj - 1)];
    hxv[q] = hx[ii - 1];
    hyv[q] = hy[jj - 1];
  }
#pragma unroll
  for (int q = 0; q < kDiagTiles; ++q) {
    if (!has_s) so[q] = 0;
    if (!has_hx) hxv[q] = dx;
    if (!has_hy) hyv[q] = dy;
  }
  double r[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < kDiagTiles; ++q) {
    if (!in[q]) continue;
    const int i = ci[q], j = cj[q];
    if (i <= nx + 2 && j <= ny + 1) {
      const double a = au[q];
      if (!isfinite(a)) r[2] += 1.0;
      if (i >= 1 && i <= nx + 1 && j >= 1 && j <= ny) r[0] = fmax(r[0], fabs(a));
      if (has_prev) r[3] = fmax(r[3], fabs(a - aup[q]));
    }
    if (i <= nx + 1 && j <= ny + 2) {
      const double a = av[q];
      if (!isfinite(a)) r[2] += 1.0;
      if (i >= 1 && i <= nx && j >= 1 && j <= ny + 1) r[1] = fmax(r[1], fabs(a));
      if (has_prev) r[4] = fmax(r[4], fabs(a - avp[q]));
    }
    if (has_p && i <= nx + 1 && j <= ny + 1) {
      if (!isfinite(ap[q])) r[2] += 1.0;
    }
    if (i >= 1 && i <= nx && j >= 1 && j <= ny) {
      if (!so[q]) {
        const double d = (ub[q] - ua[q]) / hxv[q] + (vb[q] - va[q]) / hyv[q];
        r[5] = fmax(r[5], fabs(d));
      }
      const double uc = 0.5 * (ua[q] + ub[q]), vc = 0.5 * (va[q] + vb[q]);
      double ke = uc * uc + vc * vc;
      if (has_hx) ke = ke * (hxv[q] * hyv[q]);
      r[6] += ke;
    }
  }
  const int ops[7] = {kMax, kMax, kSum, kMax, kMax, kMax, kSum};
  block_partials<7>(r, ops, partial, blk);
}

struct EpiDiag {
  static constexpr int K = 7;
  double* out;
  __device__ void ops(int (&o)[K]) const {
    o[0] = kMax; o[1] = kMax; o[2] = kSum; o[3] = kMax; o[4] = kMax; o[5] = kMax; o[6] = kSum;
  }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&v)[K]) const {
    for (int q = 0; q < K; ++q) out[q] = v[q];
  }
};

// ----------------------------------------------------------------------------
// Host launchers.
// ----------------------------------------------------------------------------
static inline dim3 grid_for(int nxr, int nyr) { return dim3(num_tiles(nxr, nyr)); }

int launch_eddy_viscosity(hipStream_t s, const LesParams& P, const double* u0, const double* ue,
                          const double* v0, const double* ve, double ca, double cb, double* nuc,
                          double* nu_corner) {
  const dim3 g = grid_for(P.nx + 2, P.ny + 2);
  const bool st = P.g.hx != nullptr;
  if (ue) {
    const CombLoad U{u0, ue, P.ldu, ca, cb}, V{v0, ve, P.ldv, ca, cb};
    if (st)
      hipLaunchKernelGGL((k_eddy_viscosity<true, CombLoad, CombLoad>), g, dim3(kBlock), 0, s, U,
                         V, P, nuc, nu_corner);
    else
      hipLaunchKernelGGL((k_eddy_viscosity<false, CombLoad, CombLoad>), g, dim3(kBlock), 0, s, U,
                         V, P, nuc, nu_corner);
  } else {
    const PlainLoad U{u0, P.ldu}, V{v0, P.ldv};
    if (st)
      hipLaunchKernelGGL((k_eddy_viscosity<true, PlainLoad, PlainLoad>), g, dim3(kBlock), 0, s,
                         U, V, P, nuc, nu_corner);
    else
      hipLaunchKernelGGL((k_eddy_viscosity<false, PlainLoad, PlainLoad>), g, dim3(kBlock), 0, s,
                         U, V, P, nuc, nu_corner);
  }
  PCFD_LAUNCH_CHECK();
  if (nu_corner) {
    hipLaunchKernelGGL(k_corner_viscosity, grid_for(P.nx + 1, P.ny + 1), dim3(kBlock), 0, s, nuc,
                       P.nx, P.ny, nu_corner);
    PCFD_LAUNCH_CHECK();
  }
  return kOk;
}

int launch_momentum_rhs(hipStream_t s, const double* u, const double* v, const MomParams& P,
                        double* rhs_u, double* rhs_v) {
  const dim3 g = grid_for(P.nx + 1, P.ny + 1);
  if (P.g.hx)
    hipLaunchKernelGGL((k_momentum_rhs<true, false>), g, dim3(kBlock), 0, s, u, v, P, rhs_u,
                       rhs_v);
  else if (P.pow2)
    hipLaunchKernelGGL((k_momentum_rhs<false, true>), g, dim3(kBlock), 0, s, u, v, P, rhs_u,
                       rhs_v);
  else
    hipLaunchKernelGGL((k_momentum_rhs<false, false>), g, dim3(kBlock), 0, s, u, v, P, rhs_u,
                       rhs_v);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int launch_predictor(hipStream_t s, const double* u, const double* v, const MomParams& P,
                     const UpdRange& R, const StepScalars& T, double* us, double* vs,
                     const LaunchClock& clk) {
  const dim3 g = grid_for(P.nx + 3, P.ny + 3);
  const dim3 gr = grid_rows(P.nx + 3, P.ny + 3, kPredRows);
  const PlainLoad U{u, P.ldu}, V{v, P.ldv};
  if (P.g.hx)
    PCFD_LAUNCH_TIMED(clk, (k_predictor<true, false>), g, dim3(kBlock), 0, s, u, v, P, R, T, us,
                      vs);
  else if (P.nuc)
    PCFD_LAUNCH_TIMED(clk, (k_predictor_rows<false, true, kPredRows, PlainLoad>), gr,
                      dim3(kBlock), 0, s, U, V, P, R, T, us, vs);
  else if (P.pow2)
    PCFD_LAUNCH_TIMED(clk, (k_predictor_rows<true, false, kPredRows, PlainLoad>), gr,
                      dim3(kBlock), 0, s, U, V, P, R, T, us, vs);
  else
    PCFD_LAUNCH_TIMED(clk, (k_predictor_rows<false, false, kPredRows, PlainLoad>), gr,
                      dim3(kBlock), 0, s, U, V, P, R, T, us, vs);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int launch_predictor_comb(hipStream_t s, const double* u0, const double* ue, const double* v0,
                          const double* ve, double ca, double cb, const MomParams& P,
                          const UpdRange& R, const StepScalars& T, double* us, double* vs,
                          const LaunchClock& clk) {
  const dim3 g = grid_for(P.nx + 3, P.ny + 3);
  const dim3 gr = grid_rows(P.nx + 3, P.ny + 3, kPredRowsComb);
  const CombLoad U{u0, ue, P.ldu, ca, cb}, V{v0, ve, P.ldv, ca, cb};
  if (P.g.hx)
    PCFD_LAUNCH_TIMED(clk, (k_predictor_comb<true, false>), g, dim3(kBlock), 0, s, u0, ue, v0, ve,
                      ca, cb, P, R, T, us, vs);
  else if (P.nuc)
    PCFD_LAUNCH_TIMED(clk, (k_predictor_rows<false, true, kPredRowsComb, CombLoad>), gr,
                      dim3(kBlock), 0, s, U, V, P, R, T, us, vs);
  else if (P.pow2)
    PCFD_LAUNCH_TIMED(clk, (k_predictor_rows<true, false, kPredRowsComb, CombLoad>), gr,
                      dim3(kBlock), 0, s, U, V, P, R, T, us, vs);
  else
    PCFD_LAUNCH_TIMED(clk, (k_predictor_rows<false, false, kPredRowsComb, CombLoad>), gr,
                      dim3(kBlock), 0, s, U, V, P, R, T, us, vs);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int launch_combine(hipStream_t s, const double* u0, const double* ue, const double* v0,
                   const double* ve, double ca, double cb, int nx, int ny, int ldu, int ldv,
                   double* uo, double* vo, const LaunchClock& clk) {
  PCFD_LAUNCH_TIMED(clk, k_combine, grid_for(nx + 3, ny + 3), dim3(kBlock), 0, s, u0, ue, v0, ve,
                    ca, cb, nx, ny, ldu, ldv, uo, vo);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int launch_correct(hipStream_t s, const double* p, int ldp, int nx, int ny, int ldu, int ldv,
                   double dx, double dy, const Metrics& g, const UpdRange& R,
                   const StepScalars& T, double* u, double* v) {
  hipLaunchKernelGGL(k_correct, grid_for(nx + 3, ny + 3), dim3(kBlock), 0, s, p, ldp, nx, ny, ldu,
                     ldv, dx, dy, g, R, T, u, v);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int launch_correct_ghosts(hipStream_t s, double* p, int ldp, const BCParams& B, int ldu, int ldv,
                          double dx, double dy, const Metrics& g, const UpdRange& R,
                          const StepScalars& T, double* u, double* v, const LaunchClock& clk) {
  PGhost G{B.nx, B.ny, ldp, {}, {}, B.periodic_x, B.periodic_y};
  for (int w = 0; w < 4; ++w) {
    G.kind[w] = B.kind[w];
    G.pref[w] = B.pref[w];
  }
  const bool p2 = is_pow2(dx) && is_pow2(dy);
  const dim3 grid = grid_for(B.nx + 3, B.ny + 3);
  if (p2)
    PCFD_LAUNCH_TIMED(clk, k_correct_ghosts<true>, grid, dim3(kBlock), 0, s, p, G, ldu, ldv, dx,
                      dy, 1.0 / dx, 1.0 / dy, g, R, T, u, v);
  else
    PCFD_LAUNCH_TIMED(clk, k_correct_ghosts<false>, grid, dim3(kBlock), 0, s, p, G, ldu, ldv, dx,
                      dy, 1.0 / dx, 1.0 / dy, g, R, T, u, v);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int launch_divergence(hipStream_t s, const double* u, const double* v, int nx, int ny, int ldu,
                      int ldv, double dx, double dy, const Metrics& g, double* out) {
  hipLaunchKernelGGL(k_divergence, grid_for(nx, ny), dim3(kBlock), 0, s, u, v, nx, ny, ldu, ldv,
                     dx, dy, g, out);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

size_t diagnostics_partials(int nx, int ny) { return (size_t)diag_blocks(nx, ny) * 7; }

int launch_diagnostics(hipStream_t s, const double* u, const double* v, const double* p,
                       const double* up, const double* vp, const uint8_t* solid, int nx, int ny,
                       int ldu, int ldv, int ldp, double dx, double dy, const Metrics& g,
                       double* partial, unsigned int* counter, double* out,
                       const LaunchClock& clk) {
  (void)counter;
  const int nb = diag_blocks(nx, ny);
  // Stand-ins for absent inputs: u's allocation ((nx+3) x ldu doubles,
  // ldu >= ny+2) holds every index the kernel forms for p (rows <= nx+1,
  // columns <= ny+1, row pitch ldu), the solid flags (nx*ny bytes), hx (nx)
  // and hy (ny); v's the same for v_prev.  Their values are never used.
  const bool prev = up && vp;
  const int have = (p ? kHaveP : 0) | (prev ? kHavePrev : 0) | (solid ? kHaveSolid : 0) |
                   (g.hx ? kHaveHx : 0) | (g.hy ? kHaveHy : 0);
  PCFD_LAUNCH_TIMED(clk, k_diagnostics, dim3(nb), dim3(kBlock), 0, s, u, v, p ? p : u,
                    prev ? up : u, prev ? vp : v, solid ? solid : (const uint8_t*)u,
                    g.hx ? g.hx : u, g.hy ? g.hy : u, have, nx, ny, ldu, ldv, p ? ldp : ldu, dx,
                    dy, partial);
  PCFD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fold<EpiDiag>, dim3(1), dim3(kBlock), 0, s, (const double*)partial, nb,
                     EpiDiag{out});
  PCFD_LAUNCH_CHECK();
  return kOk;
}

}  // namespace pcfd
