// Device helpers of the matrix-free fine Poisson operator, shared by the
// solver kernels and the fused multigrid kernels.
#pragma once

#include "poisson.hpp"

namespace pcfd {

// ----------------------------------------------------------------------------
// Fine operator helpers
// ----------------------------------------------------------------------------

// Coefficients of the four faces of cell (i, j): 1/dx^2, 1/dy^2 on a uniform
// mesh; on a stretched one the reference's face area / centre distance / mean
// cell area (pressure.py:193-207), with its operation order, so the operator
// is the reference's matrix bit for bit.  A face beyond a non-periodic wall
// takes the wall face's coefficient (the Dirichlet terms use it).
struct FaceC {
  double l, r, b, t;
};

__device__ __forceinline__ FaceC fine_faces(const FineOp& A, int i, int j) {
  if (!A.hxc) return {A.cx, A.cx, A.cy, A.cy};
  const double hy = A.hyc[j], hx = A.hxc[i];
  return {hy / A.hxv[i], hy / A.hxv[i + 1], hx / A.hyv[j], hx / A.hyv[j + 1]};
}

// Row scale of the right-hand side (pressure.py:207, 300-308): 1 on a
// uniform mesh, cell area / mean cell area on a stretched one.
__device__ __forceinline__ double fine_row_scale(const FineOp& A, int i, int j) {
  return A.hxc ? (A.hxc[i] * A.hyc[j]) / A.vbar : 1.0;
}

__device__ __forceinline__ double fine_diag(const FineOp& A, const FaceC& c, uint8_t f, int i,
                                            int j) {
  if (!(f & F_SOLV)) return 1.0;
  // pressure.py:219-238: one -coeff per active face, -2*coeff per Dirichlet
  // wall face, accumulated left, right, bottom, top.
  double d = 0.0;
  if (f & F_L) d -= c.l;
  if (A.dir[0] && i == 0) d -= 2.0 * c.l;
  if (f & F_R) d -= c.r;
  if (A.dir[1] && i == A.nx - 1) d -= 2.0 * c.r;
  if (f & F_B) d -= c.b;
  if (A.dir[2] && j == 0) d -= 2.0 * c.b;
  if (f & F_T) d -= c.t;
  if (A.dir[3] && j == A.ny - 1) d -= 2.0 * c.t;
  return d;
}

__device__ __forceinline__ double fine_diag(const FineOp& A, uint8_t f, int i, int j) {
  return fine_diag(A, fine_faces(A, i, j), f, i, j);
}

struct Nb {
  int kL, kR, kB, kT;
};

__device__ __forceinline__ Nb fine_neighbours(const FineOp& A, int i, int j) {
  const int ny = A.ny, nx = A.nx, k = i * ny + j;
  Nb n;
  n.kL = (i > 0) ? k - ny : (A.px ? k + (nx - 1) * ny : k);
  n.kR = (i < nx - 1) ? k + ny : (A.px ? k - (nx - 1) * ny : k);
  n.kB = (j > 0) ? k - 1 : (A.py ? k + ny - 1 : k);
  n.kT = (j < ny - 1) ? k + 1 : (A.py ? k - (ny - 1) : k);
  return n;
}

// y = (A x)[k], terms in CSR column order for a non-wrapping row.
template <typename T, class X>
__device__ __forceinline__ T fine_apply(const FaceC& c, uint8_t f, const Nb& nb, int k, T d,
                                        const X& x) {
  T y = (T)0;
  if (f & F_L) y += (T)c.l * x(nb.kL);
  if (f & F_B) y += (T)c.b * x(nb.kB);
  y += d * x(k);
  if (f & F_T) y += (T)c.t * x(nb.kT);
  if (f & F_R) y += (T)c.r * x(nb.kR);
  return y;
}

__device__ __forceinline__ bool cell_of_tile(int n_x, int n_y, Tile2D& t) {
  t = tile_coords(n_y);
  return t.i < n_x && t.j < n_y;
}

__device__ __forceinline__ double dirichlet_rhs(const FineOp& A, const FaceC& c, int i, int j) {
  // pressure.py:233-238: bc -= 2*coeff*p_ref on the faces of a Dirichlet wall.
  double bc = 0.0;
  if (A.dir[0] && i == 0) bc -= 2.0 * c.l * A.pref[0];
  if (A.dir[1] && i == A.nx - 1) bc -= 2.0 * c.r * A.pref[1];
  if (A.dir[2] && j == 0) bc -= 2.0 * c.b * A.pref[2];
  if (A.dir[3] && j == A.ny - 1) bc -= 2.0 * c.t * A.pref[3];
  return bc;
}

}  // namespace pcfd
