// Device helpers of the matrix-free fine Poisson operator, shared by the
// solver kernels and the fused multigrid kernels.
#pragma once

#include "poisson.hpp"

namespace pcfd {

// ----------------------------------------------------------------------------
// Fine operator helpers
// ----------------------------------------------------------------------------
__device__ __forceinline__ double fine_diag(const FineOp& A, uint8_t f, int i, int j) {
  if (!(f & F_SOLV)) return 1.0;
  // pressure.py:219-238: one -coeff per active face, -2*coeff per Dirichlet
  // wall face, accumulated left, right, bottom, top.
  double d = 0.0;
  if (f & F_L) d -= A.cx;
  if (A.dir[0] && i == 0) d -= 2.0 * A.cx;
  if (f & F_R) d -= A.cx;
  if (A.dir[1] && i == A.nx - 1) d -= 2.0 * A.cx;
  if (f & F_B) d -= A.cy;
  if (A.dir[2] && j == 0) d -= 2.0 * A.cy;
  if (f & F_T) d -= A.cy;
  if (A.dir[3] && j == A.ny - 1) d -= 2.0 * A.cy;
  return d;
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
__device__ __forceinline__ T fine_apply(const FineOp& A, uint8_t f, const Nb& nb, int k, T d,
                                        const X& x) {
  T y = (T)0;
  if (f & F_L) y += (T)A.cx * x(nb.kL);
  if (f & F_B) y += (T)A.cy * x(nb.kB);
  y += d * x(k);
  if (f & F_T) y += (T)A.cy * x(nb.kT);
  if (f & F_R) y += (T)A.cx * x(nb.kR);
  return y;
}

__device__ __forceinline__ bool cell_of_tile(int n_x, int n_y, Tile2D& t) {
  t = tile_coords(n_y);
  return t.i < n_x && t.j < n_y;
}

__device__ __forceinline__ double dirichlet_rhs(const FineOp& A, int i, int j) {
  // pressure.py:233-238: bc -= 2*coeff*p_ref on the faces of a Dirichlet wall.
  double bc = 0.0;
  if (A.dir[0] && i == 0) bc -= 2.0 * A.cx * A.pref[0];
  if (A.dir[1] && i == A.nx - 1) bc -= 2.0 * A.cx * A.pref[1];
  if (A.dir[2] && j == 0) bc -= 2.0 * A.cy * A.pref[2];
  if (A.dir[3] && j == A.ny - 1) bc -= 2.0 * A.cy * A.pref[3];
  return bc;
}

}  // namespace pcfd
