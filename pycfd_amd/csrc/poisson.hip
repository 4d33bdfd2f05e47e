// Pressure-Poisson solvers on the GPU.
//
// The operator is the reference's 5-point div(grad) over exactly the
// correctable faces (pycfd/core/pressure.py:110-270), applied matrix-free
// from one flag byte per cell.  Solver semantics follow the reference:
//   CG    : scipy.sparse.linalg.cg with M = x/diag (pressure.py:366-398);
//           stopping test ||r|| < tol*||b|| on the recursive residual.
//   Jacobi: damped sweeps, mean removal, residual every 20 (pressure.py:410-454)
//   SOR   : red/black sweeps, mean removal, residual every 5 (pressure.py:457-526)
//   MG    : V-cycles until the residual meets tol, with the stall guard
//           (multigrid.py:481-511)
//   MGCG  : CG preconditioned by one V-cycle (pressure.py:577-632)
// Constant-mode removal ("subtract the fluid mean") is carried as a pending
// scalar shift that the next kernel applies on the fly; for the singular
// Neumann operator A*1 = 0 on fluid rows, so this is the same iteration.
#include "poisson.hpp"

#include <algorithm>
#include <cstring>

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

// ----------------------------------------------------------------------------
// Right-hand side assembly (solver.py:659-661, pressure.py:300-308)
// ----------------------------------------------------------------------------
__device__ __forceinline__ double dirichlet_rhs(const FineOp& A, int i, int j) {
  // pressure.py:233-238: bc -= 2*coeff*p_ref on the faces of a Dirichlet wall.
  double bc = 0.0;
  if (A.dir[0] && i == 0) bc -= 2.0 * A.cx * A.pref[0];
  if (A.dir[1] && i == A.nx - 1) bc -= 2.0 * A.cx * A.pref[1];
  if (A.dir[2] && j == 0) bc -= 2.0 * A.cy * A.pref[2];
  if (A.dir[3] && j == A.ny - 1) bc -= 2.0 * A.cy * A.pref[3];
  return bc;
}

__global__ void __launch_bounds__(kBlock)
k_prepare(FineOp A, const double* __restrict__ u, const double* __restrict__ v, int ldu, int ldv,
          double dx, double dy, double dt, const double* __restrict__ rhs, double* __restrict__ b,
          double* partial, unsigned int* counter, SolveState* st, int singular, int nfluid) {
  Tile2D t;
  double s[1] = {0.0};
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int i = t.i, j = t.j, k = i * A.ny + j;
    const uint8_t f = A.flags[k];
    double val;
    if (rhs) {
      val = rhs[k];
    } else {
      const int a = i + 1, c = j + 1;
      const double div = (u[(size_t)(a + 1) * ldu + c] - u[(size_t)a * ldu + c]) / dx +
                         (v[(size_t)a * ldv + c + 1] - v[(size_t)a * ldv + c]) / dy;
      val = div / dt;
    }
    val = val * 1.0 + ((f & F_SOLV) ? dirichlet_rhs(A, i, j) : 0.0);
    if (!(f & F_SOLV)) val = 0.0;
    b[k] = val;
    s[0] = val;
  }
  const int ops[1] = {kSum};
  if (grid_reduce<1>(s, ops, partial, counter, t.tile, gridDim.x) && threadIdx.x == 0) {
    st->mean_b = (singular && nfluid > 0) ? s[0] / (double)nfluid : 0.0;
  }
}

// Solver start: b <- b - mean_b on fluid rows; x <- guess (non-fluid rows
// zeroed) or 0; r = b - A x.  Also ||b||^2, ||r||^2 and, per mode, r.(r/d)
// (Jacobi-CG) or sum(r) (multigrid-CG input projection).
enum InitMode : int { IM_JPCG = 0, IM_MGCG = 1, IM_RELAX = 2, IM_MG = 3 };

__global__ void __launch_bounds__(kBlock)
k_init(FineOp A, double* __restrict__ b, double* __restrict__ x, double* __restrict__ r,
       int zero_guess, int mode, double tol, int maxiter, int n_active, double* partial,
       unsigned int* counter, SolveState* st) {
  Tile2D t;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  const double mb = st->mean_b;
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int i = t.i, j = t.j, k = i * A.ny + j;
    const uint8_t f = A.flags[k];
    const Nb nb = fine_neighbours(A, i, j);
    const double bk = (f & F_SOLV) ? b[k] - mb : 0.0;
    auto xv = [&](int q) -> double {
      if (zero_guess) return 0.0;
      return (A.flags[q] & F_SOLV) ? x[q] : 0.0;
    };
    const double d = fine_diag(A, f, i, j);
    const double rk = bk - fine_apply<double>(A, f, nb, k, d, xv);
    s[0] = bk * bk;
    s[1] = rk * rk;
    if (mode == IM_JPCG) s[2] = rk * (rk / d);
    if (mode == IM_MGCG && (f & F_SOLV)) s[2] = rk;
    // every thread has read its neighbours' x before any write below is
    // visible to them only if writes go to other buffers: b and r are
    // separate from x, and x is written only where it changes (zeroing).
    b[k] = bk;
    if (r) r[k] = rk;
  }
  const int ops[4] = {kSum, kSum, kSum, kSum};
  if (grid_reduce<4>(s, ops, partial, counter, t.tile, gridDim.x) && threadIdx.x == 0) {
    const double bn = sqrt(s[0]);
    st->bnorm = bn;
    st->zero_rhs = (mode == IM_JPCG || mode == IM_MGCG) && (bn == 0.0);
    st->rr = s[1];
    st->iter = 0;
    st->stalled = 0;
    st->mean_x = 0.0;
    st->atol = tol * bn;
    const double res = sqrt(s[1]) / (bn > 1.0e-300 ? bn : 1.0);
    st->res = res;
    st->res_prev = res;
    int done = 0, conv = 0;
    if (mode == IM_JPCG || mode == IM_MGCG) {
      st->rho = s[2];
      st->mean_r = (n_active > 0) ? s[2] / (double)n_active : 0.0;
      if (bn == 0.0) {
        done = 1;
        conv = 1;
      } else if (sqrt(s[1]) < st->atol) {
        done = 1;
        conv = 1;
      } else if (maxiter <= 0) {
        done = 1;
      }
    } else if (mode == IM_MG) {
      if (!(res >= tol)) {
        done = 1;
        conv = 1;
      } else if (maxiter <= 0) {
        done = 1;
      }
    } else {
      if (maxiter <= 0) done = 1;
    }
    st->done = done;
    st->converged = conv;
  }
}

// Non-fluid rows of the warm-start guess are zeroed in a separate pass so
// that k_init never writes a vector its neighbours read.
__global__ void __launch_bounds__(kBlock)
k_zero_nonfluid(FineOp A, double* __restrict__ x) {
  Tile2D t;
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int k = t.i * A.ny + t.j;
    if (!(A.flags[k] & F_SOLV)) x[k] = 0.0;
  }
}

// ----------------------------------------------------------------------------
// Jacobi-preconditioned CG (scipy.sparse.linalg.cg iteration)
// ----------------------------------------------------------------------------
// p = z + beta*p_old (z = r/d), q = A p, and p.q.
__global__ void __launch_bounds__(kBlock)
k_jpcg_dir(FineOp A, const double* __restrict__ r, const double* __restrict__ p_old,
           double* __restrict__ p_new, double* __restrict__ q, double* partial,
           unsigned int* counter, SolveState* st) {
  if (st->done) return;
  Tile2D t;
  double s[1] = {0.0};
  const bool first = st->iter == 0;
  const double beta = st->beta;
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int i = t.i, j = t.j, k = i * A.ny + j;
    const uint8_t f = A.flags[k];
    const Nb nb = fine_neighbours(A, i, j);
    auto pv = [&](int m) -> double {
      const int mi = m / A.ny, mj = m - mi * A.ny;
      const double z = r[m] / fine_diag(A, A.flags[m], mi, mj);
      return first ? z : p_old[m] * beta + z;
    };
    const double d = fine_diag(A, f, i, j);
    const double pk = pv(k);
    const double qk = fine_apply<double>(A, f, nb, k, d, pv);
    p_new[k] = pk;
    q[k] = qk;
    s[0] = pk * qk;
  }
  const int ops[1] = {kSum};
  if (grid_reduce<1>(s, ops, partial, counter, t.tile, gridDim.x) && threadIdx.x == 0) {
    st->alpha = st->rho / s[0];
  }
}

// x += alpha p, r -= alpha q; then ||r||^2 and the next rho.
__global__ void __launch_bounds__(kBlock)
k_cg_update(FineOp A, double* __restrict__ x, double* __restrict__ r,
            const double* __restrict__ p, const double* __restrict__ q, int mode, int maxiter,
            int n_active, double* partial, unsigned int* counter, SolveState* st) {
  if (st->done) return;
  Tile2D t;
  double s[2] = {0.0, 0.0};
  const double alpha = st->alpha;
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int i = t.i, j = t.j, k = i * A.ny + j;
    const uint8_t f = A.flags[k];
    const double xk = x[k] + alpha * p[k];
    const double rk = r[k] - alpha * q[k];
    x[k] = xk;
    r[k] = rk;
    s[0] = rk * rk;
    if (mode == IM_JPCG) s[1] = rk * (rk / fine_diag(A, f, i, j));
    else if (f & F_SOLV) s[1] = rk;
  }
  const int ops[2] = {kSum, kSum};
  if (grid_reduce<2>(s, ops, partial, counter, t.tile, gridDim.x) && threadIdx.x == 0) {
    const int it = st->iter + 1;
    st->iter = it;
    st->rr = s[0];
    if (mode == IM_JPCG) {
      st->rho_prev = st->rho;
      st->rho = s[1];
      st->beta = st->rho / st->rho_prev;
    } else {
      st->mean_r = (n_active > 0) ? s[1] / (double)n_active : 0.0;
    }
    if (sqrt(s[0]) < st->atol) {
      st->done = 1;
      st->converged = 1;
    } else if (it >= maxiter) {
      st->done = 1;
    }
  }
}

// Multigrid-CG direction: p = zp + beta*p_old with zp the projected V-cycle
// output (z - mean_z on the active set), q = A p, p.q.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_mgcg_dir(FineOp A, const T* __restrict__ z, const double* __restrict__ p_old,
           double* __restrict__ p_new, double* __restrict__ q, double* partial,
           unsigned int* counter, SolveState* st) {
  if (st->done) return;
  Tile2D t;
  double s[1] = {0.0};
  const bool first = st->iter == 0;
  const double beta = st->beta;
  const double mz = st->mean_z;
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int i = t.i, j = t.j, k = i * A.ny + j;
    const uint8_t f = A.flags[k];
    const Nb nb = fine_neighbours(A, i, j);
    auto pv = [&](int m) -> double {
      const double zp = (A.flags[m] & F_SOLV) ? (double)z[m] - mz : 0.0;
      return first ? zp : p_old[m] * beta + zp;
    };
    const double d = fine_diag(A, f, i, j);
    const double pk = pv(k);
    const double qk = fine_apply<double>(A, f, nb, k, d, pv);
    p_new[k] = pk;
    q[k] = qk;
    s[0] = pk * qk;
  }
  const int ops[1] = {kSum};
  if (grid_reduce<1>(s, ops, partial, counter, t.tile, gridDim.x) && threadIdx.x == 0) {
    st->alpha = st->rho / s[0];
  }
}

// V-cycle input: b0 = r - mean(r) on the active set, 0 elsewhere
// (MultigridHierarchy._project, multigrid.py:453-458).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_project_in(FineOp A, const double* __restrict__ r, T* __restrict__ b0, SolveState* st) {
  if (st->done) return;
  Tile2D t;
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int k = t.i * A.ny + t.j;
    b0[k] = (A.flags[k] & F_SOLV) ? (T)(r[k] - st->mean_r) : (T)0;
  }
}

// ----------------------------------------------------------------------------
// Relaxation solvers
// ----------------------------------------------------------------------------
// One colour of a red/black SOR sweep (pressure.py:509-513).  Values still
// carrying the pending mean shift are corrected on read.
__global__ void __launch_bounds__(kBlock)
k_sor_colour(FineOp A, double* __restrict__ p, const double* __restrict__ b, int colour,
             double omega, int nfluid, int singular, double* partial, unsigned int* counter,
             SolveState* st) {
  if (st->done) return;
  Tile2D t;
  double s[1] = {0.0};
  const double m = st->mean_x;
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int i = t.i, j = t.j, k = i * A.ny + j;
    if (((i + j) & 1) == colour) {
      const uint8_t f = A.flags[k];
      const Nb nb = fine_neighbours(A, i, j);
      // red reads both colours shifted; black reads the fresh red unshifted
      auto pv = [&](int q) -> double {
        const int qi = q / A.ny, qj = q - qi * A.ny;
        const bool same = (((qi + qj) & 1) == colour);
        const double sh = (colour == 0 || same) ? m : 0.0;
        return (A.flags[q] & F_SOLV) ? p[q] - sh : p[q];
      };
      const double d = fine_diag(A, f, i, j);
      // off-diagonal part only: A_rb p_b = (A p) - d p
      double off = 0.0;
      if (f & F_L) off += A.cx * pv(nb.kL);
      if (f & F_B) off += A.cy * pv(nb.kB);
      if (f & F_T) off += A.cy * pv(nb.kT);
      if (f & F_R) off += A.cx * pv(nb.kR);
      const double pk = pv(k);
      const double val = (1.0 - omega) * pk + omega * (b[k] - off) / d;
      p[k] = val;
      if (f & F_SOLV) s[0] = val;
    }
  }
  const int ops[1] = {kSum};
  if (grid_reduce<1>(s, ops, partial, counter, t.tile, gridDim.x) && threadIdx.x == 0) {
    if (colour == 0) {
      st->sum_red = s[0];
    } else {
      st->iter = st->iter + 1;
      st->mean_x = (singular && nfluid > 0) ? (st->sum_red + s[0]) / (double)nfluid : 0.0;
    }
  }
}

// Damped Jacobi sweep p_out = p + omega*dinv*(b - A p) (pressure.py:439-441).
__global__ void __launch_bounds__(kBlock)
k_jacobi_sweep(FineOp A, const double* __restrict__ p, double* __restrict__ pn,
               const double* __restrict__ b, double omega, int nfluid, int singular,
               double* partial, unsigned int* counter, SolveState* st) {
  if (st->done) return;
  Tile2D t;
  double s[1] = {0.0};
  const double m = st->mean_x;
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int i = t.i, j = t.j, k = i * A.ny + j;
    const uint8_t f = A.flags[k];
    const Nb nb = fine_neighbours(A, i, j);
    auto pv = [&](int q) -> double { return (A.flags[q] & F_SOLV) ? p[q] - m : p[q]; };
    const double d = fine_diag(A, f, i, j);
    const double ap = fine_apply<double>(A, f, nb, k, d, pv);
    const double val = pv(k) + omega * (1.0 / d) * (b[k] - ap);
    pn[k] = val;
    if (f & F_SOLV) s[0] = val;
  }
  const int ops[1] = {kSum};
  if (grid_reduce<1>(s, ops, partial, counter, t.tile, gridDim.x) && threadIdx.x == 0) {
    st->iter = st->iter + 1;
    st->mean_x = (singular && nfluid > 0) ? s[0] / (double)nfluid : 0.0;
  }
}

// Relative residual of (x - pending shift) against b; decides termination.
enum CheckMode : int { CM_RELAX = 0, CM_MG = 1, CM_FINAL = 2 };

__global__ void __launch_bounds__(kBlock)
k_check(FineOp A, const double* __restrict__ x, const double* __restrict__ b, int mode,
        double tol, int maxiter, int nfluid, int singular, double* partial,
        unsigned int* counter, SolveState* st) {
  if (mode != CM_FINAL && st->done) return;
  Tile2D t;
  double s[2] = {0.0, 0.0};
  const double m = st->mean_x;
  const bool zero = (mode == CM_FINAL) && st->zero_rhs;
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int i = t.i, j = t.j, k = i * A.ny + j;
    const uint8_t f = A.flags[k];
    const Nb nb = fine_neighbours(A, i, j);
    auto xv = [&](int q) -> double {
      if (zero) return 0.0;
      return (A.flags[q] & F_SOLV) ? x[q] - m : x[q];
    };
    const double d = fine_diag(A, f, i, j);
    const double rk = b[k] - fine_apply<double>(A, f, nb, k, d, xv);
    s[0] = rk * rk;
    if (f & F_SOLV) s[1] = xv(k);
  }
  const int ops[2] = {kSum, kSum};
  if (grid_reduce<2>(s, ops, partial, counter, t.tile, gridDim.x) && threadIdx.x == 0) {
    const double bn = st->bnorm;
    const double res = sqrt(s[0]) / (bn > 1.0e-300 ? bn : 1.0);
    if (mode == CM_FINAL) {
      st->res_final = res;
      st->mean_final = (singular && nfluid > 0) ? s[1] / (double)nfluid : 0.0;
      return;
    }
    const double prev = st->res;
    st->res_prev = prev;
    st->res = res;
    if (mode == CM_RELAX) {
      if (res < tol) {
        st->done = 1;
        st->converged = 1;
      } else if (st->iter >= maxiter) {
        st->done = 1;
      }
    } else {  // multigrid.py:504-510
      const int cycles = st->iter + 1;
      st->iter = cycles;
      st->stalled = (res > 0.995 * prev) ? st->stalled + 1 : 0;
      if (!(res >= tol)) {
        st->done = 1;
        st->converged = 1;
      } else if (cycles >= maxiter || st->stalled >= 3) {
        st->done = 1;
      }
    }
  }
}

// Normalised pressure into the ghosted array (solver.py:659-665 with
// pressure.py:310-316): fluid rows minus the fluid mean, others zero, ghost
// ring zero.  The same values become the next warm-start guess.
__global__ void __launch_bounds__(kBlock)
k_scatter(FineOp A, double* __restrict__ x, const SolveState* st, double* __restrict__ p,
          int ldp, double* __restrict__ out) {
  const Tile2D t = tile_coords(A.ny + 2);
  const int i = t.i, j = t.j;
  if (i > A.nx + 1 || j > A.ny + 1) return;
  const bool interior = i >= 1 && i <= A.nx && j >= 1 && j <= A.ny;
  double val = 0.0;
  if (interior) {
    const int k = (i - 1) * A.ny + (j - 1);
    if (A.flags[k] & F_SOLV) {
      const double xk = st->zero_rhs ? 0.0 : x[k] - st->mean_x;
      val = xk - st->mean_final;
    }
    x[k] = val;
    if (out) out[k] = val;
  }
  if (p) p[(size_t)i * ldp + j] = val;
}

// ----------------------------------------------------------------------------
// Multigrid V-cycle kernels.  Level 0 uses the matrix-free fine operator;
// coarser levels use their stored Galerkin stencils.  T is the storage and
// arithmetic precision of the cycle.
// ----------------------------------------------------------------------------
template <typename T>
struct FineLevelArgs {
  FineOp A;
  T weight;
};

template <typename T>
__device__ __forceinline__ T fine_dinv(const FineOp& A, int k) {
  const int i = k / A.ny, j = k - i * A.ny;
  return (T)(1.0 / fine_diag(A, A.flags[k], i, j));
}

// Neighbour index of offset s, or -1 outside a non-periodic grid.
template <typename T>
__device__ __forceinline__ int class_nb(const ClassLevel<T>& L, int i, int j, int s) {
  int ii = i + L.di[s], jj = j + L.dj[s];
  if (L.px) ii = (ii % L.nx + L.nx) % L.nx;
  else if (ii < 0 || ii >= L.nx) return -1;
  if (L.py) jj = (jj % L.ny + L.ny) % L.ny;
  else if (jj < 0 || jj >= L.ny) return -1;
  return ii * L.ny + jj;
}

// (A x)[k] for a coarse level; `xv(q)` supplies x at cell q.
template <typename T, class X>
__device__ __forceinline__ T class_apply(const ClassLevel<T>& L, int i, int j, int c,
                                         const X& xv) {
  const T* row = L.table + (size_t)c * L.S;
  T y = (T)0;
  for (int s = 0; s < L.S; ++s) {
    const T a = row[s];
    if (a == (T)0) continue;
    const int q = class_nb(L, i, j, s);
    if (q >= 0) y += a * xv(q);
  }
  return y;
}

// Damped Jacobi sweep (multigrid.py:424-428); x_in == nullptr means a zero
// start.  Level 0 may apply a pending mean shift to x_in, and the last
// sweep of a cycle may reduce sum(x), sum(b*x), sum(b) over the active set.
enum SmoothReduce : int { SR_NONE = 0, SR_MG = 1, SR_MGCG = 2 };

__device__ __forceinline__ void cycle_reduce(double (&s)[3], int reduce, int n_active,
                                             double* partial, unsigned int* counter,
                                             SolveState* st, int tile) {
  const int ops[3] = {kSum, kSum, kSum};
  if (grid_reduce<3>(s, ops, partial, counter, tile, gridDim.x) && threadIdx.x == 0) {
    const double mz = (n_active > 0) ? s[0] / (double)n_active : 0.0;
    if (reduce == SR_MG) {
      st->mean_x = mz;
    } else {
      // r.z_proj with z_proj = z - mean_z and the input b0 = r - mean_r:
      // sum(b0*z) - mean_z*sum(b0) is the same inner product.
      st->mean_z = mz;
      const double rho = s[1] - mz * s[2];
      st->rho_prev = st->rho;
      st->rho = rho;
      st->beta = rho / st->rho_prev;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_smooth_fine(FineLevelArgs<T> L, const T* __restrict__ b, const T* __restrict__ x_in,
              T* __restrict__ x_out, int shift_in, int reduce, int n_active, double* partial,
              unsigned int* counter, SolveState* st) {
  if (st->done) return;
  const FineOp& A = L.A;
  Tile2D t;
  double s[3] = {0.0, 0.0, 0.0};
  const T m = shift_in ? (T)st->mean_x : (T)0;
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int i = t.i, j = t.j, k = i * A.ny + j;
    const uint8_t f = A.flags[k];
    const T d = (T)fine_diag(A, f, i, j);
    const T wd = L.weight * ((T)1 / d);
    T val;
    if (x_in) {
      const Nb nb = fine_neighbours(A, i, j);
      auto xv = [&](int q) -> T { return (A.flags[q] & F_SOLV) ? x_in[q] - m : x_in[q]; };
      val = xv(k) + wd * (b[k] - fine_apply<T>(A, f, nb, k, d, xv));
    } else {
      val = wd * b[k];
    }
    x_out[k] = val;
    if (reduce != SR_NONE && (f & F_SOLV)) {
      s[0] = (double)val;
      s[1] = (double)b[k] * (double)val;
      s[2] = (double)b[k];
    }
  }
  if (reduce == SR_NONE) return;
  cycle_reduce(s, reduce, n_active, partial, counter, st, t.tile);
}

// Single-level hierarchy: the coarsest solve is the whole cycle, so the
// projection sums are taken in a pass of their own.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cycle_reduce(FineOp A, const T* __restrict__ b, const T* __restrict__ z, int reduce,
               int n_active, double* partial, unsigned int* counter, SolveState* st) {
  if (st->done) return;
  Tile2D t;
  double s[3] = {0.0, 0.0, 0.0};
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int k = t.i * A.ny + t.j;
    if (A.flags[k] & F_SOLV) {
      s[0] = (double)z[k];
      s[1] = (double)b[k] * (double)z[k];
      s[2] = (double)b[k];
    }
  }
  cycle_reduce(s, reduce, n_active, partial, counter, st, t.tile);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_smooth_cls(ClassLevel<T> L, T weight, const T* __restrict__ b, const T* __restrict__ x_in,
             T* __restrict__ x_out, const SolveState* st) {
  if (st->done) return;
  Tile2D t;
  if (!cell_of_tile(L.nx, L.ny, t)) return;
  const int k = t.i * L.ny + t.j;
  const int c = L.cls[k];
  const T wd = weight * L.cdinv[c];
  if (x_in) {
    auto xv = [&](int q) -> T { return x_in[q]; };
    x_out[k] = x_in[k] + wd * (b[k] - class_apply(L, t.i, t.j, c, xv));
  } else {
    x_out[k] = wd * b[k];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_residual_fine(FineOp A, const T* __restrict__ b, const T* __restrict__ x, T* __restrict__ r,
                const SolveState* st) {
  if (st->done) return;
  Tile2D t;
  if (!cell_of_tile(A.nx, A.ny, t)) return;
  const int i = t.i, j = t.j, k = i * A.ny + j;
  const uint8_t f = A.flags[k];
  const Nb nb = fine_neighbours(A, i, j);
  auto xv = [&](int q) -> T { return x[q]; };
  r[k] = b[k] - fine_apply<T>(A, f, nb, k, (T)fine_diag(A, f, i, j), xv);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_residual_cls(ClassLevel<T> L, const T* __restrict__ b, const T* __restrict__ x,
               T* __restrict__ r, const SolveState* st) {
  if (st->done) return;
  Tile2D t;
  if (!cell_of_tile(L.nx, L.ny, t)) return;
  const int k = t.i * L.ny + t.j;
  auto xv = [&](int q) -> T { return x[q]; };
  r[k] = b[k] - class_apply(L, t.i, t.j, (int)L.cls[k], xv);
}

// Restriction b_c = P^T r with P = (I - w D^-1 A) P_tent
// (multigrid.py:248-273): b_c[I] = sum over active a in block I of
// (r - w A D^-1 r)[a].  One thread per coarse cell.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_restrict_fine(FineLevelArgs<T> L, int fx, int fy, int cnx, int cny, const T* __restrict__ r,
                T* __restrict__ bc, const SolveState* st) {
  if (st->done) return;
  const FineOp& A = L.A;
  Tile2D t;
  if (!cell_of_tile(cnx, cny, t)) return;
  T acc = (T)0;
  for (int a = 0; a < fx; ++a) {
    const int i = t.i * fx + a;
    if (i >= A.nx) break;
    for (int c = 0; c < fy; ++c) {
      const int j = t.j * fy + c;
      if (j >= A.ny) break;
      const int k = i * A.ny + j;
      const uint8_t f = A.flags[k];
      if (!(f & F_SOLV)) continue;
      const Nb nb = fine_neighbours(A, i, j);
      auto tv = [&](int q) -> T { return r[q] * fine_dinv<T>(A, q); };
      const T atr = fine_apply<T>(A, f, nb, k, (T)fine_diag(A, f, i, j), tv);
      acc += r[k] - L.weight * atr;
    }
  }
  bc[t.i * cny + t.j] = acc;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_restrict_cls(ClassLevel<T> L, T weight, int fx, int fy, int cnx, int cny,
               const T* __restrict__ r, T* __restrict__ bc, const SolveState* st) {
  if (st->done) return;
  Tile2D t;
  if (!cell_of_tile(cnx, cny, t)) return;
  T acc = (T)0;
  auto tv = [&](int q) -> T { return r[q] * L.cdinv[L.cls[q]]; };
  for (int a = 0; a < fx; ++a) {
    const int i = t.i * fx + a;
    if (i >= L.nx) break;
    for (int c = 0; c < fy; ++c) {
      const int j = t.j * fy + c;
      if (j >= L.ny) break;
      const int k = i * L.ny + j;
      const int cl = L.cls[k];
      if (!L.cactive[cl]) continue;
      acc += r[k] - weight * class_apply(L, i, j, cl, tv);
    }
  }
  bc[t.i * cny + t.j] = acc;
}

// Prolongation x += P x_c = e - w D^-1 A e with e = P_tent x_c.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_prolong_fine(FineLevelArgs<T> L, int fx, int fy, int cny, const T* __restrict__ xc,
               T* __restrict__ x, const SolveState* st) {
  if (st->done) return;
  const FineOp& A = L.A;
  Tile2D t;
  if (!cell_of_tile(A.nx, A.ny, t)) return;
  const int i = t.i, j = t.j, k = i * A.ny + j;
  const uint8_t f = A.flags[k];
  const Nb nb = fine_neighbours(A, i, j);
  auto ev = [&](int q) -> T {
    if (!(A.flags[q] & F_SOLV)) return (T)0;
    const int qi = q / A.ny, qj = q - qi * A.ny;
    return xc[(qi / fx) * cny + (qj / fy)];
  };
  const T d = (T)fine_diag(A, f, i, j);
  const T ae = fine_apply<T>(A, f, nb, k, d, ev);
  x[k] += ev(k) - L.weight * ((T)1 / d) * ae;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_prolong_cls(ClassLevel<T> L, T weight, int fx, int fy, int cny, const T* __restrict__ xc,
              T* __restrict__ x, const SolveState* st) {
  if (st->done) return;
  Tile2D t;
  if (!cell_of_tile(L.nx, L.ny, t)) return;
  const int i = t.i, j = t.j, k = i * L.ny + j;
  auto ev = [&](int q) -> T {
    if (!L.cactive[L.cls[q]]) return (T)0;
    const int qi = q / L.ny, qj = q - qi * L.ny;
    return xc[(qi / fx) * cny + (qj / fy)];
  };
  const int c = L.cls[k];
  const T ae = class_apply(L, i, j, c, ev);
  x[k] += ev(k) - weight * L.cdinv[c] * ae;
}

// Coarsest level: x = C b with C the dense FP64 pseudo-inverse
// (multigrid.py:514-543).  One workgroup.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_coarse_solve(const double* __restrict__ C, int n, const T* __restrict__ b, T* __restrict__ x,
               const SolveState* st) {
  if (st->done) return;
  extern __shared__ double sb[];
  for (int k = threadIdx.x; k < n; k += blockDim.x) sb[k] = (double)b[k];
  __syncthreads();
  for (int row = threadIdx.x; row < n; row += blockDim.x) {
    double acc = 0.0;
    const double* c = C + (size_t)row * n;
    for (int k = 0; k < n; ++k) acc += c[k] * sb[k];
    x[row] = (T)acc;
  }
}

// Fine-level operator / preconditioner application for the C-ABI.
__global__ void __launch_bounds__(kBlock)
k_apply_fine(FineOp A, const double* __restrict__ x, double* __restrict__ y) {
  Tile2D t;
  if (!cell_of_tile(A.nx, A.ny, t)) return;
  const int i = t.i, j = t.j, k = i * A.ny + j;
  const uint8_t f = A.flags[k];
  const Nb nb = fine_neighbours(A, i, j);
  auto xv = [&](int q) -> double { return x[q]; };
  y[k] = fine_apply<double>(A, f, nb, k, fine_diag(A, f, i, j), xv);
}

// Opens the state for a single preconditioner application: mean of r over
// the active set, loop flags cleared.
__global__ void __launch_bounds__(kBlock)
k_open_for_apply(FineOp A, const double* __restrict__ r, int n_active, double* partial,
                 unsigned int* counter, SolveState* st) {
  Tile2D t;
  double s[1] = {0.0};
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int k = t.i * A.ny + t.j;
    if (A.flags[k] & F_SOLV) s[0] = r[k];
  }
  const int ops[1] = {kSum};
  if (grid_reduce<1>(s, ops, partial, counter, t.tile, gridDim.x) && threadIdx.x == 0) {
    st->mean_r = (n_active > 0) ? s[0] / (double)n_active : 0.0;
    st->done = 0;
    st->iter = 0;
    st->rho = 1.0;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cast_out(FineOp A, const T* __restrict__ z, const SolveState* st, double* __restrict__ out) {
  Tile2D t;
  if (!cell_of_tile(A.nx, A.ny, t)) return;
  const int k = t.i * A.ny + t.j;
  out[k] = (A.flags[k] & F_SOLV) ? (double)z[k] - st->mean_z : 0.0;
}

// ============================================================================
// Host side
// ============================================================================
static inline dim3 tiles_for(int nx, int ny) { return dim3(num_tiles(nx, ny)); }

PoissonEngine::~PoissonEngine() {
  free_levels();
  hipFree(b_);
  hipFree(x_);
  hipFree(r_);
  hipFree(pa_);
  hipFree(pb_);
  hipFree(q_);
  hipFree(z_);
  hipFree(st_);
  hipFree(partial_);
  hipFree(counter_);
  hipFree(fine_active_);
  if (st_host_) hipHostFree(st_host_);
  if (ev_) hipEventDestroy(ev_);
}

void PoissonEngine::free_levels() {
  for (auto* v : {&lv64_, &lv32_}) {
    for (auto& L : *v) {
      hipFree(L.b);
      hipFree(L.x);
      hipFree(L.y);
      hipFree(L.r);
      hipFree(L.di);
      hipFree(L.dj);
      hipFree(L.cls);
      hipFree(L.table);
      hipFree(L.cdinv);
      hipFree(L.cactive);
      hipFree(L.cinv);
    }
    v->clear();
  }
}

int PoissonEngine::init(const FineOp& op, int n_fluid, int singular, hipStream_t stream) {
  op_ = op;
  s_ = stream;
  n_ = op.nx * op.ny;
  n_fluid_ = n_fluid;
  singular_ = singular;
  const size_t bytes = (size_t)n_ * sizeof(double);
  PCFD_HIP_CHECK(hipMalloc(&b_, bytes));
  PCFD_HIP_CHECK(hipMalloc(&x_, bytes));
  PCFD_HIP_CHECK(hipMalloc(&r_, bytes));
  PCFD_HIP_CHECK(hipMalloc(&pa_, bytes));
  PCFD_HIP_CHECK(hipMalloc(&pb_, bytes));
  PCFD_HIP_CHECK(hipMalloc(&q_, bytes));
  PCFD_HIP_CHECK(hipMalloc(&z_, bytes));
  for (double* v : {b_, x_, r_, pa_, pb_, q_, z_}) PCFD_HIP_CHECK(hipMemsetAsync(v, 0, bytes, s_));
  const size_t ntile = (size_t)num_tiles(op.nx + 3, op.ny + 3) + 64;
  PCFD_HIP_CHECK(hipMalloc(&partial_, ntile * 8 * sizeof(double)));
  PCFD_HIP_CHECK(hipMalloc(&counter_, 64 * sizeof(unsigned int)));
  PCFD_HIP_CHECK(hipMemsetAsync(counter_, 0, 64 * sizeof(unsigned int), s_));
  PCFD_HIP_CHECK(hipMalloc(&st_, sizeof(SolveState)));
  PCFD_HIP_CHECK(hipMemsetAsync(st_, 0, sizeof(SolveState), s_));
  PCFD_HIP_CHECK(hipHostMalloc(&st_host_, sizeof(SolveState), hipHostMallocDefault));
  PCFD_HIP_CHECK(hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
  return kOk;
}

template <typename T>
static int upload_level(const MGLevelHost& h, bool fine, PoissonEngine* /*unused*/, void** b,
                        void** x, void** y, void** r, int8_t** di, int8_t** dj, uint16_t** cls,
                        void** table, void** cdinv, uint8_t** cactive, double** cinv,
                        hipStream_t s) {
  const size_t n = (size_t)h.nx * h.ny;
  PCFD_HIP_CHECK(hipMalloc(b, n * sizeof(T)));
  PCFD_HIP_CHECK(hipMalloc(x, n * sizeof(T)));
  PCFD_HIP_CHECK(hipMalloc(y, n * sizeof(T)));
  PCFD_HIP_CHECK(hipMalloc(r, n * sizeof(T)));
  PCFD_HIP_CHECK(hipMemsetAsync(*x, 0, n * sizeof(T), s));
  PCFD_HIP_CHECK(hipMemsetAsync(*y, 0, n * sizeof(T), s));
  if (!fine) {
    const size_t S = h.di.size(), nc = h.cdinv.size();
    std::vector<T> tab(h.table.size()), cd(nc);
    for (size_t k = 0; k < tab.size(); ++k) tab[k] = (T)h.table[k];
    for (size_t k = 0; k < nc; ++k) cd[k] = (T)h.cdinv[k];
    PCFD_HIP_CHECK(hipMalloc(di, S));
    PCFD_HIP_CHECK(hipMalloc(dj, S));
    PCFD_HIP_CHECK(hipMalloc(cls, n * sizeof(uint16_t)));
    PCFD_HIP_CHECK(hipMalloc(table, tab.size() * sizeof(T)));
    PCFD_HIP_CHECK(hipMalloc(cdinv, nc * sizeof(T)));
    PCFD_HIP_CHECK(hipMalloc(cactive, nc));
    PCFD_HIP_CHECK(hipMemcpy(*di, h.di.data(), S, hipMemcpyHostToDevice));
    PCFD_HIP_CHECK(hipMemcpy(*dj, h.dj.data(), S, hipMemcpyHostToDevice));
    PCFD_HIP_CHECK(hipMemcpy(*cls, h.cls.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice));
    PCFD_HIP_CHECK(hipMemcpy(*table, tab.data(), tab.size() * sizeof(T), hipMemcpyHostToDevice));
    PCFD_HIP_CHECK(hipMemcpy(*cdinv, cd.data(), nc * sizeof(T), hipMemcpyHostToDevice));
    PCFD_HIP_CHECK(hipMemcpy(*cactive, h.cactive.data(), nc, hipMemcpyHostToDevice));
  }
  if (!h.coarse_inv.empty()) {
    PCFD_HIP_CHECK(hipMalloc(cinv, h.coarse_inv.size() * sizeof(double)));
    PCFD_HIP_CHECK(hipMemcpy(*cinv, h.coarse_inv.data(), h.coarse_inv.size() * sizeof(double),
                             hipMemcpyHostToDevice));
  }
  return kOk;
}

int PoissonEngine::configure(const PoissonConfig& cfg, const std::vector<MGLevelHost>& levels) {
  cfg_ = cfg;
  free_levels();
  last_iters_ = 0;
  predicted_ = 0;
  const bool need_mg = cfg.kind == SK_MULTIGRID || cfg.kind == SK_MGCG || cfg.kind == SK_DIRECT;
  if (!need_mg) return kOk;
  if (levels.empty()) {
    set_last_error("multigrid solver configured without a hierarchy");
    return kErrBadArg;
  }
  n_active0_ = 0;
  for (uint8_t a : levels[0].active) n_active0_ += a ? 1 : 0;
  // multigrid.py:453-458: the mean is removed only when the operator is singular
  n_proj_ = singular_ ? n_active0_ : 0;
  weight0_ = levels[0].weight;
  for (int pass = 0; pass < 2; ++pass) {
    const bool f32 = pass == 1;
    if (f32 && cfg.precision == 0) break;
    auto& dst = f32 ? lv32_ : lv64_;
    for (size_t l = 0; l < levels.size(); ++l) {
      const MGLevelHost& h = levels[l];
      LevelBuf L{};
      L.nx = h.nx;
      L.ny = h.ny;
      L.n = h.nx * h.ny;
      L.fx = h.fx;
      L.fy = h.fy;
      L.weight = h.weight;
      L.S = (int)h.di.size();
      L.ncls = (int)h.cdinv.size();
      int rc = f32 ? upload_level<float>(h, l == 0, this, &L.b, &L.x, &L.y, &L.r, &L.di, &L.dj,
                                         &L.cls, &L.table, &L.cdinv, &L.cactive, &L.cinv, s_)
                   : upload_level<double>(h, l == 0, this, &L.b, &L.x, &L.y, &L.r, &L.di,
                                          &L.dj, &L.cls, &L.table, &L.cdinv, &L.cactive,
                                          &L.cinv, s_);
      if (rc) return rc;
      dst.push_back(L);
    }
  }
  PCFD_HIP_CHECK(hipStreamSynchronize(s_));
  return kOk;
}

int PoissonEngine::prepare_from_velocity(const double* u, const double* v, int ldu, int ldv,
                                         double dx, double dy, double dt) {
  hipLaunchKernelGGL(k_prepare, tiles_for(op_.nx, op_.ny), dim3(kBlock), 0, s_, op_, u, v, ldu,
                     ldv, dx, dy, dt, (const double*)nullptr, b_, partial_, counter_, st_,
                     singular_, n_fluid_);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int PoissonEngine::prepare_from_rhs(const double* rhs) {
  hipLaunchKernelGGL(k_prepare, tiles_for(op_.nx, op_.ny), dim3(kBlock), 0, s_, op_,
                     (const double*)nullptr, (const double*)nullptr, 0, 0, 1.0, 1.0, 1.0, rhs,
                     b_, partial_, counter_, st_, singular_, n_fluid_);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int PoissonEngine::reset_guess() {
  PCFD_HIP_CHECK(hipMemsetAsync(x_, 0, (size_t)n_ * sizeof(double), s_));
  return kOk;
}

int PoissonEngine::set_guess(const double* x0) {
  PCFD_HIP_CHECK(
      hipMemcpyAsync(x_, x0, (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
  return kOk;
}

int PoissonEngine::read_state(SolveState* st) {
  PCFD_HIP_CHECK(hipMemcpyAsync(st_host_, st_, sizeof(SolveState), hipMemcpyDeviceToHost, s_));
  PCFD_HIP_CHECK(hipStreamSynchronize(s_));
  std::memcpy(st, st_host_, sizeof(SolveState));
  return kOk;
}

int PoissonEngine::poll(int* iter, int* done) {
  PCFD_HIP_CHECK(hipMemcpyAsync(st_host_, st_, sizeof(SolveState), hipMemcpyDeviceToHost, s_));
  PCFD_HIP_CHECK(hipEventRecord(ev_, s_));
  PCFD_HIP_CHECK(hipEventSynchronize(ev_));
  *iter = st_host_->iter;
  *done = st_host_->done;
  return kOk;
}

int PoissonEngine::solve(SolveInfo* info, bool sync_info) {
  int rc = kOk;
  switch (cfg_.kind) {
    case SK_CG: rc = run_cg(); break;
    case SK_DIRECT:
    case SK_MGCG: rc = run_mgcg(); break;
    case SK_SOR: rc = run_sor(); break;
    case SK_JACOBI: rc = run_jacobi(); break;
    case SK_MULTIGRID: rc = run_multigrid(); break;
    default: set_last_error("unknown solver kind %d", cfg_.kind); return kErrBadArg;
  }
  if (rc) return rc;
  const bool recompute = cfg_.kind == SK_CG || cfg_.kind == SK_MGCG || cfg_.kind == SK_DIRECT;
  rc = finalize(recompute);
  if (rc) return rc;
  if (info && sync_info) {
    SolveState st;
    rc = read_state(&st);
    if (rc) return rc;
    info->iterations = st.iter;
    info->residual = recompute ? st.res_final : st.res;
    info->converged = st.converged;
    info->status = kOk;
    last_iters_ = st.iter;
  }
  return kOk;
}

int PoissonEngine::finalize(bool recompute_residual) {
  (void)recompute_residual;
  hipLaunchKernelGGL(k_check, tiles_for(op_.nx, op_.ny), dim3(kBlock), 0, s_, op_, x_, b_,
                     (int)CM_FINAL, cfg_.tol, cfg_.maxiter, n_fluid_, singular_, partial_,
                     counter_, st_);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int PoissonEngine::scatter_pressure(double* p, int ldp) {
  hipLaunchKernelGGL(k_scatter, tiles_for(op_.nx + 2, op_.ny + 2), dim3(kBlock), 0, s_, op_, x_,
                     st_, p, ldp, (double*)nullptr);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int PoissonEngine::write_interior(double* out) {
  hipLaunchKernelGGL(k_scatter, tiles_for(op_.nx + 2, op_.ny + 2), dim3(kBlock), 0, s_, op_, x_,
                     st_, (double*)nullptr, 0, out);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int PoissonEngine::apply_operator(const double* x, double* y) {
  hipLaunchKernelGGL(k_apply_fine, tiles_for(op_.nx, op_.ny), dim3(kBlock), 0, s_, op_, x, y);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

// Iteration driver shared by the solvers: enqueue a chunk of iterations,
// then poll the device flag.  Kernels turn into no-ops once the flag is set,
// so over-enqueueing only costs launch slots; the chunk follows the previous
// solve's count (warm-started solves barely change) to keep polls rare.
// PER is the number of iterations one BODY advances; MAXB bounds the bodies.
#define PCFD_ITERATE(PER, MAXB, ...)                                          \
  do {                                                                        \
    int _it = 0, _done = 0, _bodies = 0;                                      \
    int _rc = poll(&_it, &_done);                                             \
    if (_rc) return _rc;                                                      \
    int _chunk = std::max(1, (predicted_ > 0 ? predicted_ - 1 : 4) / (PER));  \
    while (!_done && _bodies < (MAXB)) {                                      \
      for (int _c = 0; _c < _chunk && _bodies < (MAXB); ++_c, ++_bodies) {    \
        __VA_ARGS__;                                                          \
      }                                                                       \
      _rc = poll(&_it, &_done);                                               \
      if (_rc) return _rc;                                                    \
      _chunk = std::max(1, std::min(64, _it / (8 * (PER))));                  \
    }                                                                         \
    predicted_ = _it;                                                         \
  } while (0)

int PoissonEngine::run_cg() {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  hipLaunchKernelGGL(k_zero_nonfluid, g, dim3(kBlock), 0, s_, op_, x_);
  hipLaunchKernelGGL(k_init, g, dim3(kBlock), 0, s_, op_, b_, x_, r_, 0, (int)IM_JPCG, cfg_.tol,
                     cfg_.maxiter, n_fluid_, partial_, counter_, st_);
  PCFD_LAUNCH_CHECK();
  double* p_old = pa_;
  double* p_new = pb_;
  PCFD_ITERATE(1, cfg_.maxiter, {
    hipLaunchKernelGGL(k_jpcg_dir, g, dim3(kBlock), 0, s_, op_, r_, p_old, p_new, q_, partial_,
                       counter_, st_);
    hipLaunchKernelGGL(k_cg_update, g, dim3(kBlock), 0, s_, op_, x_, r_, p_new, q_,
                       (int)IM_JPCG, cfg_.maxiter, n_fluid_, partial_, counter_, st_);
    PCFD_LAUNCH_CHECK();
    std::swap(p_old, p_new);
  });
  return kOk;
}

int PoissonEngine::run_sor() {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  hipLaunchKernelGGL(k_zero_nonfluid, g, dim3(kBlock), 0, s_, op_, x_);
  hipLaunchKernelGGL(k_init, g, dim3(kBlock), 0, s_, op_, b_, x_, (double*)nullptr, 0,
                     (int)IM_RELAX, cfg_.tol, cfg_.maxiter, n_fluid_, partial_, counter_, st_);
  PCFD_LAUNCH_CHECK();
  int sweeps_enqueued = 0;
  PCFD_ITERATE(5, (cfg_.maxiter + 4) / 5, {
    const int until = std::min(cfg_.maxiter, (sweeps_enqueued / 5 + 1) * 5);
    for (; sweeps_enqueued < until; ++sweeps_enqueued) {
      hipLaunchKernelGGL(k_sor_colour, g, dim3(kBlock), 0, s_, op_, x_, b_, 0, cfg_.omega,
                         n_fluid_, singular_, partial_, counter_, st_);
      hipLaunchKernelGGL(k_sor_colour, g, dim3(kBlock), 0, s_, op_, x_, b_, 1, cfg_.omega,
                         n_fluid_, singular_, partial_, counter_, st_);
    }
    hipLaunchKernelGGL(k_check, g, dim3(kBlock), 0, s_, op_, x_, b_, (int)CM_RELAX, cfg_.tol,
                       cfg_.maxiter, n_fluid_, singular_, partial_, counter_, st_);
    PCFD_LAUNCH_CHECK();
  });
  return kOk;
}

int PoissonEngine::run_jacobi() {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  hipLaunchKernelGGL(k_zero_nonfluid, g, dim3(kBlock), 0, s_, op_, x_);
  hipLaunchKernelGGL(k_init, g, dim3(kBlock), 0, s_, op_, b_, x_, (double*)nullptr, 0,
                     (int)IM_RELAX, cfg_.tol, cfg_.maxiter, n_fluid_, partial_, counter_, st_);
  PCFD_LAUNCH_CHECK();
  int sweeps_enqueued = 0;
  double* cur = x_;
  double* nxt = pa_;
  PCFD_ITERATE(20, (cfg_.maxiter + 19) / 20, {
    const int until = std::min(cfg_.maxiter, (sweeps_enqueued / 20 + 1) * 20);
    for (; sweeps_enqueued < until; ++sweeps_enqueued) {
      hipLaunchKernelGGL(k_jacobi_sweep, g, dim3(kBlock), 0, s_, op_, cur, nxt, b_, cfg_.omega,
                         n_fluid_, singular_, partial_, counter_, st_);
      std::swap(cur, nxt);
    }
    hipLaunchKernelGGL(k_check, g, dim3(kBlock), 0, s_, op_, cur, b_, (int)CM_RELAX, cfg_.tol,
                       cfg_.maxiter, n_fluid_, singular_, partial_, counter_, st_);
    PCFD_LAUNCH_CHECK();
  });
  if (cur != x_) {
    PCFD_HIP_CHECK(
        hipMemcpyAsync(x_, cur, (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
  }
  return kOk;
}

// One V-cycle (multigrid.py:430-451) over the levels in precision T.  Level
// 0's right-hand side is lv[0].b; the result is left in lv[0].x (unprojected,
// with the active-set mean recorded by the last sweep).
template <typename T>
int PoissonEngine::enqueue_vcycle_t(bool nonzero_start) {
  auto& lv = (sizeof(T) == 4) ? lv32_ : lv64_;
  const int nl = (int)lv.size();
  const int nu = cfg_.sweeps;
  const int red = (cfg_.kind == SK_MULTIGRID) ? SR_MG : SR_MGCG;
  FineLevelArgs<T> F{op_, (T)lv[0].weight};
  auto dia = [&](LevelBuf& L) {
    ClassLevel<T> D;
    D.nx = L.nx;
    D.ny = L.ny;
    D.n = L.n;
    D.px = op_.px;
    D.py = op_.py;
    D.S = L.S;
    D.di = L.di;
    D.dj = L.dj;
    D.cls = L.cls;
    D.table = (const T*)L.table;
    D.cdinv = (const T*)L.cdinv;
    D.cactive = L.cactive;
    return D;
  };
  if (nl == 1) {
    LevelBuf& L = lv[0];
    hipLaunchKernelGGL(k_coarse_solve<T>, dim3(1), dim3(kBlock), L.n * sizeof(double), s_,
                       (const double*)L.cinv, L.n, (const T*)L.b, (T*)L.x, st_);
    hipLaunchKernelGGL(k_cycle_reduce<T>, tiles_for(op_.nx, op_.ny), dim3(kBlock), 0, s_, op_,
                       (const T*)L.b, (const T*)L.x, red, n_proj_, partial_, counter_, st_);
    PCFD_LAUNCH_CHECK();
    return kOk;
  }
  // Downward leg.
  for (int l = 0; l < nl - 1; ++l) {
    LevelBuf& L = lv[l];
    LevelBuf& C = lv[l + 1];
    const bool nz = (l == 0) && nonzero_start;
    const dim3 g = tiles_for(L.nx, L.ny);
    for (int s = 0; s < nu; ++s) {
      const T* xin = (s == 0 && !nz) ? nullptr : (const T*)L.x;
      if (l == 0) {
        hipLaunchKernelGGL(k_smooth_fine<T>, g, dim3(kBlock), 0, s_, F, (const T*)L.b, xin,
                           (T*)L.y, (s == 0 && nz) ? 1 : 0, (int)SR_NONE, n_proj_, partial_,
                           counter_, st_);
      } else {
        hipLaunchKernelGGL(k_smooth_cls<T>, g, dim3(kBlock), 0, s_, dia(L), (T)L.weight,
                           (const T*)L.b, xin, (T*)L.y, st_);
      }
      std::swap(L.x, L.y);
    }
    if (l == 0) {
      hipLaunchKernelGGL(k_residual_fine<T>, g, dim3(kBlock), 0, s_, op_, (const T*)L.b,
                         (const T*)L.x, (T*)L.r, st_);
      hipLaunchKernelGGL(k_restrict_fine<T>, tiles_for(C.nx, C.ny), dim3(kBlock), 0, s_, F, L.fx,
                         L.fy, C.nx, C.ny, (const T*)L.r, (T*)C.b, st_);
    } else {
      hipLaunchKernelGGL(k_residual_cls<T>, g, dim3(kBlock), 0, s_, dia(L), (const T*)L.b,
                         (const T*)L.x, (T*)L.r, st_);
      hipLaunchKernelGGL(k_restrict_cls<T>, tiles_for(C.nx, C.ny), dim3(kBlock), 0, s_, dia(L),
                         (T)L.weight, L.fx, L.fy, C.nx, C.ny, (const T*)L.r, (T*)C.b, st_);
    }
    PCFD_LAUNCH_CHECK();
  }
  {
    LevelBuf& L = lv[nl - 1];
    hipLaunchKernelGGL(k_coarse_solve<T>, dim3(1), dim3(kBlock), L.n * sizeof(double), s_,
                       (const double*)L.cinv, L.n, (const T*)L.b, (T*)L.x, st_);
    PCFD_LAUNCH_CHECK();
  }
  // Upward leg.
  for (int l = nl - 2; l >= 0; --l) {
    LevelBuf& L = lv[l];
    LevelBuf& C = lv[l + 1];
    const dim3 g = tiles_for(L.nx, L.ny);
    if (l == 0) {
      hipLaunchKernelGGL(k_prolong_fine<T>, g, dim3(kBlock), 0, s_, F, L.fx, L.fy, C.ny,
                         (const T*)C.x, (T*)L.x, st_);
    } else {
      hipLaunchKernelGGL(k_prolong_cls<T>, g, dim3(kBlock), 0, s_, dia(L), (T)L.weight, L.fx,
                         L.fy, C.ny, (const T*)C.x, (T*)L.x, st_);
    }
    for (int s = 0; s < nu; ++s) {
      if (l == 0) {
        const int rmode = (s == nu - 1) ? red : (int)SR_NONE;
        hipLaunchKernelGGL(k_smooth_fine<T>, g, dim3(kBlock), 0, s_, F, (const T*)L.b,
                           (const T*)L.x, (T*)L.y, 0, rmode, n_proj_, partial_, counter_,
                           st_);
      } else {
        hipLaunchKernelGGL(k_smooth_cls<T>, g, dim3(kBlock), 0, s_, dia(L), (T)L.weight,
                           (const T*)L.b, (const T*)L.x, (T*)L.y, st_);
      }
      std::swap(L.x, L.y);
    }
    PCFD_LAUNCH_CHECK();
  }
  return kOk;
}

int PoissonEngine::enqueue_vcycle(bool nonzero_start) {
  if (cfg_.precision == 1) return enqueue_vcycle_t<float>(nonzero_start);
  return enqueue_vcycle_t<double>(nonzero_start);
}

int PoissonEngine::run_mgcg() {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  hipLaunchKernelGGL(k_zero_nonfluid, g, dim3(kBlock), 0, s_, op_, x_);
  hipLaunchKernelGGL(k_init, g, dim3(kBlock), 0, s_, op_, b_, x_, r_, 0, (int)IM_MGCG, cfg_.tol,
                     cfg_.maxiter, n_proj_, partial_, counter_, st_);
  PCFD_LAUNCH_CHECK();
  double* p_old = pa_;
  double* p_new = pb_;
  const bool f32 = cfg_.precision == 1;
  PCFD_ITERATE(1, cfg_.maxiter, {
    if (f32) {
      hipLaunchKernelGGL(k_project_in<float>, g, dim3(kBlock), 0, s_, op_, r_, (float*)lv32_[0].b,
                         st_);
    } else {
      hipLaunchKernelGGL(k_project_in<double>, g, dim3(kBlock), 0, s_, op_, r_,
                         (double*)lv64_[0].b, st_);
    }
    int rc = enqueue_vcycle(false);
    if (rc) return rc;
    if (f32) {
      hipLaunchKernelGGL(k_mgcg_dir<float>, g, dim3(kBlock), 0, s_, op_,
                         (const float*)lv32_[0].x, p_old, p_new, q_, partial_, counter_, st_);
    } else {
      hipLaunchKernelGGL(k_mgcg_dir<double>, g, dim3(kBlock), 0, s_, op_,
                         (const double*)lv64_[0].x, p_old, p_new, q_, partial_, counter_, st_);
    }
    hipLaunchKernelGGL(k_cg_update, g, dim3(kBlock), 0, s_, op_, x_, r_, p_new, q_, (int)IM_MGCG,
                       cfg_.maxiter, n_proj_, partial_, counter_, st_);
    PCFD_LAUNCH_CHECK();
    std::swap(p_old, p_new);
  });
  return kOk;
}

int PoissonEngine::run_multigrid() {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  hipLaunchKernelGGL(k_zero_nonfluid, g, dim3(kBlock), 0, s_, op_, x_);
  hipLaunchKernelGGL(k_init, g, dim3(kBlock), 0, s_, op_, b_, x_, (double*)nullptr, 0,
                     (int)IM_MG, cfg_.tol, cfg_.maxiter, n_proj_, partial_, counter_, st_);
  PCFD_LAUNCH_CHECK();
  LevelBuf& L0 = lv64_[0];
  // the hierarchy's level-0 vectors alias the solver's: b and the iterate
  const size_t bytes = (size_t)n_ * sizeof(double);
  PCFD_HIP_CHECK(hipMemcpyAsync(L0.b, b_, bytes, hipMemcpyDeviceToDevice, s_));
  PCFD_HIP_CHECK(hipMemcpyAsync(L0.x, x_, bytes, hipMemcpyDeviceToDevice, s_));
  PCFD_ITERATE(1, cfg_.maxiter, {
    int rc = enqueue_vcycle_t<double>(true);
    if (rc) return rc;
    hipLaunchKernelGGL(k_check, g, dim3(kBlock), 0, s_, op_, (const double*)L0.x, b_,
                       (int)CM_MG, cfg_.tol, cfg_.maxiter, n_fluid_, singular_, partial_,
                       counter_, st_);
    PCFD_LAUNCH_CHECK();
  });
  PCFD_HIP_CHECK(hipMemcpyAsync(x_, L0.x, bytes, hipMemcpyDeviceToDevice, s_));
  return kOk;
}

int PoissonEngine::apply_preconditioner(const double* r, double* z) {
  // One V-cycle from zero applied to r, projected on both sides
  // (MultigridHierarchy.apply, multigrid.py:460-474).
  if (lv64_.empty() && lv32_.empty()) {
    set_last_error("no multigrid hierarchy configured");
    return kErrBadArg;
  }
  const dim3 g = tiles_for(op_.nx, op_.ny);
  PCFD_HIP_CHECK(hipMemcpyAsync(r_, r, (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
  hipLaunchKernelGGL(k_open_for_apply, g, dim3(kBlock), 0, s_, op_, (const double*)r_,
                     n_proj_, partial_, counter_, st_);
  const bool f32 = cfg_.precision == 1;
  if (f32) {
    hipLaunchKernelGGL(k_project_in<float>, g, dim3(kBlock), 0, s_, op_, r_, (float*)lv32_[0].b,
                       st_);
  } else {
    hipLaunchKernelGGL(k_project_in<double>, g, dim3(kBlock), 0, s_, op_, r_, (double*)lv64_[0].b,
                       st_);
  }
  int rc = enqueue_vcycle(false);
  if (rc) return rc;
  if (f32) {
    hipLaunchKernelGGL(k_cast_out<float>, g, dim3(kBlock), 0, s_, op_, (const float*)lv32_[0].x,
                       st_, z);
  } else {
    hipLaunchKernelGGL(k_cast_out<double>, g, dim3(kBlock), 0, s_, op_,
                       (const double*)lv64_[0].x, st_, z);
  }
  PCFD_LAUNCH_CHECK();
  PCFD_HIP_CHECK(hipStreamSynchronize(s_));
  return kOk;
}

int PoissonEngine::time_kernel(int which, int reps, double* avg_ms, double* bytes) {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  const double n = (double)n_;
  SolveState* st = nullptr;
  PCFD_HIP_CHECK(hipMalloc(&st, sizeof(SolveState)));
  PCFD_HIP_CHECK(hipMemsetAsync(st, 0, sizeof(SolveState), s_));
  hipEvent_t e0, e1;
  PCFD_HIP_CHECK(hipEventCreate(&e0));
  PCFD_HIP_CHECK(hipEventCreate(&e1));
  const bool have_mg = !lv64_.empty();
  double B = 0.0;
  auto launch = [&]() -> int {
    switch (which) {
      case 0:  // CG update: read x, p, r, q + flags, write x, r
        hipLaunchKernelGGL(k_cg_update, g, dim3(kBlock), 0, s_, op_, x_, r_, pa_, q_,
                           (int)IM_MGCG, 1 << 30, n_active0_, partial_, counter_, st);
        B = n * (6 * 8 + 1);
        break;
      case 1:  // multigrid-CG direction + SpMV: read z, p_old + flags, write p, q
        hipLaunchKernelGGL(k_mgcg_dir<double>, g, dim3(kBlock), 0, s_, op_, (const double*)z_,
                           pa_, pb_, q_, partial_, counter_, st);
        B = n * (4 * 8 + 1);
        break;
      case 2: {  // fine damped-Jacobi sweep: read b, x + flags, write x'
        if (!have_mg) return kErrBadArg;
        FineLevelArgs<double> F{op_, lv64_[0].weight};
        hipLaunchKernelGGL(k_smooth_fine<double>, g, dim3(kBlock), 0, s_, F,
                           (const double*)b_, (const double*)x_, z_, 0, (int)SR_NONE,
                           n_active0_, partial_, counter_, st);
        B = n * (3 * 8 + 1);
        break;
      }
      case 3:  // fine residual: read b, x + flags, write r
        hipLaunchKernelGGL(k_residual_fine<double>, g, dim3(kBlock), 0, s_, op_,
                           (const double*)b_, (const double*)x_, z_, st);
        B = n * (3 * 8 + 1);
        break;
      default:
        return kErrBadArg;
    }
    return kOk;
  };
  int rc = launch();  // warm
  if (rc) return rc;
  PCFD_HIP_CHECK(hipEventRecord(e0, s_));
  for (int k = 0; k < reps; ++k) launch();
  PCFD_HIP_CHECK(hipEventRecord(e1, s_));
  PCFD_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  PCFD_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  PCFD_LAUNCH_CHECK();
  *avg_ms = (double)ms / reps;
  *bytes = B;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(st);
  return kOk;
}

}  // namespace pcfd
