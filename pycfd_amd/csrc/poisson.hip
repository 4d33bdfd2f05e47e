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
//
// Every reduction is two launches: the producing kernel writes per-tile
// partials, k_fold sums them in tile order and runs the epilogue (the scalar
// recurrences of the solver).  Solver scalars live in a device SolveState;
// kernels turn into no-ops once `done` is set, so iterations are enqueued
// without waiting for the host.
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

__device__ __forceinline__ double dirichlet_rhs(const FineOp& A, int i, int j) {
  // pressure.py:233-238: bc -= 2*coeff*p_ref on the faces of a Dirichlet wall.
  double bc = 0.0;
  if (A.dir[0] && i == 0) bc -= 2.0 * A.cx * A.pref[0];
  if (A.dir[1] && i == A.nx - 1) bc -= 2.0 * A.cx * A.pref[1];
  if (A.dir[2] && j == 0) bc -= 2.0 * A.cy * A.pref[2];
  if (A.dir[3] && j == A.ny - 1) bc -= 2.0 * A.cy * A.pref[3];
  return bc;
}

// ----------------------------------------------------------------------------
// Reduction epilogues (run by k_fold in one thread after the partials are in)
// ----------------------------------------------------------------------------
enum InitMode : int { IM_JPCG = 0, IM_MGCG = 1, IM_RELAX = 2, IM_MG = 3 };
enum CheckMode : int { CM_RELAX = 0, CM_MG = 1, CM_FINAL = 2 };
enum SmoothReduce : int { SR_NONE = 0, SR_MG = 1, SR_MGCG = 2 };

struct EpiPrepare {  // fluid mean of the right-hand side (pressure.py:306-307)
  static constexpr int K = 1;
  SolveState* st;
  int singular, nfluid;
  __device__ void ops(int (&o)[K]) const { o[0] = kSum; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&v)[K]) const {
    st->mean_b = (singular && nfluid > 0) ? v[0] / (double)nfluid : 0.0;
  }
};

struct EpiInit {  // ||b||, ||r||, first rho, termination at iteration 0
  static constexpr int K = 4;
  SolveState* st;
  int mode;
  double tol;
  int maxiter, n_proj;
  __device__ void ops(int (&o)[K]) const { o[0] = o[1] = o[2] = o[3] = kSum; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&v)[K]) const {
    const double bn = sqrt(v[0]);
    st->bnorm = bn;
    st->zero_rhs = (mode == IM_JPCG || mode == IM_MGCG) && (bn == 0.0);
    st->rr = v[1];
    st->iter = 0;
    st->stalled = 0;
    st->mean_x = 0.0;
    st->mean_z = 0.0;
    st->atol = tol * bn;
    const double res = sqrt(v[1]) / (bn > 1.0e-300 ? bn : 1.0);
    st->res = res;
    st->res_prev = res;
    int done = 0, conv = 0;
    if (mode == IM_JPCG || mode == IM_MGCG) {
      st->rho = v[2];
      st->mean_r = (n_proj > 0) ? v[3] / (double)n_proj : 0.0;
      if (bn == 0.0 || sqrt(v[1]) < st->atol) {
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
    } else if (maxiter <= 0) {
      done = 1;
    }
    st->done = done;
    st->converged = conv;
  }
};

struct EpiAlpha {  // alpha = rho / p.q
  static constexpr int K = 1;
  SolveState* st;
  __device__ void ops(int (&o)[K]) const { o[0] = kSum; }
  __device__ bool active() const { return !st->done; }
  __device__ void operator()(const double (&v)[K]) const { st->alpha = st->rho / v[0]; }
};

struct EpiUpdate {  // after x += alpha p, r -= alpha q (scipy cg loop)
  static constexpr int K = 2;
  SolveState* st;
  int mode, maxiter, n_proj;
  __device__ void ops(int (&o)[K]) const { o[0] = o[1] = kSum; }
  __device__ bool active() const { return !st->done; }
  __device__ void operator()(const double (&v)[K]) const {
    const int it = st->iter + 1;
    st->iter = it;
    st->rr = v[0];
    if (mode == IM_JPCG) {
      st->rho_prev = st->rho;
      st->rho = v[1];
      st->beta = st->rho / st->rho_prev;
    } else {
      st->mean_r = (n_proj > 0) ? v[1] / (double)n_proj : 0.0;
    }
    if (sqrt(v[0]) < st->atol) {
      st->done = 1;
      st->converged = 1;
    } else if (it >= maxiter) {
      st->done = 1;
    }
  }
};

struct EpiSor {  // colour sums -> the mean removed after a full sweep
  static constexpr int K = 1;
  SolveState* st;
  int colour, singular, nfluid;
  __device__ void ops(int (&o)[K]) const { o[0] = kSum; }
  __device__ bool active() const { return !st->done; }
  __device__ void operator()(const double (&v)[K]) const {
    if (colour == 0) {
      st->sum_red = v[0];
    } else {
      st->iter = st->iter + 1;
      st->mean_x = (singular && nfluid > 0) ? (st->sum_red + v[0]) / (double)nfluid : 0.0;
    }
  }
};

struct EpiJacobi {
  static constexpr int K = 1;
  SolveState* st;
  int singular, nfluid;
  __device__ void ops(int (&o)[K]) const { o[0] = kSum; }
  __device__ bool active() const { return !st->done; }
  __device__ void operator()(const double (&v)[K]) const {
    st->iter = st->iter + 1;
    st->mean_x = (singular && nfluid > 0) ? v[0] / (double)nfluid : 0.0;
  }
};

struct EpiCheck {  // relative residual; termination of the relaxation loops
  static constexpr int K = 2;
  SolveState* st;
  int mode;
  double tol;
  int maxiter, singular, nfluid;
  __device__ void ops(int (&o)[K]) const { o[0] = o[1] = kSum; }
  __device__ bool active() const { return mode == CM_FINAL || !st->done; }
  __device__ void operator()(const double (&v)[K]) const {
    const double bn = st->bnorm;
    const double res = sqrt(v[0]) / (bn > 1.0e-300 ? bn : 1.0);
    if (mode == CM_FINAL) {
      st->res_final = res;
      st->mean_final = (singular && nfluid > 0) ? v[1] / (double)nfluid : 0.0;
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
};

struct EpiCycle {  // projection of the V-cycle output; r.z for multigrid-CG
  static constexpr int K = 3;
  SolveState* st;
  int reduce, n_proj;
  __device__ void ops(int (&o)[K]) const { o[0] = o[1] = o[2] = kSum; }
  __device__ bool active() const { return !st->done; }
  __device__ void operator()(const double (&v)[K]) const {
    const double mz = (n_proj > 0) ? v[0] / (double)n_proj : 0.0;
    if (reduce == SR_MG) {
      st->mean_x = mz;
    } else {
      // r.z_proj with z_proj = z - mean_z and the cycle input b0 = r - mean_r:
      // sum(b0*z) - mean_z*sum(b0) is the same inner product.
      st->mean_z = mz;
      const double rho = v[1] - mz * v[2];
      st->rho_prev = st->rho;
      st->rho = rho;
      st->beta = rho / st->rho_prev;
    }
  }
};

struct EpiOpen {  // opens the state for one preconditioner application
  static constexpr int K = 1;
  SolveState* st;
  int n_proj;
  __device__ void ops(int (&o)[K]) const { o[0] = kSum; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&v)[K]) const {
    st->mean_r = (n_proj > 0) ? v[0] / (double)n_proj : 0.0;
    st->done = 0;
    st->iter = 0;
    st->rho = 1.0;
  }
};

// ----------------------------------------------------------------------------
// Right-hand side assembly (solver.py:659-661, pressure.py:300-308)
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_prepare(FineOp A, const double* __restrict__ u, const double* __restrict__ v, int ldu, int ldv,
          double dx, double dy, double dt, const double* __restrict__ rhs, double* __restrict__ b,
          double* partial) {
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
  block_partials<1>(s, ops, partial, t.tile);
}

// Solver start: b <- b - mean_b on fluid rows; r = b - A x (x = guess with
// non-fluid rows already zeroed); partials ||b||^2, ||r||^2 and, per mode,
// r.(r/d) (Jacobi-CG) or sum(r) (input projection of multigrid-CG).
__global__ void __launch_bounds__(kBlock)
k_init(FineOp A, double* __restrict__ b, const double* __restrict__ x, double* __restrict__ r,
       int mode, const SolveState* st, double* partial) {
  Tile2D t;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  const double mb = st->mean_b;
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int i = t.i, j = t.j, k = i * A.ny + j;
    const uint8_t f = A.flags[k];
    const Nb nb = fine_neighbours(A, i, j);
    const double bk = (f & F_SOLV) ? b[k] - mb : 0.0;
    auto xv = [&](int q) -> double { return x[q]; };
    const double d = fine_diag(A, f, i, j);
    const double rk = bk - fine_apply<double>(A, f, nb, k, d, xv);
    s[0] = bk * bk;
    s[1] = rk * rk;
    if (mode == IM_JPCG) s[2] = rk * (rk / d);
    if (mode == IM_MGCG && (f & F_SOLV)) s[3] = rk;
    b[k] = bk;
    if (r) r[k] = rk;
  }
  const int ops[4] = {kSum, kSum, kSum, kSum};
  block_partials<4>(s, ops, partial, t.tile);
}

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
           const SolveState* st) {
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
  block_partials<1>(s, ops, partial, t.tile);
}

// x += alpha p, r -= alpha q; partials ||r||^2 and r.(r/d) or sum(r).
__global__ void __launch_bounds__(kBlock)
k_cg_update(FineOp A, double* __restrict__ x, double* __restrict__ r,
            const double* __restrict__ p, const double* __restrict__ q, int mode,
            double* partial, const SolveState* st) {
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
  block_partials<2>(s, ops, partial, t.tile);
}

// Multigrid-CG direction: p = zp + beta*p_old with zp the projected V-cycle
// output (z - mean_z on the active set), q = A p, p.q.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_mgcg_dir(FineOp A, const T* __restrict__ z, const double* __restrict__ p_old,
           double* __restrict__ p_new, double* __restrict__ q, double* partial,
           const SolveState* st) {
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
  block_partials<1>(s, ops, partial, t.tile);
}

// V-cycle input: b0 = r - mean(r) on the active set, 0 elsewhere
// (MultigridHierarchy._project, multigrid.py:453-458).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_project_in(FineOp A, const double* __restrict__ r, T* __restrict__ b0, const SolveState* st) {
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
             double omega, double* partial, const SolveState* st) {
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
      double off = 0.0;  // A_rb p_b in CSR column order
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
  block_partials<1>(s, ops, partial, t.tile);
}

// Damped Jacobi sweep p_out = p + omega*dinv*(b - A p) (pressure.py:439-441).
__global__ void __launch_bounds__(kBlock)
k_jacobi_sweep(FineOp A, const double* __restrict__ p, double* __restrict__ pn,
               const double* __restrict__ b, double omega, double* partial,
               const SolveState* st) {
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
  block_partials<1>(s, ops, partial, t.tile);
}

// Residual of (x - pending shift) against b, and the fluid sum of the same x.
__global__ void __launch_bounds__(kBlock)
k_check(FineOp A, const double* __restrict__ x, const double* __restrict__ b, int mode,
        double* partial, const SolveState* st) {
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
  block_partials<2>(s, ops, partial, t.tile);
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
// Multigrid, level 0: the matrix-free fine operator.  T is the cycle's
// storage and arithmetic precision.
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

// Damped Jacobi sweep (multigrid.py:424-428); x_in == nullptr means a zero
// start.  A pending mean shift may be applied to x_in, and the last sweep of
// a cycle may take the sums sum(x), sum(b*x), sum(b) over the active set.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_smooth_fine(FineLevelArgs<T> L, const T* __restrict__ b, const T* __restrict__ x_in,
              T* __restrict__ x_out, int shift_in, int reduce, double* partial,
              const SolveState* st) {
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
  const int ops[3] = {kSum, kSum, kSum};
  block_partials<3>(s, ops, partial, t.tile);
}

// Projection sums of a single-level hierarchy (the coarsest solve is the
// whole cycle).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cycle_sums(FineOp A, const T* __restrict__ b, const T* __restrict__ z, double* partial,
             const SolveState* st) {
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
  const int ops[3] = {kSum, kSum, kSum};
  block_partials<3>(s, ops, partial, t.tile);
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

// ----------------------------------------------------------------------------
// Multigrid, levels >= 1: stored Galerkin operators (row dictionary), one
// row per LPR lanes; the lanes split the row's nonzeros and combine with
// cross-lane shuffles.
// ----------------------------------------------------------------------------
enum RowOp : int {
  RO_SMOOTH0_RES = 0,  // x = w D^-1 b (zero start), r = b - A x, t = D^-1 r
  RO_SMOOTH = 1,       // x_out = x + w D^-1 (b - A x)
  RO_RES = 2,          // r = b - A x, t = D^-1 r
  RO_RESTRICT_Y = 3,   // y = r - w A t on active rows, 0 elsewhere
  RO_PROLONG = 4,      // x += e - w D^-1 A e, e = P_tent x_c
};

template <typename T>
__device__ __forceinline__ int cls_nb(const ClassLevel<T>& L, int row, int e) {
  if (!(L.px | L.py)) return row + L.ent_lin[e];
  const int i = row / L.ny, j = row - i * L.ny;
  int ii = i + L.ent_di[e], jj = j + L.ent_dj[e];
  if (L.px) ii = (ii % L.nx + L.nx) % L.nx;
  if (L.py) jj = (jj % L.ny + L.ny) % L.ny;
  return ii * L.ny + jj;
}

template <typename T, int LPR, int OP>
__global__ void __launch_bounds__(kBlock)
k_cls_rows(ClassLevel<T> L, T w, const T* __restrict__ b, const T* __restrict__ xin,
           const T* __restrict__ xc, const int* __restrict__ agg, T* __restrict__ o1,
           T* __restrict__ o2, T* __restrict__ o3, const SolveState* st) {
  if (st->done) return;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = gid / LPR;
  const int lane = gid % LPR;
  const bool valid = row < L.n;
  T acc = (T)0;
  int c = 0;
  if (valid) {
    c = L.cls[row];
    const int e1 = L.row_ptr[c + 1];
    for (int e = L.row_ptr[c] + lane; e < e1; e += LPR) {
      const int q = cls_nb(L, row, e);
      T xv;
      if (OP == RO_SMOOTH0_RES) xv = w * L.cdinv[L.cls[q]] * b[q];
      else if (OP == RO_PROLONG) xv = (agg[q] >= 0) ? xc[agg[q]] : (T)0;
      else xv = xin[q];
      acc += L.ent_val[e] * xv;
    }
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
  if (!valid || lane != 0) return;
  const T dinv = L.cdinv[c];
  if (OP == RO_SMOOTH0_RES) {
    const T bk = b[row];
    const T rk = bk - acc;
    o1[row] = w * dinv * bk;
    o2[row] = rk;
    o3[row] = dinv * rk;
  } else if (OP == RO_SMOOTH) {
    o1[row] = xin[row] + w * dinv * (b[row] - acc);
  } else if (OP == RO_RES) {
    const T rk = b[row] - acc;
    o2[row] = rk;
    o3[row] = dinv * rk;
  } else if (OP == RO_RESTRICT_Y) {
    o1[row] = L.cactive[c] ? b[row] - w * acc : (T)0;
  } else {  // RO_PROLONG
    const T e = (agg[row] >= 0) ? xc[agg[row]] : (T)0;
    o1[row] = o1[row] + e - w * dinv * acc;
  }
}

// b_c[I] = sum of y over the cells of block I (P_tent^T y).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_block_sum(const T* __restrict__ y, int nx, int ny, int fx, int fy, int cnx, int cny,
            T* __restrict__ bc, const SolveState* st) {
  if (st->done) return;
  Tile2D t;
  if (!cell_of_tile(cnx, cny, t)) return;
  T acc = (T)0;
  for (int a = 0; a < fx; ++a) {
    const int i = t.i * fx + a;
    if (i >= nx) break;
    for (int c = 0; c < fy; ++c) {
      const int j = t.j * fy + c;
      if (j >= ny) break;
      acc += y[i * ny + j];
    }
  }
  bc[t.i * cny + t.j] = acc;
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

// ----------------------------------------------------------------------------
// Building blocks for the C-ABI
// ----------------------------------------------------------------------------
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

__global__ void __launch_bounds__(kBlock)
k_active_sum(FineOp A, const double* __restrict__ r, double* partial) {
  Tile2D t;
  double s[1] = {0.0};
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int k = t.i * A.ny + t.j;
    if (A.flags[k] & F_SOLV) s[0] = r[k];
  }
  const int ops[1] = {kSum};
  block_partials<1>(s, ops, partial, t.tile);
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
static inline int ntiles(int nx, int ny) { return num_tiles(nx, ny); }

template <class Epi>
static int fold(hipStream_t s, double* partial, int n, const Epi& e) {
  hipLaunchKernelGGL(k_fold<Epi>, dim3(1), dim3(kBlock), 0, s, (const double*)partial, n, e);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

PoissonEngine::~PoissonEngine() {
  free_levels();
  for (double* v : {b_, x_, r_, pa_, pb_, q_, z_}) hipFree(v);
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
      for (void* p : {L.b, L.x, L.y, L.r, L.t, L.ent_val, L.cdinv}) hipFree(p);
      hipFree(L.cls);
      hipFree(L.row_ptr);
      hipFree(L.ent_lin);
      hipFree(L.agg);
      hipFree(L.ent_di);
      hipFree(L.ent_dj);
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
  for (double** v : {&b_, &x_, &r_, &pa_, &pb_, &q_, &z_}) {
    PCFD_HIP_CHECK(hipMalloc(v, bytes));
    PCFD_HIP_CHECK(hipMemsetAsync(*v, 0, bytes, s_));
  }
  const size_t nt = (size_t)num_tiles(op.nx + 3, op.ny + 3) + 64;
  PCFD_HIP_CHECK(hipMalloc(&partial_, nt * 8 * sizeof(double)));
  PCFD_HIP_CHECK(hipMalloc(&counter_, 64 * sizeof(unsigned int)));
  PCFD_HIP_CHECK(hipMemsetAsync(counter_, 0, 64 * sizeof(unsigned int), s_));
  PCFD_HIP_CHECK(hipMalloc(&st_, sizeof(SolveState)));
  PCFD_HIP_CHECK(hipMemsetAsync(st_, 0, sizeof(SolveState), s_));
  PCFD_HIP_CHECK(hipHostMalloc(&st_host_, sizeof(SolveState), hipHostMallocDefault));
  PCFD_HIP_CHECK(hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
  return kOk;
}

// Uploads one level in precision T: vectors, and for coarse levels the row
// dictionary built from the dense class table (zeros dropped) plus the
// aggregation map towards the next level.
template <typename T>
static int upload_level(const MGLevelHost& h, const MGLevelHost* next, bool fine, LevelBuf& L,
                        hipStream_t s) {
  const size_t n = (size_t)h.nx * h.ny;
  for (void** v : {&L.b, &L.x, &L.y, &L.r, &L.t}) {
    PCFD_HIP_CHECK(hipMalloc(v, n * sizeof(T)));
    PCFD_HIP_CHECK(hipMemsetAsync(*v, 0, n * sizeof(T), s));
  }
  if (!fine) {
    const size_t S = h.di.size(), nc = h.cdinv.size();
    std::vector<int> rp(nc + 1, 0), lin;
    std::vector<int8_t> edi, edj;
    std::vector<T> val;
    int max_row = 0;
    for (size_t c = 0; c < nc; ++c) {
      for (size_t q = 0; q < S; ++q) {
        const double a = h.table[c * S + q];
        if (a == 0.0) continue;
        edi.push_back(h.di[q]);
        edj.push_back(h.dj[q]);
        lin.push_back((int)h.di[q] * h.ny + (int)h.dj[q]);
        val.push_back((T)a);
      }
      rp[c + 1] = (int)val.size();
      max_row = std::max(max_row, rp[c + 1] - rp[c]);
    }
    std::vector<T> cd(nc);
    for (size_t k = 0; k < nc; ++k) cd[k] = (T)h.cdinv[k];
    const size_t nnz = std::max<size_t>(val.size(), 1);
    PCFD_HIP_CHECK(hipMalloc(&L.cls, n * sizeof(uint16_t)));
    PCFD_HIP_CHECK(hipMalloc(&L.row_ptr, (nc + 1) * sizeof(int)));
    PCFD_HIP_CHECK(hipMalloc(&L.ent_lin, nnz * sizeof(int)));
    PCFD_HIP_CHECK(hipMalloc(&L.ent_di, nnz));
    PCFD_HIP_CHECK(hipMalloc(&L.ent_dj, nnz));
    PCFD_HIP_CHECK(hipMalloc(&L.ent_val, nnz * sizeof(T)));
    PCFD_HIP_CHECK(hipMalloc(&L.cdinv, nc * sizeof(T)));
    PCFD_HIP_CHECK(hipMalloc(&L.cactive, nc));
    PCFD_HIP_CHECK(hipMemcpy(L.cls, h.cls.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice));
    PCFD_HIP_CHECK(hipMemcpy(L.row_ptr, rp.data(), (nc + 1) * sizeof(int), hipMemcpyHostToDevice));
    PCFD_HIP_CHECK(hipMemcpy(L.ent_lin, lin.data(), val.size() * sizeof(int), hipMemcpyHostToDevice));
    PCFD_HIP_CHECK(hipMemcpy(L.ent_di, edi.data(), val.size(), hipMemcpyHostToDevice));
    PCFD_HIP_CHECK(hipMemcpy(L.ent_dj, edj.data(), val.size(), hipMemcpyHostToDevice));
    PCFD_HIP_CHECK(hipMemcpy(L.ent_val, val.data(), val.size() * sizeof(T), hipMemcpyHostToDevice));
    PCFD_HIP_CHECK(hipMemcpy(L.cdinv, cd.data(), nc * sizeof(T), hipMemcpyHostToDevice));
    PCFD_HIP_CHECK(hipMemcpy(L.cactive, h.cactive.data(), nc, hipMemcpyHostToDevice));
    L.ncls = (int)nc;
    L.nnz = (int)val.size();
    L.max_row = max_row;
    if (next) {
      std::vector<int> agg(n, -1);
      for (int i = 0; i < h.nx; ++i)
        for (int j = 0; j < h.ny; ++j) {
          const size_t k = (size_t)i * h.ny + j;
          if (h.active[k]) agg[k] = (i / h.fx) * next->ny + (j / h.fy);
        }
      PCFD_HIP_CHECK(hipMalloc(&L.agg, n * sizeof(int)));
      PCFD_HIP_CHECK(hipMemcpy(L.agg, agg.data(), n * sizeof(int), hipMemcpyHostToDevice));
    }
  }
  if (!h.coarse_inv.empty()) {
    PCFD_HIP_CHECK(hipMalloc(&L.cinv, h.coarse_inv.size() * sizeof(double)));
    PCFD_HIP_CHECK(hipMemcpy(L.cinv, h.coarse_inv.data(), h.coarse_inv.size() * sizeof(double),
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
      const MGLevelHost* next = (l + 1 < levels.size()) ? &levels[l + 1] : nullptr;
      LevelBuf L;
      L.nx = h.nx;
      L.ny = h.ny;
      L.n = h.nx * h.ny;
      L.fx = h.fx;
      L.fy = h.fy;
      L.weight = h.weight;
      int rc = f32 ? upload_level<float>(h, next, l == 0, L, s_)
                   : upload_level<double>(h, next, l == 0, L, s_);
      if (rc) return rc;
      L.lanes = L.max_row <= 16 ? 1 : L.max_row <= 64 ? 4 : L.max_row <= 256 ? 16 : 64;
      dst.push_back(L);
    }
  }
  PCFD_HIP_CHECK(hipStreamSynchronize(s_));
  return kOk;
}

int PoissonEngine::prepare_from_velocity(const double* u, const double* v, int ldu, int ldv,
                                         double dx, double dy, double dt) {
  hipLaunchKernelGGL(k_prepare, tiles_for(op_.nx, op_.ny), dim3(kBlock), 0, s_, op_, u, v, ldu,
                     ldv, dx, dy, dt, (const double*)nullptr, b_, partial_);
  PCFD_LAUNCH_CHECK();
  return fold(s_, partial_, ntiles(op_.nx, op_.ny), EpiPrepare{st_, singular_, n_fluid_});
}

int PoissonEngine::prepare_from_rhs(const double* rhs) {
  hipLaunchKernelGGL(k_prepare, tiles_for(op_.nx, op_.ny), dim3(kBlock), 0, s_, op_,
                     (const double*)nullptr, (const double*)nullptr, 0, 0, 1.0, 1.0, 1.0, rhs,
                     b_, partial_);
  PCFD_LAUNCH_CHECK();
  return fold(s_, partial_, ntiles(op_.nx, op_.ny), EpiPrepare{st_, singular_, n_fluid_});
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
                     (int)CM_FINAL, partial_, (const SolveState*)st_);
  PCFD_LAUNCH_CHECK();
  return fold(s_, partial_, ntiles(op_.nx, op_.ny),
              EpiCheck{st_, (int)CM_FINAL, cfg_.tol, cfg_.maxiter, singular_, n_fluid_});
}

int PoissonEngine::scatter_pressure(double* p, int ldp) {
  hipLaunchKernelGGL(k_scatter, tiles_for(op_.nx + 2, op_.ny + 2), dim3(kBlock), 0, s_, op_, x_,
                     (const SolveState*)st_, p, ldp, (double*)nullptr);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int PoissonEngine::write_interior(double* out) {
  hipLaunchKernelGGL(k_scatter, tiles_for(op_.nx + 2, op_.ny + 2), dim3(kBlock), 0, s_, op_, x_,
                     (const SolveState*)st_, (double*)nullptr, 0, out);
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
// PER is the number of iterations one body advances; MAXB bounds the bodies.
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

int PoissonEngine::start(int mode, double* r) {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  hipLaunchKernelGGL(k_zero_nonfluid, g, dim3(kBlock), 0, s_, op_, x_);
  hipLaunchKernelGGL(k_init, g, dim3(kBlock), 0, s_, op_, b_, (const double*)x_, r, mode,
                     (const SolveState*)st_, partial_);
  PCFD_LAUNCH_CHECK();
  return fold(s_, partial_, ntiles(op_.nx, op_.ny),
              EpiInit{st_, mode, cfg_.tol, cfg_.maxiter, n_proj_});
}

int PoissonEngine::run_cg() {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  const int nt = ntiles(op_.nx, op_.ny);
  if (int rc = start(IM_JPCG, r_)) return rc;
  double* p_old = pa_;
  double* p_new = pb_;
  PCFD_ITERATE(1, cfg_.maxiter, {
    hipLaunchKernelGGL(k_jpcg_dir, g, dim3(kBlock), 0, s_, op_, r_, p_old, p_new, q_, partial_,
                       (const SolveState*)st_);
    if (int rc = fold(s_, partial_, nt, EpiAlpha{st_})) return rc;
    hipLaunchKernelGGL(k_cg_update, g, dim3(kBlock), 0, s_, op_, x_, r_, p_new, q_,
                       (int)IM_JPCG, partial_, (const SolveState*)st_);
    if (int rc = fold(s_, partial_, nt, EpiUpdate{st_, (int)IM_JPCG, cfg_.maxiter, n_proj_}))
      return rc;
    std::swap(p_old, p_new);
  });
  return kOk;
}

int PoissonEngine::run_sor() {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  const int nt = ntiles(op_.nx, op_.ny);
  if (int rc = start(IM_RELAX, nullptr)) return rc;
  int sweeps_enqueued = 0;
  PCFD_ITERATE(5, (cfg_.maxiter + 4) / 5, {
    const int until = std::min(cfg_.maxiter, (sweeps_enqueued / 5 + 1) * 5);
    for (; sweeps_enqueued < until; ++sweeps_enqueued) {
      for (int colour = 0; colour < 2; ++colour) {
        hipLaunchKernelGGL(k_sor_colour, g, dim3(kBlock), 0, s_, op_, x_, b_, colour,
                           cfg_.omega, partial_, (const SolveState*)st_);
        if (int rc = fold(s_, partial_, nt, EpiSor{st_, colour, singular_, n_fluid_}))
          return rc;
      }
    }
    hipLaunchKernelGGL(k_check, g, dim3(kBlock), 0, s_, op_, x_, b_, (int)CM_RELAX, partial_,
                       (const SolveState*)st_);
    if (int rc = fold(s_, partial_, nt,
                      EpiCheck{st_, (int)CM_RELAX, cfg_.tol, cfg_.maxiter, singular_, n_fluid_}))
      return rc;
  });
  return kOk;
}

int PoissonEngine::run_jacobi() {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  const int nt = ntiles(op_.nx, op_.ny);
  if (int rc = start(IM_RELAX, nullptr)) return rc;
  int sweeps_enqueued = 0;
  double* cur = x_;
  double* nxt = pa_;
  PCFD_ITERATE(20, (cfg_.maxiter + 19) / 20, {
    const int until = std::min(cfg_.maxiter, (sweeps_enqueued / 20 + 1) * 20);
    for (; sweeps_enqueued < until; ++sweeps_enqueued) {
      hipLaunchKernelGGL(k_jacobi_sweep, g, dim3(kBlock), 0, s_, op_, cur, nxt, b_, cfg_.omega,
                         partial_, (const SolveState*)st_);
      if (int rc = fold(s_, partial_, nt, EpiJacobi{st_, singular_, n_fluid_})) return rc;
      std::swap(cur, nxt);
    }
    hipLaunchKernelGGL(k_check, g, dim3(kBlock), 0, s_, op_, cur, b_, (int)CM_RELAX, partial_,
                       (const SolveState*)st_);
    if (int rc = fold(s_, partial_, nt,
                      EpiCheck{st_, (int)CM_RELAX, cfg_.tol, cfg_.maxiter, singular_, n_fluid_}))
      return rc;
  });
  if (cur != x_) {
    PCFD_HIP_CHECK(
        hipMemcpyAsync(x_, cur, (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
  }
  return kOk;
}

template <typename T>
static ClassLevel<T> class_level(const LevelBuf& L, int px, int py) {
  ClassLevel<T> D;
  D.nx = L.nx;
  D.ny = L.ny;
  D.n = L.n;
  D.px = px;
  D.py = py;
  D.cls = L.cls;
  D.row_ptr = L.row_ptr;
  D.ent_lin = L.ent_lin;
  D.ent_di = L.ent_di;
  D.ent_dj = L.ent_dj;
  D.ent_val = (const T*)L.ent_val;
  D.cdinv = (const T*)L.cdinv;
  D.cactive = L.cactive;
  return D;
}

// Launches one row operation on a coarse level with the lane count chosen
// for its widest row.
template <typename T, int OP>
static void cls_rows(hipStream_t s, int lanes, const ClassLevel<T>& L, T w, const T* b,
                     const T* xin, const T* xc, const int* agg, T* o1, T* o2, T* o3,
                     const SolveState* st) {
  const long threads = (long)L.n * lanes;
  const dim3 grid((unsigned)((threads + kBlock - 1) / kBlock));
  switch (lanes) {
    case 1:
      hipLaunchKernelGGL((k_cls_rows<T, 1, OP>), grid, dim3(kBlock), 0, s, L, w, b, xin, xc,
                         agg, o1, o2, o3, st);
      break;
    case 4:
      hipLaunchKernelGGL((k_cls_rows<T, 4, OP>), grid, dim3(kBlock), 0, s, L, w, b, xin, xc,
                         agg, o1, o2, o3, st);
      break;
    case 16:
      hipLaunchKernelGGL((k_cls_rows<T, 16, OP>), grid, dim3(kBlock), 0, s, L, w, b, xin, xc,
                         agg, o1, o2, o3, st);
      break;
    default:
      hipLaunchKernelGGL((k_cls_rows<T, 64, OP>), grid, dim3(kBlock), 0, s, L, w, b, xin, xc,
                         agg, o1, o2, o3, st);
      break;
  }
}

// One V-cycle (multigrid.py:430-451) over the levels in precision T.  Level
// 0's right-hand side is lv[0].b; the result is left in lv[0].x (unprojected;
// the last sweep's sums give its active-set mean).
template <typename T>
int PoissonEngine::enqueue_vcycle_t(bool nonzero_start) {
  auto& lv = (sizeof(T) == 4) ? lv32_ : lv64_;
  const int nl = (int)lv.size();
  const int nu = cfg_.sweeps;
  const int red = (cfg_.kind == SK_MULTIGRID) ? SR_MG : SR_MGCG;
  const int nt0 = ntiles(op_.nx, op_.ny);
  const SolveState* cst = st_;
  FineLevelArgs<T> F{op_, (T)lv[0].weight};
  if (nl == 1) {
    LevelBuf& L = lv[0];
    hipLaunchKernelGGL(k_coarse_solve<T>, dim3(1), dim3(kBlock), L.n * sizeof(double), s_,
                       (const double*)L.cinv, L.n, (const T*)L.b, (T*)L.x, cst);
    hipLaunchKernelGGL(k_cycle_sums<T>, tiles_for(op_.nx, op_.ny), dim3(kBlock), 0, s_, op_,
                       (const T*)L.b, (const T*)L.x, partial_, cst);
    PCFD_LAUNCH_CHECK();
    return fold(s_, partial_, nt0, EpiCycle{st_, red, n_proj_});
  }
  // Downward leg.
  for (int l = 0; l < nl - 1; ++l) {
    LevelBuf& L = lv[l];
    LevelBuf& C = lv[l + 1];
    const bool nz = (l == 0) && nonzero_start;
    const dim3 g = tiles_for(L.nx, L.ny);
    if (l == 0) {
      for (int s = 0; s < nu; ++s) {
        const T* xin = (s == 0 && !nz) ? nullptr : (const T*)L.x;
        hipLaunchKernelGGL(k_smooth_fine<T>, g, dim3(kBlock), 0, s_, F, (const T*)L.b, xin,
                           (T*)L.y, (s == 0 && nz) ? 1 : 0, (int)SR_NONE, partial_, cst);
        std::swap(L.x, L.y);
      }
      hipLaunchKernelGGL(k_residual_fine<T>, g, dim3(kBlock), 0, s_, op_, (const T*)L.b,
                         (const T*)L.x, (T*)L.r, cst);
      hipLaunchKernelGGL(k_restrict_fine<T>, tiles_for(C.nx, C.ny), dim3(kBlock), 0, s_, F, L.fx,
                         L.fy, C.nx, C.ny, (const T*)L.r, (T*)C.b, cst);
    } else {
      const ClassLevel<T> D = class_level<T>(L, op_.px, op_.py);
      const T w = (T)L.weight;
      cls_rows<T, RO_SMOOTH0_RES>(s_, L.lanes, D, w, (const T*)L.b, nullptr, nullptr, nullptr,
                                  (T*)L.x, (T*)L.r, (T*)L.t, cst);
      if (nu > 1) {
        for (int s = 1; s < nu; ++s) {
          cls_rows<T, RO_SMOOTH>(s_, L.lanes, D, w, (const T*)L.b, (const T*)L.x, nullptr,
                                 nullptr, (T*)L.y, nullptr, nullptr, cst);
          std::swap(L.x, L.y);
        }
        cls_rows<T, RO_RES>(s_, L.lanes, D, w, (const T*)L.b, (const T*)L.x, nullptr, nullptr,
                            nullptr, (T*)L.r, (T*)L.t, cst);
      }
      // y = r - w A t into L.y, then the block sums into the next level's b
      cls_rows<T, RO_RESTRICT_Y>(s_, L.lanes, D, w, (const T*)L.r, (const T*)L.t, nullptr,
                                 nullptr, (T*)L.y, nullptr, nullptr, cst);
      hipLaunchKernelGGL(k_block_sum<T>, tiles_for(C.nx, C.ny), dim3(kBlock), 0, s_,
                         (const T*)L.y, L.nx, L.ny, L.fx, L.fy, C.nx, C.ny, (T*)C.b, cst);
    }
    PCFD_LAUNCH_CHECK();
  }
  {
    LevelBuf& L = lv[nl - 1];
    hipLaunchKernelGGL(k_coarse_solve<T>, dim3(1), dim3(kBlock), L.n * sizeof(double), s_,
                       (const double*)L.cinv, L.n, (const T*)L.b, (T*)L.x, cst);
    PCFD_LAUNCH_CHECK();
  }
  // Upward leg.
  for (int l = nl - 2; l >= 0; --l) {
    LevelBuf& L = lv[l];
    LevelBuf& C = lv[l + 1];
    const dim3 g = tiles_for(L.nx, L.ny);
    if (l == 0) {
      hipLaunchKernelGGL(k_prolong_fine<T>, g, dim3(kBlock), 0, s_, F, L.fx, L.fy, C.ny,
                         (const T*)C.x, (T*)L.x, cst);
      for (int s = 0; s < nu; ++s) {
        const int rmode = (s == nu - 1) ? red : (int)SR_NONE;
        hipLaunchKernelGGL(k_smooth_fine<T>, g, dim3(kBlock), 0, s_, F, (const T*)L.b,
                           (const T*)L.x, (T*)L.y, 0, rmode, partial_, cst);
        std::swap(L.x, L.y);
      }
      PCFD_LAUNCH_CHECK();
      if (int rc = fold(s_, partial_, nt0, EpiCycle{st_, red, n_proj_})) return rc;
    } else {
      const ClassLevel<T> D = class_level<T>(L, op_.px, op_.py);
      const T w = (T)L.weight;
      cls_rows<T, RO_PROLONG>(s_, L.lanes, D, w, nullptr, nullptr, (const T*)C.x, L.agg,
                              (T*)L.x, nullptr, nullptr, cst);
      for (int s = 0; s < nu; ++s) {
        cls_rows<T, RO_SMOOTH>(s_, L.lanes, D, w, (const T*)L.b, (const T*)L.x, nullptr, nullptr,
                               (T*)L.y, nullptr, nullptr, cst);
        std::swap(L.x, L.y);
      }
      PCFD_LAUNCH_CHECK();
    }
  }
  return kOk;
}

int PoissonEngine::enqueue_vcycle(bool nonzero_start) {
  if (cfg_.precision == 1) return enqueue_vcycle_t<float>(nonzero_start);
  return enqueue_vcycle_t<double>(nonzero_start);
}

int PoissonEngine::run_mgcg() {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  const int nt = ntiles(op_.nx, op_.ny);
  if (int rc = start(IM_MGCG, r_)) return rc;
  double* p_old = pa_;
  double* p_new = pb_;
  const bool f32 = cfg_.precision == 1;
  PCFD_ITERATE(1, cfg_.maxiter, {
    if (f32) {
      hipLaunchKernelGGL(k_project_in<float>, g, dim3(kBlock), 0, s_, op_, r_, (float*)lv32_[0].b,
                         (const SolveState*)st_);
    } else {
      hipLaunchKernelGGL(k_project_in<double>, g, dim3(kBlock), 0, s_, op_, r_,
                         (double*)lv64_[0].b, (const SolveState*)st_);
    }
    if (int rc = enqueue_vcycle(false)) return rc;
    if (f32) {
      hipLaunchKernelGGL(k_mgcg_dir<float>, g, dim3(kBlock), 0, s_, op_,
                         (const float*)lv32_[0].x, p_old, p_new, q_, partial_,
                         (const SolveState*)st_);
    } else {
      hipLaunchKernelGGL(k_mgcg_dir<double>, g, dim3(kBlock), 0, s_, op_,
                         (const double*)lv64_[0].x, p_old, p_new, q_, partial_,
                         (const SolveState*)st_);
    }
    if (int rc = fold(s_, partial_, nt, EpiAlpha{st_})) return rc;
    hipLaunchKernelGGL(k_cg_update, g, dim3(kBlock), 0, s_, op_, x_, r_, p_new, q_, (int)IM_MGCG,
                       partial_, (const SolveState*)st_);
    if (int rc = fold(s_, partial_, nt, EpiUpdate{st_, (int)IM_MGCG, cfg_.maxiter, n_proj_}))
      return rc;
    std::swap(p_old, p_new);
  });
  return kOk;
}

int PoissonEngine::run_multigrid() {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  const int nt = ntiles(op_.nx, op_.ny);
  if (int rc = start(IM_MG, nullptr)) return rc;
  LevelBuf& L0 = lv64_[0];
  // the hierarchy's level-0 vectors hold b and the iterate
  const size_t bytes = (size_t)n_ * sizeof(double);
  PCFD_HIP_CHECK(hipMemcpyAsync(L0.b, b_, bytes, hipMemcpyDeviceToDevice, s_));
  PCFD_HIP_CHECK(hipMemcpyAsync(L0.x, x_, bytes, hipMemcpyDeviceToDevice, s_));
  PCFD_ITERATE(1, cfg_.maxiter, {
    if (int rc = enqueue_vcycle_t<double>(true)) return rc;
    hipLaunchKernelGGL(k_check, g, dim3(kBlock), 0, s_, op_, (const double*)L0.x, b_,
                       (int)CM_MG, partial_, (const SolveState*)st_);
    if (int rc = fold(s_, partial_, nt,
                      EpiCheck{st_, (int)CM_MG, cfg_.tol, cfg_.maxiter, singular_, n_fluid_}))
      return rc;
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
  hipLaunchKernelGGL(k_active_sum, g, dim3(kBlock), 0, s_, op_, (const double*)r_, partial_);
  PCFD_LAUNCH_CHECK();
  if (int rc = fold(s_, partial_, ntiles(op_.nx, op_.ny), EpiOpen{st_, n_proj_})) return rc;
  const bool f32 = cfg_.precision == 1;
  if (f32) {
    hipLaunchKernelGGL(k_project_in<float>, g, dim3(kBlock), 0, s_, op_, r_, (float*)lv32_[0].b,
                       (const SolveState*)st_);
  } else {
    hipLaunchKernelGGL(k_project_in<double>, g, dim3(kBlock), 0, s_, op_, r_, (double*)lv64_[0].b,
                       (const SolveState*)st_);
  }
  if (int rc = enqueue_vcycle(false)) return rc;
  if (f32) {
    hipLaunchKernelGGL(k_cast_out<float>, g, dim3(kBlock), 0, s_, op_, (const float*)lv32_[0].x,
                       (const SolveState*)st_, z);
  } else {
    hipLaunchKernelGGL(k_cast_out<double>, g, dim3(kBlock), 0, s_, op_,
                       (const double*)lv64_[0].x, (const SolveState*)st_, z);
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
                           (int)IM_MGCG, partial_, (const SolveState*)st);
        B = n * (6 * 8 + 1);
        break;
      case 1:  // multigrid-CG direction + SpMV: read z, p_old + flags, write p, q
        hipLaunchKernelGGL(k_mgcg_dir<double>, g, dim3(kBlock), 0, s_, op_, (const double*)z_,
                           pa_, pb_, q_, partial_, (const SolveState*)st);
        B = n * (4 * 8 + 1);
        break;
      case 2: {  // fine damped-Jacobi sweep: read b, x + flags, write x'
        if (!have_mg) return kErrBadArg;
        FineLevelArgs<double> F{op_, lv64_[0].weight};
        hipLaunchKernelGGL(k_smooth_fine<double>, g, dim3(kBlock), 0, s_, F,
                           (const double*)b_, (const double*)x_, z_, 0, (int)SR_NONE, partial_,
                           (const SolveState*)st);
        B = n * (3 * 8 + 1);
        break;
      }
      case 3:  // fine residual: read b, x + flags, write r
        hipLaunchKernelGGL(k_residual_fine<double>, g, dim3(kBlock), 0, s_, op_,
                           (const double*)b_, (const double*)x_, z_, (const SolveState*)st);
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
