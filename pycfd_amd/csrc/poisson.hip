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
#include "poisson_dev.hpp"
#include "vcycle.hpp"

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <climits>
#include <cstring>
#include <map>
#include <mutex>

#include <hip/hip_cooperative_groups.h>

namespace pcfd {

// ----------------------------------------------------------------------------
// Reduction epilogues (run by k_fold in one thread after the partials are in)
// ----------------------------------------------------------------------------
enum InitMode : int { IM_JPCG = 0, IM_MGCG = 1, IM_RELAX = 2, IM_MG = 3 };
enum CheckMode : int { CM_RELAX = 0, CM_MG = 1, CM_FINAL = 2 };

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
      if (mode == IM_JPCG) st->rho_hist[0] = v[2];  // rho of iteration 0 (k_jpcg_dir)
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

struct EpiUpdateJ {  // Jacobi-CG: the last update's partials when no direction kernel follows
  static constexpr int K = 2;
  SolveState* st;
  int k, maxiter;  // updates done (the iteration count once this one is counted)
  __device__ void ops(int (&o)[K]) const { o[0] = o[1] = kSum; }
  __device__ bool active() const { return !st->done; }
  __device__ void operator()(const double (&v)[K]) const {
    const double rho_prev = st->rho_hist[(k - 1) & 1];
    st->iter = k;
    st->rr = v[0];
    st->rho_prev = rho_prev;
    st->rho = v[1];
    st->beta = v[1] / rho_prev;
    st->rho_hist[k & 1] = v[1];
    if (sqrt(v[0]) < st->atol) {
      st->done = 1;
      st->converged = 1;
    } else if (k >= maxiter) {
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
  SolveLog* log = nullptr;  // CM_FINAL: the solve's record, or none
  __device__ void ops(int (&o)[K]) const { o[0] = o[1] = kSum; }
  __device__ bool active() const { return mode == CM_FINAL || !st->done; }
  __device__ void operator()(const double (&v)[K]) const {
    const double bn = st->bnorm;
    const double res = sqrt(v[0]) / (bn > 1.0e-300 ? bn : 1.0);
    if (mode == CM_FINAL) {
      st->res_final = res;
      st->mean_final = (singular && nfluid > 0) ? v[1] / (double)nfluid : 0.0;
      if (log) {
        log->iter = st->iter;
        log->converged = st->converged;
        log->done = st->done;
        log->res = st->res;
        log->res_final = res;
      }
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
// The once-per-solve full-grid passes (right-hand side, start, final check)
// take kMultiTiles tiles per workgroup: every load of a thread's cells is
// issued before any of its arithmetic and stores, and one row of partials
// serves the workgroup's tiles (a quarter of the rows to fold).
// ----------------------------------------------------------------------------
constexpr int kMultiTiles = 4;

struct MultiTile {
  int blk;
  int i[kMultiTiles], j[kMultiTiles], k[kMultiTiles];
  bool in[kMultiTiles];  // false: past the grid (i, j, k then name cell 0)
};

__device__ __forceinline__ MultiTile multi_tile(int nx, int ny) {
  MultiTile m;
  m.blk = xcd_remap(blockIdx.x, gridDim.x);
  const int ty = tiles_y(ny), ntile = tiles_x(nx) * ty;
#pragma unroll
  for (int u = 0; u < kMultiTiles; ++u) {
    const int t = m.blk * kMultiTiles + u;
    const int i = (t / ty) * kTileX + (int)threadIdx.x / kTileY;
    const int j = (t % ty) * kTileY + (int)threadIdx.x % kTileY;
    m.in[u] = t < ntile && i < nx && j < ny;
    m.i[u] = m.in[u] ? i : 0;
    m.j[u] = m.in[u] ? j : 0;
    m.k[u] = m.i[u] * ny + m.j[u];
  }
  return m;
}

static inline int multi_blocks(int nx, int ny) {
  return (num_tiles(nx, ny) + kMultiTiles - 1) / kMultiTiles;
}

// y = (A x)[k] from the five values already loaded, in fine_apply's order.
__device__ __forceinline__ double apply5(const FaceC& c, uint8_t f, double d, double xc,
                                         double xl, double xr, double xb, double xt) {
  double y = 0.0;
  if (f & F_L) y += c.l * xl;
  if (f & F_B) y += c.b * xb;
  y += d * xc;
  if (f & F_T) y += c.t * xt;
  if (f & F_R) y += c.r * xr;
  return y;
}

// ----------------------------------------------------------------------------
// Right-hand side assembly (solver.py:659-661, pressure.py:300-308)
// ----------------------------------------------------------------------------
template <bool P2>
__global__ void __launch_bounds__(kBlock)
k_prepare(FineOp A, const double* __restrict__ u, const double* __restrict__ v, int ldu, int ldv,
          double dx, double dy, double idx, double idy, StepScalars T,
          const double* __restrict__ rhs, double* __restrict__ b, double* __restrict__ x,
          double* partial) {
  const double dt = T.get_dt();
  const MultiTile m = multi_tile(A.nx, A.ny);
  uint8_t f[kMultiTiles];
  double g0[kMultiTiles], g1[kMultiTiles], g2[kMultiTiles], g3[kMultiTiles];
#pragma unroll
  for (int q = 0; q < kMultiTiles; ++q) {
    const int i = m.i[q], j = m.j[q];
    f[q] = A.flags[m.k[q]];
    if (rhs) {
      g0[q] = rhs[m.k[q]];
      g1[q] = g2[q] = g3[q] = 0.0;
    } else {
      const int a = i + 1, c = j + 1;
      g0[q] = u[(size_t)(a + 1) * ldu + c];
      g1[q] = u[(size_t)a * ldu + c];
      g2[q] = v[(size_t)a * ldv + c + 1];
      g3[q] = v[(size_t)a * ldv + c];
    }
  }
  double s[1] = {0.0};
#pragma unroll
  for (int q = 0; q < kMultiTiles; ++q) {
    if (!m.in[q]) continue;
    const int i = m.i[q], j = m.j[q];
    double val;
    if (rhs) {
      val = g0[q];
    } else {
      double div;
      if (A.hxc) div = (g0[q] - g1[q]) / A.hxc[i] + (g2[q] - g3[q]) / A.hyc[j];
      else if constexpr (P2) div = (g0[q] - g1[q]) * idx + (g2[q] - g3[q]) * idy;
      else div = (g0[q] - g1[q]) / dx + (g2[q] - g3[q]) / dy;
      val = div / dt;
    }
    val = val * fine_row_scale(A, i, j) +
          ((f[q] & F_SOLV) ? dirichlet_rhs(A, fine_faces(A, i, j), i, j) : 0.0);
    if (!(f[q] & F_SOLV)) {
      val = 0.0;
      x[m.k[q]] = 0.0;  // the start guess holds zero off the solvable set
    }
    b[m.k[q]] = val;
    s[0] += val;
  }
  const int ops[1] = {kSum};
  block_partials<1>(s, ops, partial, m.blk);
}

// Solver start: b <- b - mean_b on fluid rows; r = b - A x (x = guess with
// non-fluid rows already zeroed); partials ||b||^2, ||r||^2 and, per mode,
// r.(r/d) (Jacobi-CG) or sum(r) (input projection of multigrid-CG).
__global__ void __launch_bounds__(kBlock)
k_init(FineOp A, double* __restrict__ b, const double* __restrict__ x, double* __restrict__ r,
       int mode, const SolveState* st, double* partial) {
  const MultiTile m = multi_tile(A.nx, A.ny);
  const double mb = st->mean_b;
  uint8_t f[kMultiTiles];
  double bv[kMultiTiles], xc[kMultiTiles], xl[kMultiTiles], xr[kMultiTiles], xb[kMultiTiles],
      xt[kMultiTiles];
#pragma unroll
  for (int u = 0; u < kMultiTiles; ++u) {
    const int k = m.k[u];
    const Nb nb = fine_neighbours(A, m.i[u], m.j[u]);
    f[u] = A.flags[k];
    bv[u] = b[k];
    xc[u] = x[k];
    xl[u] = x[nb.kL];
    xr[u] = x[nb.kR];
    xb[u] = x[nb.kB];
    xt[u] = x[nb.kT];
  }
  double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int u = 0; u < kMultiTiles; ++u) {
    if (!m.in[u]) continue;
    const int i = m.i[u], j = m.j[u], k = m.k[u];
    const double bk = (f[u] & F_SOLV) ? bv[u] - mb : 0.0;
    const FaceC fc = fine_faces(A, i, j);
    const double d = fine_diag(A, fc, f[u], i, j);
    const double rk = bk - apply5(fc, f[u], d, xc[u], xl[u], xr[u], xb[u], xt[u]);
    s[0] += bk * bk;
    s[1] += rk * rk;
    if (mode == IM_JPCG) s[2] += rk * (rk / d);
    if (mode == IM_MGCG && (f[u] & F_SOLV)) s[3] += rk;
    b[k] = bk;
    if (r) r[k] = rk;
  }
  const int ops[4] = {kSum, kSum, kSum, kSum};
  block_partials<4>(s, ops, partial, m.blk);
}

// ----------------------------------------------------------------------------
// Jacobi-preconditioned CG (scipy.sparse.linalg.cg iteration)
// ----------------------------------------------------------------------------
// Jacobi-CG state across launches: rho of iteration k (r.z after k updates)
// sits in st->rho_hist[k & 1]; the start writes rho_hist[0].  Iteration k's
// direction kernel reads rho_hist[(k - 1) & 1] and writes rho_hist[k & 1],
// so no workgroup reads what its own launch writes.
//
// The previous update's partials (r.r, r.z), folded by the next direction
// kernel (rows > 0), or already folded by EpiUpdateJ (rows = 0).  `k` is the
// iteration the direction kernel starts (updates done so far).
struct JpcgFold {
  const double* partial;
  int rows, k, maxiter;
};

// p = z + beta*p_old (z = r/d), q = A p, and p.q.  For k > 0 with rows > 0
// every workgroup first folds the previous update's partials itself in
// k_fold's fixed order: the update epilogue's arithmetic (iteration count,
// r.r, rho, beta = rho_k / rho_(k-1), the stopping test), bit for bit,
// without its launch; workgroup 0 records it.
__global__ void __launch_bounds__(kBlock)
k_jpcg_dir(FineOp A, const double* __restrict__ r, const double* __restrict__ p_old,
           double* __restrict__ p_new, double* __restrict__ q, double* partial,
           SolveState* st, JpcgFold jf) {
  if (st->done) return;
  Tile2D t;
  double s[1] = {0.0};
  const bool first = jf.k == 0;
  double beta = 0.0;
  if (!first && jf.rows > 0) {
    double v[2];
    const int op2[2] = {kSum, kSum};
    fold_partials<2>(v, op2, jf.partial, jf.rows);
    __shared__ double s_beta;
    __shared__ int s_done;
    if (threadIdx.x == 0) {
      const double rho_prev = st->rho_hist[(jf.k - 1) & 1];
      const double rho = v[1];
      int done = 0, conv = 0;
      if (sqrt(v[0]) < st->atol) {
        done = 1;
        conv = 1;
      } else if (jf.k >= jf.maxiter) {
        done = 1;
      }
      s_beta = rho / rho_prev;
      s_done = done;
      if (blockIdx.x == 0) {
        st->iter = jf.k;
        st->rr = v[0];
        st->rho_prev = rho_prev;
        st->rho = rho;
        st->beta = s_beta;
        st->rho_hist[jf.k & 1] = rho;
        if (done) {
          st->done = 1;
          st->converged = conv;
        }
      }
    }
    __syncthreads();
    if (s_done) return;
    beta = s_beta;
  } else if (!first) {
    beta = st->rho_hist[jf.k & 1] / st->rho_hist[(jf.k - 1) & 1];
  }
  if (cell_of_tile(A.nx, A.ny, t)) {
    const int i = t.i, j = t.j, k = i * A.ny + j;
    const uint8_t f = A.flags[k];
    const Nb nb = fine_neighbours(A, i, j);
    auto pv = [&](int m) -> double {
      const int mi = m / A.ny, mj = m - mi * A.ny;
      const double z = r[m] / fine_diag(A, A.flags[m], mi, mj);
      return first ? z : p_old[m] * beta + z;
    };
    const FaceC fc = fine_faces(A, i, j);
    const double d = fine_diag(A, fc, f, i, j);
    const double pk = pv(k);
    const double qk = fine_apply<double>(fc, f, nb, k, d, pv);
    p_new[k] = pk;
    q[k] = qk;
    s[0] = pk * qk;
  }
  const int ops[1] = {kSum};
  block_partials<1>(s, ops, partial, t.tile);
}

// Jacobi-CG update with the step length folded in: every workgroup first
// issues its loads, then folds the direction kernel's p.q partials itself in
// k_fold's fixed order (alpha = rho_k / p.q, the bits a fold launch would
// give), then x += alpha p, r -= alpha q, partials r.r and r.(r/d).
__global__ void __launch_bounds__(kBlock)
k_jpcg_update_f(FineOp A, double* __restrict__ x, double* __restrict__ r,
                const double* __restrict__ p, const double* __restrict__ q,
                const double* __restrict__ dir_partial, int dir_tiles, int k_iter,
                double* partial, SolveState* st) {
  if (st->done) return;
  Tile2D t;
  const bool in = cell_of_tile(A.nx, A.ny, t);
  const int i = in ? t.i : 0, j = in ? t.j : 0, k = i * A.ny + j;
  const uint8_t f = A.flags[k];
  const double xv = x[k], pv = p[k], rv = r[k], qv = q[k];
  double pq[1];
  const int op1[1] = {kSum};
  fold_partials<1>(pq, op1, dir_partial, dir_tiles);
  __shared__ double s_alpha;
  if (threadIdx.x == 0) s_alpha = st->rho_hist[k_iter & 1] / pq[0];
  __syncthreads();
  const double alpha = s_alpha;
  if (t.tile == 0 && threadIdx.x == 0) st->alpha = alpha;
  double s[2] = {0.0, 0.0};
  if (in) {
    const double xk = xv + alpha * pv;
    const double rk = rv - alpha * qv;
    x[k] = xk;
    r[k] = rk;
    s[0] = rk * rk;
    s[1] = rk * (rk / fine_diag(A, f, i, j));
  }
  const int ops[2] = {kSum, kSum};
  block_partials<2>(s, ops, partial, t.tile);
}

// The update with alpha from the state (no fold of its own): what
// pcfd_time_kernel times as the plain CG update.
__global__ void __launch_bounds__(kBlock)
k_jpcg_update(FineOp A, double* __restrict__ x, double* __restrict__ r,
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

// Multigrid-CG update with the step length folded in: every workgroup first
// issues the loads of its kCgTiles tiles, then folds the direction kernel's
// p.q partials itself in k_fold's fixed order (so each computes the alpha a
// separate fold launch would, bit for bit), then x += alpha p, r -= alpha q
// and one row of partials r.r, sum(r) over its tiles.  Saves the alpha fold's
// launch; workgroup 0 records alpha in the state.
constexpr int kCgTiles = 4;

__global__ void __launch_bounds__(kBlock)
k_cg_update(FineOp A, double* __restrict__ x, double* __restrict__ r,
              const double* __restrict__ p, const double* __restrict__ q,
              const double* __restrict__ dir_partial, int dir_tiles, double* __restrict__ partial,
              SolveState* st) {
  if (st->done) return;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int ty = tiles_y(A.ny), ntile = tiles_x(A.nx) * ty;
  double xv[kCgTiles], pv[kCgTiles], rv[kCgTiles], qv[kCgTiles];
  uint8_t fv[kCgTiles];
  int kk[kCgTiles];
  bool in[kCgTiles];
#pragma unroll
  for (int u = 0; u < kCgTiles; ++u) {
    const int t = blk * kCgTiles + u;
    const int i = (t / ty) * kTileX + (int)threadIdx.x / kTileY;
    const int j = (t % ty) * kTileY + (int)threadIdx.x % kTileY;
    in[u] = t < ntile && i < A.nx && j < A.ny;
    kk[u] = in[u] ? i * A.ny + j : 0;
    xv[u] = x[kk[u]];
    pv[u] = p[kk[u]];
    rv[u] = r[kk[u]];
    qv[u] = q[kk[u]];
    fv[u] = A.flags[kk[u]];
  }
  double pq[1];
  const int op1[1] = {kSum};
  fold_partials<1>(pq, op1, dir_partial, dir_tiles);
  __shared__ double s_alpha;
  if (threadIdx.x == 0) s_alpha = st->rho / pq[0];
  __syncthreads();
  const double alpha = s_alpha;
  if (blk == 0 && threadIdx.x == 0) st->alpha = alpha;
  double s[2] = {0.0, 0.0};
#pragma unroll
  for (int u = 0; u < kCgTiles; ++u) {
    if (!in[u]) continue;
    const double xk = xv[u] + alpha * pv[u];
    const double rk = rv[u] - alpha * qv[u];
    x[kk[u]] = xk;
    r[kk[u]] = rk;
    s[0] += rk * rk;
    if (fv[u] & F_SOLV) s[1] += rk;
  }
  const int op2[2] = {kSum, kSum};
  block_partials<2>(s, op2, partial, blk);
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
      const FaceC fc = fine_faces(A, i, j);
    const double d = fine_diag(A, fc, f, i, j);
      double off = 0.0;  // A_rb p_b in CSR column order
      if (f & F_L) off += fc.l * pv(nb.kL);
      if (f & F_B) off += fc.b * pv(nb.kB);
      if (f & F_T) off += fc.t * pv(nb.kT);
      if (f & F_R) off += fc.r * pv(nb.kR);
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
    const FaceC fc = fine_faces(A, i, j);
    const double d = fine_diag(A, fc, f, i, j);
    const double ap = fine_apply<double>(fc, f, nb, k, d, pv);
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
  const MultiTile mt = multi_tile(A.nx, A.ny);
  const double m = st->mean_x;
  const bool zero = (mode == CM_FINAL) && st->zero_rhs;
  // x at a cell less the pending shift on solvable cells (zero for a zero rhs)
  auto xv = [&](double xq, uint8_t fq) -> double {
    if (zero) return 0.0;
    return (fq & F_SOLV) ? xq - m : xq;
  };
  // Without a pending shift (the Krylov solvers' case) x needs no flags of the
  // neighbours: their loads are skipped (a uniform branch).
  const bool plain = m == 0.0 && !zero;
  uint8_t f[kMultiTiles], fl[kMultiTiles], fr[kMultiTiles], fb[kMultiTiles], ft[kMultiTiles];
  double bv[kMultiTiles], xc[kMultiTiles], xl[kMultiTiles], xr[kMultiTiles], xb[kMultiTiles],
      xt[kMultiTiles];
#pragma unroll
  for (int u = 0; u < kMultiTiles; ++u) {
    const int k = mt.k[u];
    const Nb nb = fine_neighbours(A, mt.i[u], mt.j[u]);
    f[u] = A.flags[k];
    fl[u] = fr[u] = fb[u] = ft[u] = 0;
    if (!plain) {
      fl[u] = A.flags[nb.kL];
      fr[u] = A.flags[nb.kR];
      fb[u] = A.flags[nb.kB];
      ft[u] = A.flags[nb.kT];
    }
    bv[u] = b[k];
    xc[u] = x[k];
    xl[u] = x[nb.kL];
    xr[u] = x[nb.kR];
    xb[u] = x[nb.kB];
    xt[u] = x[nb.kT];
  }
  double s[2] = {0.0, 0.0};
#pragma unroll
  for (int u = 0; u < kMultiTiles; ++u) {
    if (!mt.in[u]) continue;
    const int i = mt.i[u], j = mt.j[u];
    const FaceC fc = fine_faces(A, i, j);
    const double d = fine_diag(A, fc, f[u], i, j);
    const double rk = bv[u] - apply5(fc, f[u], d, xv(xc[u], f[u]), xv(xl[u], fl[u]),
                                     xv(xr[u], fr[u]), xv(xb[u], fb[u]), xv(xt[u], ft[u]));
    s[0] += rk * rk;
    if (f[u] & F_SOLV) s[1] += xv(xc[u], f[u]);
  }
  const int ops[2] = {kSum, kSum};
  block_partials<2>(s, ops, partial, mt.blk);
}

// Normalised pressure into the ghosted array (solver.py:659-665 with
// pressure.py:310-316): fluid rows minus the fluid mean, others zero, ghost
// ring zero.  The same values become the next warm-start guess.
__global__ void __launch_bounds__(kBlock)
k_scatter(FineOp A, double* __restrict__ x, const SolveState* st, double* __restrict__ p,
          int ldp, double* __restrict__ out, WarmHistory wh) {
  const Tile2D t = tile_coords(A.ny + 2);
  const int i = t.i, j = t.j;
  if (i > A.nx + 1 || j > A.ny + 1) return;
  const bool interior = i >= 1 && i <= A.nx && j >= 1 && j <= A.ny;
  // the history slot this solution goes to (chosen by the guess kernel)
  double* hist = wh.slots ? wh.slots + (size_t)wh.idx[2] * A.nx * A.ny : nullptr;
  double val = 0.0;
  if (interior) {
    const int k = (i - 1) * A.ny + (j - 1);
    if (A.flags[k] & F_SOLV) {
      const double xk = st->zero_rhs ? 0.0 : x[k] - st->mean_x;
      val = xk - st->mean_final;
    }
    x[k] = val;
    if (out) out[k] = val;
    if (hist) hist[k] = val;
  }
  if (p) p[(size_t)i * ldp + j] = val;
  if (wh.slots && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    wh.idx[0] = wh.idx[2];  // the newest solution and the count, for the next guess
    wh.idx[1] = wh.idx[3];
  }
}

// Warm start by polynomial extrapolation in time: the guess of this solve is
// sum_j w_j x_(n-j) over the last order+1 solutions of the same substep (a
// ring of kWarmSlots), with the weights of equal steps, w_j = (-1)^j
// C(order + 1, j + 1): 2, -1; 3, -3, 1; 4, -6, 4, -1; ...  Fewer stored
// solutions lower the order; with one or none the guess is left as it is
// (the previous solve's pressure).  The ring's state is read from idx[0..1]
// (newest slot, count) and the next state written to idx[2..3], which only
// the scatter reads (and copies back): no launch reads what it writes.
template <int O>
__device__ __forceinline__ double extrapolate(const double* __restrict__ slots, int head, int n,
                                              int k) {
  double v[O + 1];
#pragma unroll
  for (int j = 0; j <= O; ++j)  // all loads first, then the sum in a fixed order
    v[j] = slots[(size_t)((head - j + kWarmSlots) % kWarmSlots) * n + k];
  double g = 0.0, c = 1.0;
#pragma unroll
  for (int j = 0; j <= O; ++j) {
    c = c * (O + 1 - j) / (j + 1);  // C(O + 1, j + 1), an exact integer
    g += ((j & 1) ? -c : c) * v[j];
  }
  return g;
}

__global__ void __launch_bounds__(kBlock)
k_warm_guess(int n, double* __restrict__ x, WarmHistory wh, int order) {
  const int head = wh.idx[0], cnt = wh.idx[1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    wh.idx[2] = (head + 1) % kWarmSlots;
    wh.idx[3] = min(cnt + 1, kWarmSlots);
  }
  const int o = min(order, cnt - 1);
  if (o < 1) return;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  static_assert(kWarmMaxOrder == 4, "one case per order");
  double g;
  switch (o) {
    case 1: g = extrapolate<1>(wh.slots, head, n, k); break;
    case 2: g = extrapolate<2>(wh.slots, head, n, k); break;
    case 3: g = extrapolate<3>(wh.slots, head, n, k); break;
    default: g = extrapolate<4>(wh.slots, head, n, k); break;
  }
  x[k] = g;
}

// ----------------------------------------------------------------------------
// Multigrid helpers (the level kernels themselves are in vcycle.hip)
// ----------------------------------------------------------------------------
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

// ----------------------------------------------------------------------------
// Multigrid, levels >= 1: stored Galerkin operators (row dictionary), one
// row per LPR lanes; the lanes split the row's nonzeros and combine with
// cross-lane shuffles.
// ----------------------------------------------------------------------------
enum RowOp : int {
  RO_SMOOTH0_RES = 0,  // x = w D^-1 b (zero start), r = b - A x, t = D^-1 r
  RO_SMOOTH = 1,       // x_out = x + w D^-1 (b - A x)
  RO_RES = 2,          // r = b - A x, t = D^-1 r
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

// One row operation at global lane gid (LPR lanes per row).  The row's entry
// range, 1/diag and active flag come from per-cell arrays, so the only
// dependent loads are entry -> neighbour value.
template <typename T, int LPR, int OP>
__device__ __forceinline__ void cls_rows_at(const ClassLevel<T>& L, T w, const T* __restrict__ b,
                                            const T* __restrict__ xin, const T* __restrict__ xc,
                                            const int* __restrict__ agg, T* __restrict__ o1,
                                            T* __restrict__ o2, T* __restrict__ o3, int gid) {
  (void)agg;
  const int row = gid / LPR;
  const int lane = gid % LPR;
  const bool valid = row < L.n;
  T acc = (T)0;
  const int i = valid ? row / L.ny : 0, j = valid ? row - i * L.ny : 0;
  // P_tent x_c at cell (qi, qj): its aggregate's value on active cells
  auto prolong_at = [&](int qi, int qj) -> T {
    const T ev = xc[(qi / L.fx) * L.cny + qj / L.fy];
    return L.ract[qi * L.ny + qj] ? ev : (T)0;
  };
  if (valid) {
    const int2 rr = L.rr[row];
    auto xval = [&](int e) -> T {
      if (OP == RO_PROLONG) {
        int qi = i + L.ent_di[e], qj = j + L.ent_dj[e];
        if (L.px) qi = (qi % L.nx + L.nx) % L.nx;
        if (L.py) qj = (qj % L.ny + L.ny) % L.ny;
        return prolong_at(qi, qj);
      }
      const int q = cls_nb(L, row, e);
      if (OP == RO_SMOOTH0_RES) return w * L.rdinv[q] * b[q];
      return xin[q];
    };
    int e = rr.x + lane;
    // Four entries per trip with their loads issued together; the sum keeps
    // the entry order of the plain loop.
    for (; e + 3 * LPR < rr.y; e += 4 * LPR) {
      const T a0 = L.ent_val[e], a1 = L.ent_val[e + LPR];
      const T a2 = L.ent_val[e + 2 * LPR], a3 = L.ent_val[e + 3 * LPR];
      const T x0 = xval(e), x1 = xval(e + LPR), x2 = xval(e + 2 * LPR), x3 = xval(e + 3 * LPR);
      acc += a0 * x0;
      acc += a1 * x1;
      acc += a2 * x2;
      acc += a3 * x3;
    }
    for (; e < rr.y; e += LPR) acc += L.ent_val[e] * xval(e);
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
  if (!valid || lane != 0) return;
  const T dinv = L.rdinv[row];
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
  } else {  // RO_PROLONG
    o1[row] = o1[row] + prolong_at(i, j) - w * dinv * acc;
  }
}

template <typename T, int LPR, int OP>
__global__ void __launch_bounds__(kBlock)
k_cls_rows(ClassLevel<T> L, T w, const T* __restrict__ b, const T* __restrict__ xin,
           const T* __restrict__ xc, const int* __restrict__ agg, T* __restrict__ o1,
           T* __restrict__ o2, T* __restrict__ o3, const SolveState* st) {
  if (st->done) return;
  cls_rows_at<T, LPR, OP>(L, w, b, xin, xc, agg, o1, o2, o3, blockIdx.x * blockDim.x + threadIdx.x);
}

// Restriction in one pass: b_c[I] = sum over the cells a of block I of
// (r - w A t)[a] on active rows (t = D^-1 r); the block's cells are summed in
// x-major order (P_tent^T y with y the P-smoothed residual).  A group of
// 4 * LPR lanes serves one coarse cell, LPR lanes per fine row of its block
// (up to 2 x 2 cells), so the rows of a block are evaluated side by side.
template <typename T, int LPR>
__device__ __forceinline__ void cls_restrict_at(const ClassLevel<T>& L, T w,
                                                const T* __restrict__ r, const T* __restrict__ t,
                                                int fx, int fy, int cnx, int cny,
                                                T* __restrict__ bc, int gid) {
  static_assert(4 * LPR <= kWave, "a coarse cell's lanes must share a wavefront");
  constexpr int G = 4 * LPR;
  const int crow = gid / G;
  const int sub = (gid % G) / LPR;
  const int lane = gid % LPR;
  const bool valid = crow < cnx * cny;
  const int I = valid ? crow / cny : 0, J = valid ? crow - I * cny : 0;
  const int a2 = sub / fy, c2 = sub - a2 * fy;
  const int i = I * fx + a2, j = J * fy + c2;
  const bool in = valid && sub < fx * fy && i < L.nx && j < L.ny;
  const int row = in ? i * L.ny + j : 0;
  const bool act = in && L.ract[row];
  T acc = (T)0;
  if (act) {
    const int2 rr = L.rr[row];
    const int e1 = rr.y;
    int e = rr.x + lane;
    for (; e + 3 * LPR < e1; e += 4 * LPR) {
      const int q0 = cls_nb(L, row, e), q1 = cls_nb(L, row, e + LPR);
      const int q2 = cls_nb(L, row, e + 2 * LPR), q3 = cls_nb(L, row, e + 3 * LPR);
      const T v0 = L.ent_val[e], v1 = L.ent_val[e + LPR];
      const T v2 = L.ent_val[e + 2 * LPR], v3 = L.ent_val[e + 3 * LPR];
      const T x0 = t[q0], x1 = t[q1], x2 = t[q2], x3 = t[q3];
      acc += v0 * x0;
      acc += v1 * x1;
      acc += v2 * x2;
      acc += v3 * x3;
    }
    for (; e < e1; e += LPR) acc += L.ent_val[e] * t[cls_nb(L, row, e)];
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
  const T y = act ? r[row] - w * acc : (T)0;
  // Sum the block's rows in x-major order (slot order).
  const int g0 = (int)(threadIdx.x % kWave) - (int)(threadIdx.x % G);
  T sum = (T)0;
#pragma unroll
  for (int q = 0; q < 4; ++q) sum += __shfl(y, g0 + q * LPR, kWave);
  if (valid && sub == 0 && lane == 0) bc[crow] = sum;
}

template <typename T, int LPR>
__global__ void __launch_bounds__(kBlock)
k_cls_restrict(ClassLevel<T> L, T w, const T* __restrict__ r, const T* __restrict__ t, int fx,
               int fy, int cnx, int cny, T* __restrict__ bc, const SolveState* st) {
  if (st->done) return;
  cls_restrict_at<T, LPR>(L, w, r, t, fx, fy, cnx, cny, bc, blockIdx.x * blockDim.x + threadIdx.x);
}

// The same for rows too long for four of them to share a wavefront: the
// block's rows one after another, LPR lanes each.
template <typename T, int LPR>
__device__ __forceinline__ void cls_restrict_serial_at(const ClassLevel<T>& L, T w,
                                                       const T* __restrict__ r,
                                                       const T* __restrict__ t, int fx, int fy,
                                                       int cnx, int cny, T* __restrict__ bc,
                                                       int gid) {
  const int crow = gid / LPR;
  const int lane = gid % LPR;
  const bool valid = crow < cnx * cny;
  const int I = valid ? crow / cny : 0, J = valid ? crow - I * cny : 0;
  T sum = (T)0;
  for (int a2 = 0; a2 < fx; ++a2) {
    const int i = I * fx + a2;
    for (int c2 = 0; c2 < fy; ++c2) {
      const int j = J * fy + c2;
      const bool in = valid && i < L.nx && j < L.ny;
      const int row = in ? i * L.ny + j : 0;
      const bool act = in && L.ract[row];
      T acc = (T)0;
      if (act) {
        const int2 rr = L.rr[row];
        for (int e = rr.x + lane; e < rr.y; e += LPR) acc += L.ent_val[e] * t[cls_nb(L, row, e)];
      }
#pragma unroll
      for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
      if (act) sum += r[row] - w * acc;
    }
  }
  if (valid && lane == 0) bc[crow] = sum;
}

template <typename T, int LPR>
__global__ void __launch_bounds__(kBlock)
k_cls_restrict_serial(ClassLevel<T> L, T w, const T* __restrict__ r, const T* __restrict__ t,
                      int fx, int fy, int cnx, int cny, T* __restrict__ bc, const SolveState* st) {
  if (st->done) return;
  cls_restrict_serial_at<T, LPR>(L, w, r, t, fx, fy, cnx, cny, bc,
                                 blockIdx.x * blockDim.x + threadIdx.x);
}

// ----------------------------------------------------------------------------
// The same row passes with every load of a lane issued up front.  A coarse
// pass is short and bound by its chain of dependent memory round trips (row
// -> entry -> neighbour), not by bandwidth, so here each lane holds all of
// its entries (at most kRowNE) in registers: the row's own values, its entry
// range and the entries of the level's most frequent row (`spec_e0/e1`, the
// interior rows) go out together, the neighbour values follow in one batch,
// and only rows of another class pay one more round trip for their entries.
// Per-lane sums run in the entry order of cls_rows_at at the same lane count.
// ----------------------------------------------------------------------------
constexpr int kRowNE = 8;

// Entries e0, e0 + LPR, ... of the row at cell (i, j) whose range ends at
// e1: coefficient, neighbour cell q and, with AGG, the neighbour's aggregate
// on the next level.  Lanes past the row's end load its last entry (every
// row has its diagonal), so all loads are unconditional and go out in one
// batch; the callers skip those slots.  PER (a periodic level) wraps both
// axes: on a non-periodic axis the row's own entries stay inside the grid,
// where the wrap is the identity.  Aggregation factors are 1 or 2.
template <typename T, int LPR, bool AGG, bool PER>
__device__ __forceinline__ void row_entries(const ClassLevel<T>& L, int row, int i, int j, int e0,
                                            int e1, int (&q)[kRowNE], int (&a)[kRowNE],
                                            T (&v)[kRowNE]) {
#pragma unroll
  for (int k = 0; k < kRowNE; ++k) {
    const int e = min(e0 + k * LPR, e1 - 1);
    v[k] = L.ent_val[e];
    int qi = 0, qj = 0;
    if (PER) {
      qi = i + (int)L.ent_di[e];
      qj = j + (int)L.ent_dj[e];
      qi = (qi % L.nx + L.nx) % L.nx;
      qj = (qj % L.ny + L.ny) % L.ny;
      q[k] = qi * L.ny + qj;
    } else {
      q[k] = row + L.ent_lin[e];
      if (AGG) {
        qi = q[k] / L.ny;
        qj = q[k] - qi * L.ny;
      }
    }
    a[k] = AGG ? (qi >> (L.fx >> 1)) * L.cny + (qj >> (L.fy >> 1)) : 0;
  }
}

// Holds a value in registers at this point of the program, so the loads it
// comes from are issued before what follows (the compiler may otherwise sink
// them into the branch that uses them, behind the loads that branch waits for).
template <typename V>
__device__ __forceinline__ void issue_here(V& v) {
  asm volatile("" : "+v"(v));
}

template <typename T>
__device__ __forceinline__ void issue_entries(int2& rr, int (&q)[kRowNE], T (&v)[kRowNE]) {
  issue_here(rr.x);
  issue_here(rr.y);
#pragma unroll
  for (int k = 0; k < kRowNE; ++k) {
    issue_here(q[k]);
    issue_here(v[k]);
  }
}

template <typename T, int LPR, int OP, bool PER>
__global__ void __launch_bounds__(kBlock)
k_cls_rows_reg(ClassLevel<T> L, T w, const T* __restrict__ b, const T* __restrict__ xin,
               const T* __restrict__ xc, T* __restrict__ o1, T* __restrict__ o2,
               T* __restrict__ o3, const SolveState* st) {
  constexpr bool AGG = OP == RO_PROLONG;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = gid % LPR;
  const bool valid = gid / LPR < L.n;
  const int row = valid ? gid / LPR : L.n - 1;  // lanes past the grid load row n-1, store nothing
  const int i = row / L.ny, j = row - i * L.ny;
  int2 rr = L.rr[row];
  const T dinv = L.rdinv[row];
  const T bk = OP == RO_PROLONG ? (T)0 : b[row];
  const T xk = OP == RO_SMOOTH ? xin[row] : (T)0;
  T ok = (T)0, own = (T)0;
  bool own_act = false;
  if (OP == RO_PROLONG) {
    ok = o1[row];
    own = xc[(i >> (L.fx >> 1)) * L.cny + (j >> (L.fy >> 1))];
    own_act = L.ract[row] != 0;
  }
  int q[kRowNE], a[kRowNE];
  T v[kRowNE];
  int done = st->done;
  row_entries<T, LPR, AGG, PER>(L, row, i, j, L.spec_e0 + lane, L.spec_e1, q, a, v);
  issue_entries(rr, q, v);
  issue_here(done);
  if (done) return;
  if (rr.x != L.spec_e0) row_entries<T, LPR, AGG, PER>(L, row, i, j, rr.x + lane, rr.y, q, a, v);
  T xv[kRowNE];
#pragma unroll
  for (int k = 0; k < kRowNE; ++k) {
    if (OP == RO_PROLONG) {
      // P_tent x_c at the neighbour: an off-diagonal entry couples active
      // cells only (an inactive cell's row and column are the identity's),
      // so only the diagonal needs an active flag, the row's own
      const T ev = xc[a[k]];
      xv[k] = (q[k] != row || own_act) ? ev : (T)0;
    } else if (OP == RO_SMOOTH0_RES) {
      xv[k] = w * L.rdinv[q[k]] * b[q[k]];
    } else {
      xv[k] = xin[q[k]];
    }
  }
  T acc = (T)0;
#pragma unroll
  for (int k = 0; k < kRowNE; ++k) {
    const T sum = acc + v[k] * xv[k];
    acc = rr.x + lane + k * LPR < rr.y ? sum : acc;
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
  if (!valid || lane != 0) return;
  if (OP == RO_SMOOTH0_RES) {
    const T rk = bk - acc;
    o1[row] = w * dinv * bk;
    o2[row] = rk;
    o3[row] = dinv * rk;
  } else if (OP == RO_SMOOTH) {
    o1[row] = xk + w * dinv * (bk - acc);
  } else if (OP == RO_RES) {
    const T rk = bk - acc;
    o2[row] = rk;
    o3[row] = dinv * rk;
  } else {  // RO_PROLONG
    o1[row] = ok + (own_act ? own : (T)0) - w * dinv * acc;
  }
}

// Restriction with register-resident rows: 4 * LPR lanes per coarse cell, LPR
// per fine row of its block.  LPR <= 16: the cell inside one wavefront, its
// rows summed in slot order by shuffles, as cls_restrict_at; LPR = 32, 64: the
// cell spans wavefronts of one workgroup and its rows are summed through LDS.
template <typename T, int LPR, bool PER>
__global__ void __launch_bounds__(kBlock)
k_cls_restrict_reg(ClassLevel<T> L, T w, const T* __restrict__ r, const T* __restrict__ t,
                   int fx, int fy, int cnx, int cny, T* __restrict__ bc, const SolveState* st) {
  constexpr int G = 4 * LPR;
  static_assert(G <= kBlock && kBlock % G == 0, "a coarse cell's lanes must share a workgroup");
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int crow = gid / G;
  const int sub = (gid % G) / LPR;
  const int lane = gid % LPR;
  const bool valid = crow < cnx * cny;
  const int I = valid ? crow / cny : 0, J = valid ? crow - I * cny : 0;
  const int a2 = sub / fy, c2 = sub - a2 * fy;
  const int i = I * fx + a2, j = J * fy + c2;
  const bool in = valid && sub < fx * fy && i < L.nx && j < L.ny;
  const int row = in ? i * L.ny + j : 0;
  const int ri = in ? i : 0, rj = in ? j : 0;
  int2 rr = L.rr[row];
  int ra = L.ract[row];
  const T rk = r[row];
  int q[kRowNE], a[kRowNE];
  T v[kRowNE];
  int done = st->done;
  row_entries<T, LPR, false, PER>(L, row, ri, rj, L.spec_e0 + lane, L.spec_e1, q, a, v);
  issue_entries(rr, q, v);
  issue_here(ra);
  issue_here(done);
  if (done) return;
  const bool act = in & (ra != 0);
  if (rr.x != L.spec_e0)
    row_entries<T, LPR, false, PER>(L, row, ri, rj, rr.x + lane, rr.y, q, a, v);
  T tv[kRowNE];
#pragma unroll
  for (int k = 0; k < kRowNE; ++k) tv[k] = t[q[k]];
  T acc = (T)0;
#pragma unroll
  for (int k = 0; k < kRowNE; ++k) {
    const T sum = acc + v[k] * tv[k];
    acc = rr.x + lane + k * LPR < rr.y ? sum : acc;
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
  const T y = act ? rk - w * acc : (T)0;
  T sum = (T)0;
  if constexpr (G <= kWave) {
    const int g0 = (int)(threadIdx.x % kWave) - (int)(threadIdx.x % G);
#pragma unroll
    for (int s = 0; s < 4; ++s) sum += __shfl(y, g0 + s * LPR, kWave);
  } else {  // a coarse cell spans several wavefronts: combine through LDS
    __shared__ T ys[kBlock / LPR];
    if (lane == 0) ys[threadIdx.x / LPR] = y;
    __syncthreads();
    const int c0 = (int)(threadIdx.x / G) * 4;
#pragma unroll
    for (int s = 0; s < 4; ++s) sum += ys[c0 + s];
  }
  if (valid && sub == 0 && lane == 0) bc[crow] = sum;
}

// ----------------------------------------------------------------------------
// The deepest stored levels and the dense coarsest solve in one cooperative
// launch.  These levels are small (tens of thousands of rows at most) and their
// passes are bound by launch latency, so the down leg, the coarsest solve and
// the up leg run as phases of one grid, separated by grid-wide barriers,
// instead of as a dozen dependent launches.  Every phase is the same per-row
// arithmetic as the pass-by-pass kernels, spread over the grid by striding.
// ----------------------------------------------------------------------------
constexpr int kMaxChain = 8;

template <typename T>
struct ChainLevel {
  ClassLevel<T> L;
  T w;
  int lanes, fx, fy;
  T *b, *x, *y, *r, *t;
  const int* agg;
};

template <typename T>
struct Chain {
  int n;  // stored levels in the chain, finest first
  int nu;
  ChainLevel<T> lv[kMaxChain];
  int cn;             // coarsest unknowns
  const double* cinv; // dense pseudo-inverse, cn x cn
  T* cb;              // coarsest right-hand side (written by the last restriction)
  T* cy;              // coarsest solution
  int cnx, cny;       // coarsest grid
};

template <typename T, int OP>
__device__ __forceinline__ void chain_rows(const ChainLevel<T>& c, const T* b, const T* xin,
                                           const T* xc, T* o1, T* o2, T* o3, int tid,
                                           int nth) {
  const int n = c.L.n;
  switch (c.lanes) {
    case 1:
      for (int g = tid; g < n; g += nth)
        cls_rows_at<T, 1, OP>(c.L, c.w, b, xin, xc, c.agg, o1, o2, o3, g);
      break;
    case 4:
      for (int g = tid; g < n * 4; g += nth)
        cls_rows_at<T, 4, OP>(c.L, c.w, b, xin, xc, c.agg, o1, o2, o3, g);
      break;
    case 16:
      for (int g = tid; g < n * 16; g += nth)
        cls_rows_at<T, 16, OP>(c.L, c.w, b, xin, xc, c.agg, o1, o2, o3, g);
      break;
    default:
      for (int g = tid; g < n * 64; g += nth)
        cls_rows_at<T, 64, OP>(c.L, c.w, b, xin, xc, c.agg, o1, o2, o3, g);
      break;
  }
}

template <typename T>
__device__ __forceinline__ void chain_restrict(const ChainLevel<T>& c, int cnx, int cny, T* bc,
                                               int tid, int nth) {
  const int cells = cnx * cny;
  switch (c.lanes) {
    case 1:
      for (int g = tid; g < cells * 4; g += nth)
        cls_restrict_at<T, 1>(c.L, c.w, c.r, c.t, c.fx, c.fy, cnx, cny, bc, g);
      break;
    case 4:
      for (int g = tid; g < cells * 16; g += nth)
        cls_restrict_at<T, 4>(c.L, c.w, c.r, c.t, c.fx, c.fy, cnx, cny, bc, g);
      break;
    case 16:
      for (int g = tid; g < cells * 64; g += nth)
        cls_restrict_at<T, 16>(c.L, c.w, c.r, c.t, c.fx, c.fy, cnx, cny, bc, g);
      break;
    default:
      for (int g = tid; g < cells * 64; g += nth)
        cls_restrict_serial_at<T, 64>(c.L, c.w, c.r, c.t, c.fx, c.fy, cnx, cny, bc, g);
      break;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_chain(Chain<T> C, const SolveState* st) {
  // `done` is read once by every workgroup before any barrier: uniform.
  if (st->done) return;
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  extern __shared__ double cshared[];
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int nth = gridDim.x * blockDim.x;
  const int nu = C.nu;
  bool pre_in_y[kMaxChain];  // where each level's pre-smoothed iterate sits
  // Down leg.
  for (int l = 0; l < C.n; ++l) {
    const ChainLevel<T>& c = C.lv[l];
    const bool last = l + 1 == C.n;
    T* bnext = last ? C.cb : C.lv[l + 1].b;
    const int cnx = last ? C.cnx : C.lv[l + 1].L.nx;
    const int cny = last ? C.cny : C.lv[l + 1].L.ny;
    chain_rows<T, RO_SMOOTH0_RES>(c, c.b, nullptr, nullptr, c.x, c.r, c.t, tid, nth);
    grid.sync();
    T* cur = c.x;
    T* oth = c.y;
    if (nu > 1) {
      for (int sw = 1; sw < nu; ++sw) {
        chain_rows<T, RO_SMOOTH>(c, c.b, cur, nullptr, oth, nullptr, nullptr, tid, nth);
        grid.sync();
        T* tmp = cur;
        cur = oth;
        oth = tmp;
      }
      chain_rows<T, RO_RES>(c, c.b, cur, nullptr, nullptr, c.r, c.t, tid, nth);
      grid.sync();
    }
    pre_in_y[l] = cur == c.y;
    chain_restrict<T>(c, cnx, cny, bnext, tid, nth);
    grid.sync();
  }
  // Coarsest level: y = C b, one wavefront per row, b staged in LDS.
  for (int k = threadIdx.x; k < C.cn; k += blockDim.x) cshared[k] = (double)C.cb[k];
  __syncthreads();
  {
    const int lane = threadIdx.x % kWave;
    const int wv = tid / kWave, nwv = nth / kWave;
    for (int row = wv; row < C.cn; row += nwv) {
      const double* crow = C.cinv + (size_t)row * C.cn;
      double acc = 0.0;
#pragma unroll 8
      for (int k = lane; k < C.cn; k += kWave) acc += crow[k] * cshared[k];
      acc = wave_sum(acc);
      if (lane == 0) C.cy[row] = (T)acc;
    }
  }
  grid.sync();
  // Up leg: prolongate into the pre-smoothed iterate, then smooth; the
  // alternation leaves the corrected iterate in y (see the down leg).
  for (int l = C.n - 1; l >= 0; --l) {
    const ChainLevel<T>& c = C.lv[l];
    const T* xc = (l + 1 == C.n) ? C.cy : C.lv[l + 1].y;
    T* cur = pre_in_y[l] ? c.y : c.x;
    T* oth = pre_in_y[l] ? c.x : c.y;
    chain_rows<T, RO_PROLONG>(c, nullptr, nullptr, xc, cur, nullptr, nullptr, tid, nth);
    grid.sync();
    for (int sw = 0; sw < nu; ++sw) {
      chain_rows<T, RO_SMOOTH>(c, c.b, cur, nullptr, oth, nullptr, nullptr, tid, nth);
      grid.sync();
      T* tmp = cur;
      cur = oth;
      oth = tmp;
    }
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
  const FaceC fc = fine_faces(A, i, j);
  y[k] = fine_apply<double>(fc, f, nb, k, fine_diag(A, fc, f, i, j), xv);
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
static int fold(hipStream_t s, double* partial, int n, const Epi& e,
                const LaunchClock& clk = LaunchClock{}) {
  PCFD_LAUNCH_TIMED(clk, k_fold<Epi>, dim3(1), dim3(kBlock), 0, s, (const double*)partial, n, e);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

PoissonEngine::~PoissonEngine() {
  free_levels();
  for (double* v : {b_, x_, r_, pa_, pb_, q_, z_}) hipFree(v);
  hipFree(st_);
  hipFree(log_);
  if (log_host_) hipHostFree(log_host_);
  hipFree(partial_);
  hipFree(counter_);
  hipFree(fine_active_);
  if (st_host_) hipHostFree(st_host_);
  if (ev_) hipEventDestroy(ev_);
  clock_free();
}

void PoissonEngine::clock_free() {
  for (hipEvent_t e : clock_pool_) hipEventDestroy(e);
  clock_pool_.clear();
  clock_pool_used_ = 0;
  clock_log_.clear();
  clock_on_ = false;
}

int PoissonEngine::clock_start(int max_launches) {
  if (max_launches < 1) {
    set_last_error("clock_start: max_launches must be >= 1, got %d", max_launches);
    return kErrBadArg;
  }
  PCFD_HIP_CHECK(hipStreamSynchronize(s_));
  clock_free();
  clock_pool_.resize(2 * (size_t)max_launches, nullptr);
  for (hipEvent_t& e : clock_pool_) PCFD_HIP_CHECK(hipEventCreate(&e));
  clock_log_.reserve((size_t)max_launches);
  clock_on_ = true;
  return kOk;
}

LaunchClock PoissonEngine::clock_next(int kind, int level, double bytes) {
  if (!clock_on_ || clock_pool_used_ + 2 > clock_pool_.size()) return LaunchClock{};
  const int lev = level < kClockLevels ? level : kClockLevels - 1;
  LaunchClock c{clock_pool_[clock_pool_used_], clock_pool_[clock_pool_used_ + 1]};
  clock_pool_used_ += 2;
  clock_log_.push_back(ClockEntry{kind * kClockLevels + lev, clock_iter_, c.start, c.stop, bytes});
  return c;
}

LaunchClock PoissonEngine::clock_step(int kind, double bytes, int level) {
  if (!clock_on_) return LaunchClock{};
  const int it = clock_iter_;
  clock_iter_ = INT_MIN;  // an entry no trim drops
  const LaunchClock c = clock_next(kind, level, bytes);
  clock_iter_ = it;
  return c;
}

LaunchClock PoissonEngine::boundary_clock(void* self, int kind, double bytes) {
  return static_cast<PoissonEngine*>(self)->clock_step(kind == kClockWalls ? CK_WALLS : CK_BPASS,
                                                       bytes);
}

int PoissonEngine::clock_read(int32_t* launches, double* avg_ms, double* bytes) {
  PCFD_HIP_CHECK(hipStreamSynchronize(s_));
  std::vector<double> total(kClockSlots, 0.0);
  for (int q = 0; q < kClockSlots; ++q) {
    launches[q] = 0;
    bytes[q] = 0.0;
  }
  for (const ClockEntry& e : clock_log_) {
    if (e.slot < 0) continue;
    float ms = 0.f;
    PCFD_HIP_CHECK(hipEventElapsedTime(&ms, e.a, e.b));
    total[e.slot] += ms;
    ++launches[e.slot];
    bytes[e.slot] = e.bytes;
  }
  for (int q = 0; q < kClockSlots; ++q) avg_ms[q] = launches[q] ? total[q] / launches[q] : 0.0;
  clock_free();
  return kOk;
}

void PoissonEngine::free_levels() {
  for (auto* v : {&lv64_, &lv32_}) {
    for (auto& L : *v) {
      for (void* p : {L.b, L.x, L.y, L.r, L.t, L.ent_val, L.cdinv}) hipFree(p);
      hipFree(L.cls);
      hipFree(L.cls32);
      hipFree(L.row_ptr);
      hipFree(L.ent_lin);
      hipFree(L.agg);
      hipFree(L.ent_di);
      hipFree(L.ent_dj);
      hipFree(L.cactive);
      hipFree(L.cinv);
      hipFree(L.ent_off);
      hipFree(L.box);
      hipFree(L.box_dinv);
      hipFree(L.box_act);
      hipFree(L.rr);
      hipFree(L.rdinv);
      hipFree(L.ract);
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
  partial_cap_ = nt * 8;
  PCFD_HIP_CHECK(hipMalloc(&partial_, partial_cap_ * sizeof(double)));
  PCFD_HIP_CHECK(hipMalloc(&counter_, 64 * sizeof(unsigned int)));
  PCFD_HIP_CHECK(hipMemsetAsync(counter_, 0, 64 * sizeof(unsigned int), s_));
  PCFD_HIP_CHECK(hipMalloc(&st_, sizeof(SolveState)));
  PCFD_HIP_CHECK(hipMemsetAsync(st_, 0, sizeof(SolveState), s_));
  PCFD_HIP_CHECK(hipMalloc(&log_, kLogSlots * sizeof(SolveLog)));
  PCFD_HIP_CHECK(hipMemsetAsync(log_, 0, kLogSlots * sizeof(SolveLog), s_));
  PCFD_HIP_CHECK(hipHostMalloc(&log_host_, kLogSlots * sizeof(SolveLog), hipHostMallocDefault));
  PCFD_HIP_CHECK(hipHostMalloc(&st_host_, sizeof(SolveState), hipHostMallocDefault));
  PCFD_HIP_CHECK(hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
  return kOk;
}

// Uploads one level in precision T: vectors, and for coarse levels the row
// dictionary built from the dense class table (zeros dropped) plus the
// aggregation map towards the next level.
// LDS budget of one fused level workgroup (the dynamic allocation a launch
// may take without raising the kernel's attribute).
constexpr unsigned kLdsCap = 64u * 1024u;
constexpr int kFuseMaxRadius = 8;

template <typename T>
static int upload_level(const MGLevelHost& h, const MGLevelHost* next, bool fine, int nu,
                        bool fused_ok, LevelBuf& L, hipStream_t s) {
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
    int max_row = 0, min_row = 1 << 30;
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
      min_row = std::min(min_row, rp[c + 1] - rp[c]);
    }
    L.min_row = nc ? min_row : 0;
    std::vector<T> cd(nc);
    for (size_t k = 0; k < nc; ++k) cd[k] = (T)h.cdinv[k];
    L.rxs = 0;
    L.rys = 0;
    for (size_t e = 0; e < edi.size(); ++e) {
      L.rxs = std::max(L.rxs, std::abs((int)edi[e]));
      L.rys = std::max(L.rys, std::abs((int)edj[e]));
    }
    L.rxs = std::max(L.rxs, 1);
    L.rys = std::max(L.rys, 1);
    // Stored levels with a coarser one below and a narrow stencil may run
    // fused (the halo recomputation grows with the radius squared times the
    // row length; past radius 3 it costs more than the launches it saves).
    // The tile plan keeps the entry offsets to place them in its LDS region.
    const bool wide = !h.cls32.empty();
    {  // the most frequent row, whose entries the row passes load speculatively
      std::vector<int> count(nc, 0);
      if (wide) {
        for (uint32_t c : h.cls32) ++count[c];
      } else {
        for (uint16_t c : h.cls) ++count[c];
      }
      const int sc = (int)(std::max_element(count.begin(), count.end()) - count.begin());
      L.spec_e0 = rp[sc];
      L.spec_e1 = rp[sc + 1];
    }
    if (next && fused_ok && !wide && std::max(L.rxs, L.rys) <= kFuseMaxRadius) {
      L.hdi = edi;
      L.hdj = edj;
      std::vector<int> count(nc, 0);
      for (uint16_t c : h.cls) ++count[c];
      const int hot = (int)(std::max_element(count.begin(), count.end()) - count.begin());
      const int hot_n = rp[hot + 1] - rp[hot];
      if (hot_n <= kHotMax) {
        L.hot = hot;
        L.hot_n = hot_n;
        L.hot_e0 = rp[hot];
      }
    }
    const size_t nnz = std::max<size_t>(val.size(), 1);
    if (wide) {
      PCFD_HIP_CHECK(hipMalloc(&L.cls32, n * sizeof(uint32_t)));
      PCFD_HIP_CHECK(
          hipMemcpy(L.cls32, h.cls32.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    } else {
      PCFD_HIP_CHECK(hipMalloc(&L.cls, n * sizeof(uint16_t)));
      PCFD_HIP_CHECK(hipMemcpy(L.cls, h.cls.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice));
    }
    PCFD_HIP_CHECK(hipMalloc(&L.row_ptr, (nc + 1) * sizeof(int)));
    PCFD_HIP_CHECK(hipMalloc(&L.ent_lin, nnz * sizeof(int)));
    PCFD_HIP_CHECK(hipMalloc(&L.ent_di, nnz));
    PCFD_HIP_CHECK(hipMalloc(&L.ent_dj, nnz));
    PCFD_HIP_CHECK(hipMalloc(&L.ent_val, nnz * sizeof(T)));
    PCFD_HIP_CHECK(hipMalloc(&L.cdinv, nc * sizeof(T)));
    PCFD_HIP_CHECK(hipMalloc(&L.cactive, nc));
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
    {  // per-cell resolution of the class dictionary
      std::vector<int2> rr(n);
      std::vector<T> rd(n);
      std::vector<uint8_t> ra(n);
      for (size_t k = 0; k < n; ++k) {
        const size_t c = wide ? (size_t)h.cls32[k] : (size_t)h.cls[k];
        rr[k] = make_int2(rp[c], rp[c + 1]);
        rd[k] = cd[c];
        ra[k] = h.cactive[c];
      }
      PCFD_HIP_CHECK(hipMalloc(&L.rr, n * sizeof(int2)));
      PCFD_HIP_CHECK(hipMalloc(&L.rdinv, n * sizeof(T)));
      PCFD_HIP_CHECK(hipMalloc(&L.ract, n));
      PCFD_HIP_CHECK(hipMemcpy(L.rr, rr.data(), n * sizeof(int2), hipMemcpyHostToDevice));
      PCFD_HIP_CHECK(hipMemcpy(L.rdinv, rd.data(), n * sizeof(T), hipMemcpyHostToDevice));
      PCFD_HIP_CHECK(hipMemcpy(L.ract, ra.data(), n, hipMemcpyHostToDevice));
    }
    // box dictionary for the streaming legs: class rows as dense (2R+1)^2
    // boxes, then one all-zero class for cells outside the grid
    const int R = std::max(L.rxs, L.rys);
    if (next && !wide && R <= kBoxMaxRadius) {
      const int W = 2 * R + 1;
      std::vector<T> box((nc + 1) * (size_t)W * W, (T)0), bd(nc + 1, (T)0);
      std::vector<uint8_t> ba(nc + 1, 0);
      for (size_t c = 0; c < nc; ++c) {
        for (size_t q = 0; q < S; ++q)
          box[c * W * W + (size_t)(h.di[q] + R) * W + (h.dj[q] + R)] = (T)h.table[c * S + q];
        bd[c] = cd[c];
        ba[c] = h.cactive[c];
      }
      PCFD_HIP_CHECK(hipMalloc(&L.box, box.size() * sizeof(T)));
      PCFD_HIP_CHECK(hipMalloc(&L.box_dinv, bd.size() * sizeof(T)));
      PCFD_HIP_CHECK(hipMalloc(&L.box_act, ba.size()));
      PCFD_HIP_CHECK(hipMemcpy(L.box, box.data(), box.size() * sizeof(T), hipMemcpyHostToDevice));
      PCFD_HIP_CHECK(hipMemcpy(L.box_dinv, bd.data(), bd.size() * sizeof(T), hipMemcpyHostToDevice));
      PCFD_HIP_CHECK(hipMemcpy(L.box_act, ba.data(), ba.size(), hipMemcpyHostToDevice));
      L.box_radius = R;
    }
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
    L.coarse_n = (int)std::lround(std::sqrt((double)h.coarse_inv.size()));
    PCFD_HIP_CHECK(hipMalloc(&L.cinv, h.coarse_inv.size() * sizeof(double)));
    PCFD_HIP_CHECK(hipMemcpy(L.cinv, h.coarse_inv.data(), h.coarse_inv.size() * sizeof(double),
                             hipMemcpyHostToDevice));
  }
  return kOk;
}

int PoissonEngine::configure(const PoissonConfig& cfg, const std::vector<MGLevelHost>& levels) {
  cfg_ = cfg;
  free_levels();
  dir_tile_ = FusedTile{};
  chain_first_ = -1;
  last_iters_ = 0;
  predicted_ = 0;
  const bool need_mg = cfg.kind == SK_MULTIGRID || cfg.kind == SK_MGCG || cfg.kind == SK_DIRECT;
  if (!need_mg) return kOk;
  if (levels.empty()) {
    set_last_error("multigrid solver configured without a hierarchy");
    return kErrBadArg;
  }
  if ((op_.nx == 1 && op_.dir[0] && op_.dir[1]) || (op_.ny == 1 && op_.dir[2] && op_.dir[3])) {
    set_last_error("multigrid on a grid one cell wide between two pressure outlets is not "
                   "supported; use pressure_solver='cg'");
    return kErrUnsupported;
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
      int rc = f32 ? upload_level<float>(h, next, l == 0, cfg.sweeps, cfg.fused != 0, L, s_)
                   : upload_level<double>(h, next, l == 0, cfg.sweeps, cfg.fused != 0, L, s_);
      if (rc) return rc;
      L.lanes = L.max_row <= 16 ? 1 : L.max_row <= 64 ? 4 : L.max_row <= 256 ? 16 : 64;
      // register-resident passes: the fewest lanes that hold a row in kRowNE
      // entries each (0: rows too long, the looping passes run the level)
      L.row_lanes = 0;
      if (l > 0 && cfg.row_reg >= 0 && L.min_row >= 1 && (L.fx == 1 || L.fx == 2) &&
          (L.fy == 1 || L.fy == 2))
        for (int lanes : {1, 2, 4, 8, 16, 32, 64})
          if (L.max_row <= lanes * kRowNE) {
            L.row_lanes = lanes;
            break;
          }
      dst.push_back(L);
    }
    const int rc = f32 ? plan_level_tiles<float>(dst) : plan_level_tiles<double>(dst);
    if (rc) return rc;
  }
  const LevelBuf& last = lv64_.back();
  if (last.cinv == nullptr || last.coarse_n != last.n) {
    set_last_error("the coarsest level needs an n x n dense inverse (n = %d)", last.n);
    return kErrBadArg;
  }
  if ((size_t)last.n * sizeof(double) > kLdsCap) {
    set_last_error("coarsest level of %d unknowns exceeds the dense solve's %u-byte LDS stage",
                   last.n, kLdsCap);
    return kErrBadArg;
  }
  PCFD_HIP_CHECK(hipStreamSynchronize(s_));
  return kOk;
}

int PoissonEngine::ensure_partials(size_t n) {
  if (n <= partial_cap_) return kOk;
  PCFD_HIP_CHECK(hipStreamSynchronize(s_));
  PCFD_HIP_CHECK(hipFree(partial_));
  partial_ = nullptr;
  partial_cap_ = 0;
  PCFD_HIP_CHECK(hipMalloc(&partial_, n * sizeof(double)));
  partial_cap_ = n;
  ++partial_gen_;
  return kOk;
}

// Measured tile choices, kept for the life of the process: a level of the same
// shape, stencil radius, sweeps and precision gets the same plan again.
namespace {
using TuneKey = std::array<int, 14>;
struct TunePick {
  int fused, tx, ty;
  int stream = 0;  // level 0: streaming legs with this many rows per wavefront
};
std::map<TuneKey, TunePick> g_tune_cache;
std::mutex g_tune_mutex;

bool autotune_enabled() {
  const char* e = std::getenv("PCFD_AUTOTUNE");
  return !(e && e[0] == '0');
}

// Optional persistence of the measured choices across processes: with
// PCFD_TUNE_FILE set, the file's entries (one line of 14 key integers and
// 4 choice integers each) seed the cache on first use and every new
// measurement is appended, so later runs of the same workload (a profiled
// run, say) take the same plan without measuring.  Caller holds g_tune_mutex.
const char* tune_file() {
  const char* f = std::getenv("PCFD_TUNE_FILE");
  return (f && f[0]) ? f : nullptr;
}

void tune_file_load() {
  static bool loaded = false;
  if (loaded) return;
  loaded = true;
  const char* f = tune_file();
  if (!f) return;
  FILE* fp = std::fopen(f, "r");
  if (!fp) return;
  for (;;) {
    TuneKey k{};
    TunePick p{};
    int got = 0;
    for (int& v : k) got += std::fscanf(fp, "%d", &v) == 1;
    got += std::fscanf(fp, "%d %d %d %d", &p.fused, &p.tx, &p.ty, &p.stream) == 4 ? 4 : 0;
    if (got != (int)k.size() + 4) break;
    g_tune_cache[k] = p;
  }
  std::fclose(fp);
}

void tune_store(const TuneKey& k, const TunePick& p) {
  g_tune_cache[k] = p;
  const char* f = tune_file();
  if (!f || !autotune_enabled()) return;  // only measured choices persist
  FILE* fp = std::fopen(f, "a");
  if (!fp) return;
  for (int v : k) std::fprintf(fp, "%d ", v);
  std::fprintf(fp, "%d %d %d %d\n", p.fused, p.tx, p.ty, p.stream);
  std::fclose(fp);
}
}  // namespace

// Tiles of the level-l legs.  Every tiling gives the same arithmetic (each
// cell's passes read the same inputs); only speed differs.  Candidates whose
// LDS region fits are timed on this device (down + up leg, HIP events), and
// for a stored level the pass-by-pass path competes too.  PCFD_AUTOTUNE=0
// uses the static cost estimate instead.
template <typename T>
int PoissonEngine::tune_level(std::vector<LevelBuf>& lv, int l, SolveState* tst) {
  LevelBuf& L = lv[l];
  LevelBuf& C = lv[l + 1];
  const int nu = cfg_.sweeps;
  const bool fine = l == 0;
  L.fused = 0;
  L.box_rows = 0;
  // streaming legs: level 0 (uniform mesh) or a stored level with a box
  // dictionary; one sweep each way
  const bool stream_ok = cfg_.stream >= 0 &&
                         (fine ? stream_supported(op_, nu, L.fx, L.fy)
                               : (L.box_radius > 0 && nu == 1));
  if (!fine && L.hdi.empty() && !stream_ok) return kOk;
  // LDS bytes per region cell: three vectors + flags (level 0); three
  // vectors, 1/diag, entry range and active flag (stored levels)
  const unsigned per_cell = fine ? 3u * sizeof(T) + 1u : 4u * sizeof(T) + 9u;
  const unsigned fixed =
      fine ? 256u * sizeof(T)
           : (L.hot >= 0 ? (unsigned)L.hot_n * (unsigned)(sizeof(T) + 4) : 0u) + 16u;
  const int ftx = fine ? cfg_.tile_tx : 0, fty = fine ? cfg_.tile_ty : 0;
  std::vector<FusedTile> cands;
  if (fine || !L.hdi.empty())
    cands = tile_candidates(L.nx, L.ny, L.rxs, L.rys, nu + 1, L.fx, L.fy, per_cell, fixed,
                            kLdsCap, ftx, fty);
  if (fine) {  // the standalone solver's nonzero start needs one more ring
    std::vector<FusedTile> keep;
    for (const FusedTile& t : cands) {
      FusedTile t1;
      if (choose_tile(L.nx, L.ny, 1, 1, nu + 2, L.fx, L.fy, per_cell, fixed, kLdsCap, t.tx, t.ty,
                      &t1))
        keep.push_back(t);
    }
    cands.swap(keep);
    if (cands.empty()) {
      set_last_error("no fused tiling of the %dx%d level fits (tile %dx%d, %d sweeps)", L.nx,
                     L.ny, cfg_.tile_tx, cfg_.tile_ty, nu);
      return kErrBadArg;
    }
    size_t most = 0;
    for (const FusedTile& t : cands) most = std::max(most, (size_t)t.ntiles() * 3);
    if (int rc = ensure_partials(most)) return rc;
  }
  if (cands.empty() && !stream_ok) return kOk;

  auto upload_offsets = [&](const FusedTile& t) -> int {
    std::vector<int> off(std::max<size_t>(L.hdi.size(), 1), 0);
    for (size_t e = 0; e < L.hdi.size(); ++e) off[e] = (int)L.hdi[e] * t.ry + (int)L.hdj[e];
    if (!L.ent_off) PCFD_HIP_CHECK(hipMalloc(&L.ent_off, off.size() * sizeof(int)));
    PCFD_HIP_CHECK(hipMemcpyAsync(L.ent_off, off.data(), off.size() * sizeof(int),
                                  hipMemcpyHostToDevice, s_));
    PCFD_HIP_CHECK(hipStreamSynchronize(s_));
    return kOk;
  };
  // fused: tiled kernels with tile t (0: pass by pass); stream > 0: the
  // streaming legs with that many rows per wavefront (level 0 keeps the tile
  // for its nonzero-start legs)
  auto apply = [&](int fused, const FusedTile& t, int stream) -> int {
    L.fused = fused;
    if (fine) {
      L.stream_rows = stream;
      L.stream_nb = stream ? stream_blocks(L.nx, L.ny, stream) : 0;
      if (stream) {
        if (int rc = ensure_partials((size_t)L.stream_nb * 3)) return rc;
      }
    } else {
      L.box_rows = stream;
    }
    if (!fused) return kOk;
    L.down0 = t;
    L.up = t;
    if (fine) {
      if (!choose_tile(L.nx, L.ny, 1, 1, nu + 2, L.fx, L.fy, per_cell, fixed, kLdsCap, t.tx, t.ty,
                       &L.down1))
        return kErrBadArg;
      return kOk;
    }
    return upload_offsets(t);
  };
  auto legs = [&]() -> int {
    if (fine && L.stream_rows) {
      if (int rc = launch_fine_down_stream<T>(s_, op_, L.stream_rows, (T)L.weight, r_, (T*)L.b,
                                              (T*)L.t, (T*)C.b, L.fx, L.fy, C.nx, C.ny, tst))
        return rc;
      return launch_fine_up_stream<T>(s_, op_, L.stream_rows, (T)L.weight, (const T*)L.t,
                                      (const T*)L.b, (const T*)C.y, L.fx, L.fy, C.ny, (T*)L.y,
                                      (int)SR_MGCG, partial_, tst);
    }
    if (fine) {
      if (int rc = launch_fine_down<T>(s_, op_, L.down0, (T)L.weight, nu, 0, r_, nullptr, nullptr,
                                       0, (T*)L.b, (T*)L.t, (T*)C.b, L.fx, L.fy, C.nx, C.ny, tst))
        return rc;
      return launch_fine_up<T>(s_, op_, L.up, (T)L.weight, nu, (const T*)L.t, (const T*)L.b,
                               (const T*)C.y, L.fx, L.fy, C.ny, (T*)L.y, (int)SR_MGCG, partial_,
                               tst);
    }
    if (int rc = level_down<T>(L, C, tst, l)) return rc;
    return level_up<T>(L, C, tst, l);
  };
  const FusedTile none{};
  // a forced streaming choice (cfg_.stream > 0) needs no measurement
  if (cfg_.stream > 0 && stream_ok) {
    const FusedTile t =
        cands.empty() ? none
                      : *std::min_element(cands.begin(), cands.end(),
                                          [](const FusedTile& a, const FusedTile& b) {
                                            return tile_cost(a) < tile_cost(b);
                                          });
    return apply(fine ? 1 : 0, t, cfg_.stream);
  }

  const TuneKey key = {fine ? 1 : 0, L.nx, L.ny, L.rxs, L.rys, L.fx, L.fy, nu, (int)sizeof(T),
                       op_.px, op_.py, ftx, fty, op_.hxc != nullptr};
  {
    std::lock_guard<std::mutex> g(g_tune_mutex);
    tune_file_load();
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      const TunePick pk = it->second;
      const int st = stream_ok ? pk.stream : 0;
      if (!pk.fused && !fine) return apply(0, none, st);
      for (const FusedTile& t : cands)
        if (t.tx == pk.tx && t.ty == pk.ty) return apply(pk.fused, t, st);
    }
  }
  int best_fused = fine ? 1 : 0, best_stream = 0;
  FusedTile best = cands.empty() ? none
                                 : *std::min_element(cands.begin(), cands.end(),
                                                     [](const FusedTile& a, const FusedTile& b) {
                                                       return tile_cost(a) < tile_cost(b);
                                                     });
  if (!cands.empty() && !fine) best_fused = 1;
  if (autotune_enabled() && (cands.size() > 1 || !fine || stream_ok)) {
    hipEvent_t e0, e1;
    PCFD_HIP_CHECK(hipEventCreate(&e0));
    PCFD_HIP_CHECK(hipEventCreate(&e1));
    auto time_legs = [&](float* ms) -> int {
      if (int rc = legs()) return rc;  // warm-up
      PCFD_HIP_CHECK(hipEventRecord(e0, s_));
      for (int k = 0; k < 3; ++k)
        if (int rc = legs()) return rc;
      PCFD_HIP_CHECK(hipEventRecord(e1, s_));
      PCFD_HIP_CHECK(hipEventSynchronize(e1));
      PCFD_HIP_CHECK(hipEventElapsedTime(ms, e0, e1));
      return kOk;
    };
    float best_ms = 0.f;
    bool have = false;
    if (!fine) {  // the pass-by-pass path as the first contender
      if (int rc = apply(0, none, 0)) return rc;
      if (int rc = time_legs(&best_ms)) return rc;
      best_fused = 0;
      best = none;
      have = true;
    }
    for (const FusedTile& t : cands) {
      if (int rc = apply(1, t, 0)) return rc;
      float ms = 0.f;
      if (int rc = time_legs(&ms)) return rc;
      if (!have || ms < best_ms) {
        have = true;
        best_ms = ms;
        best = t;
        best_fused = 1;
      }
    }
    // the streaming legs as further contenders (level 0 keeps its best tile
    // for the nonzero-start legs)
    if (stream_ok) {
      for (int rows : {4, 8, 16, 32}) {
        if (int rc = apply(fine ? 1 : 0, fine ? best : none, rows)) return rc;
        float ms = 0.f;
        if (int rc = time_legs(&ms)) return rc;
        if (!have || ms < best_ms) {
          have = true;
          best_ms = ms;
          best_stream = rows;
          if (!fine) {
            best_fused = 0;
            best = none;
          }
        }
      }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  }
  {
    std::lock_guard<std::mutex> g(g_tune_mutex);
    TunePick pk{best_fused, best.tx, best.ty};
    pk.stream = best_stream;
    if (cfg_.stream == 0) tune_store(key, pk);
  }
  return apply(best_fused, best, best_stream);
}

// Tile of the level-0 multigrid-CG direction kernel (halo 1), measured like
// the level legs.
template <typename T>
int PoissonEngine::tune_dir(SolveState* tst) {
  const unsigned per_cell = 8u + 1u, fixed = 256u * 8u;
  std::vector<FusedTile> cands =
      tile_candidates(op_.nx, op_.ny, 1, 1, 1, 1, 1, per_cell, fixed, kLdsCap, 0, 0);
  if (cands.empty()) {
    set_last_error("no tiling of the %dx%d grid fits the direction kernel", op_.nx, op_.ny);
    return kErrBadArg;
  }
  size_t most = 0;
  for (const FusedTile& t : cands) most = std::max(most, (size_t)t.ntiles());
  if (int rc = ensure_partials(most)) return rc;
  const TuneKey key = {2, op_.nx, op_.ny, 1, 1, 1, 1, 0, (int)sizeof(T), op_.px, op_.py, 0, 0,
                       op_.hxc != nullptr};
  {
    std::lock_guard<std::mutex> g(g_tune_mutex);
    tune_file_load();
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end())
      for (const FusedTile& t : cands)
        if (t.tx == it->second.tx && t.ty == it->second.ty) {
          dir_tile_ = t;
          return kOk;
        }
  }
  FusedTile best = *std::min_element(cands.begin(), cands.end(),
                                     [](const FusedTile& a, const FusedTile& b) {
                                       return tile_cost(a) < tile_cost(b);
                                     });
  if (autotune_enabled() && cands.size() > 1) {
    hipEvent_t e0, e1;
    PCFD_HIP_CHECK(hipEventCreate(&e0));
    PCFD_HIP_CHECK(hipEventCreate(&e1));
    const T* z = (const T*)((sizeof(T) == 4) ? lv32_[0].y : lv64_[0].y);
    float best_ms = -1.f;
    for (const FusedTile& t : cands) {
      for (int k = 0; k < 4; ++k) {
        if (k == 1) PCFD_HIP_CHECK(hipEventRecord(e0, s_));
        if (int rc = launch_fine_dir<T>(s_, op_, t, z, pa_, pb_, q_, partial_, tst)) return rc;
      }
      PCFD_HIP_CHECK(hipEventRecord(e1, s_));
      PCFD_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      PCFD_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (best_ms < 0.f || ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  }
  {
    std::lock_guard<std::mutex> g(g_tune_mutex);
    tune_store(key, TunePick{1, best.tx, best.ty});
  }
  dir_tile_ = best;
  return kOk;
}

// Whether the deepest levels run as one cooperative launch, and from which
// level: the per-level legs and the chain starting at each of the last few
// levels are timed on this device (HIP events over the coarse legs) and the
// fastest kept.  The chain needs every workgroup of its grid resident at once,
// which the cooperative launch guarantees or refuses.
template <typename T>
int PoissonEngine::tune_chain(std::vector<LevelBuf>& lv, SolveState* tst) {
  chain_first_ = -1;
  const int nl = (int)lv.size();
  if (nl < 3 || cfg_.chain < 0) return kOk;
  if (cfg_.chain == 0 && !autotune_enabled()) return kOk;
  int dev = 0, cus = 0, coop = 0, per = 0;
  PCFD_HIP_CHECK(hipGetDevice(&dev));
  PCFD_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  PCFD_HIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  const size_t lds = (size_t)lv[nl - 1].n * sizeof(double);
  if (!coop || hipOccupancyMaxActiveBlocksPerMultiprocessor(
                   &per, reinterpret_cast<const void*>(&k_chain<T>), kBlock, lds) != hipSuccess ||
      per < 1) {
    (void)hipGetLastError();
    return kOk;
  }
  chain_blocks_ = cus * std::min(per, 2);
  if (cfg_.chain > 0) {  // pinned by the caller: kept if one launch is accepted
    chain_first_ = std::min(std::max(cfg_.chain, nl - 1 - kMaxChain), nl - 2);
    chain_first_ = std::max(chain_first_, 1);
    if (coarse_legs<T>(lv, 1, tst) != kOk) {  // e.g. refused while another
      (void)hipGetLastError();                // process holds the device
      (void)hipStreamSynchronize(s_);
      chain_first_ = -1;
    }
    return kOk;
  }
  hipEvent_t e0, e1;
  PCFD_HIP_CHECK(hipEventCreate(&e0));
  PCFD_HIP_CHECK(hipEventCreate(&e1));
  auto time_legs = [&](float* ms) -> int {
    for (int k = 0; k < 4; ++k) {
      if (k == 1) PCFD_HIP_CHECK(hipEventRecord(e0, s_));
      if (int rc = coarse_legs<T>(lv, 1, tst)) return rc;
    }
    PCFD_HIP_CHECK(hipEventRecord(e1, s_));
    PCFD_HIP_CHECK(hipEventSynchronize(e1));
    PCFD_HIP_CHECK(hipEventElapsedTime(ms, e0, e1));
    return kOk;
  };
  float best = 0.f;
  int rc = time_legs(&best);
  int pick = -1;
  // candidate starts: the deepest stored level up to three levels higher
  for (int f = nl - 2; f >= std::max(1, nl - 5) && rc == kOk; --f) {
    if (nl - 1 - f > kMaxChain) break;
    chain_first_ = f;
    float ms = 0.f;
    const int r2 = time_legs(&ms);
    if (r2 != kOk) {  // a refused cooperative launch just leaves the chain out
      (void)hipGetLastError();
      (void)hipStreamSynchronize(s_);
      break;
    }
    if (ms < best) {
      best = ms;
      pick = f;
    }
  }
  chain_first_ = pick;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return rc;
}

template <typename T>
int PoissonEngine::plan_level_tiles(std::vector<LevelBuf>& lv) {
  if (lv.empty()) return kOk;
  SolveState* tst = nullptr;
  PCFD_HIP_CHECK(hipMalloc(&tst, sizeof(SolveState)));
  PCFD_HIP_CHECK(hipMemsetAsync(tst, 0, sizeof(SolveState), s_));
  int rc = kOk;
  for (int l = 0; l + 1 < (int)lv.size() && rc == kOk; ++l) rc = tune_level<T>(lv, l, tst);
  // the chain serves the cycle precision's hierarchy
  if (rc == kOk && (cfg_.precision == 1) == (sizeof(T) == 4)) rc = tune_chain<T>(lv, tst);
  // the Krylov direction kernel reads the cycle output in this precision
  if (rc == kOk && (cfg_.kind == SK_MGCG || cfg_.kind == SK_DIRECT) &&
      (cfg_.precision == 1) == (sizeof(T) == 4))
    rc = tune_dir<T>(tst);
  const hipError_t e = hipStreamSynchronize(s_);
  hipFree(tst);
  if (rc) return rc;
  if (e != hipSuccess) {
    set_last_error("tile planning failed: %s", hipGetErrorString(e));
    return kErrHip;
  }
  return kOk;
}

int PoissonEngine::prepare_from_velocity(const double* u, const double* v, int ldu, int ldv,
                                         double dx, double dy, const StepScalars& T) {
  const int mb = multi_blocks(op_.nx, op_.ny);
  // algorithmic bytes: u*, v* (one face each per cell), flags; b
  const LaunchClock clk = clock_step(CK_PREPARE, (double)n_ * (8 + 8 + 1 + 8));
  if (is_pow2(dx) && is_pow2(dy))
    PCFD_LAUNCH_TIMED(clk, k_prepare<true>, dim3(mb), dim3(kBlock), 0, s_, op_, u, v, ldu, ldv,
                      dx, dy, 1.0 / dx, 1.0 / dy, T, (const double*)nullptr, b_, x_, partial_);
  else
    PCFD_LAUNCH_TIMED(clk, k_prepare<false>, dim3(mb), dim3(kBlock), 0, s_, op_, u, v, ldu, ldv,
                      dx, dy, 1.0 / dx, 1.0 / dy, T, (const double*)nullptr, b_, x_, partial_);
  PCFD_LAUNCH_CHECK();
  return fold(s_, partial_, mb, EpiPrepare{st_, singular_, n_fluid_},
              clock_step(CK_FOLD, 0.0, kFoldPrepare));
}

int PoissonEngine::prepare_from_rhs(const double* rhs) {
  const int mb = multi_blocks(op_.nx, op_.ny);
  hipLaunchKernelGGL(k_prepare<false>, dim3(mb), dim3(kBlock), 0, s_, op_,
                     (const double*)nullptr, (const double*)nullptr, 0, 0, 1.0, 1.0, 1.0, 1.0,
                     StepScalars{1.0, 0.0, nullptr}, rhs, b_, x_, partial_);
  PCFD_LAUNCH_CHECK();
  return fold(s_, partial_, mb, EpiPrepare{st_, singular_, n_fluid_});
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

int PoissonEngine::poll(int* iter, int* done, int* converged) {
  PCFD_HIP_CHECK(hipMemcpyAsync(st_host_, st_, sizeof(SolveState), hipMemcpyDeviceToHost, s_));
  PCFD_HIP_CHECK(hipEventRecord(ev_, s_));
  PCFD_HIP_CHECK(hipEventSynchronize(ev_));
  *iter = st_host_->iter;
  *done = st_host_->done;
  *converged = st_host_->converged;
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
  const int mb = multi_blocks(op_.nx, op_.ny);
  // algorithmic bytes: x, b, flags
  const LaunchClock clk = clock_step(CK_CHECK, (double)n_ * (8 + 8 + 1));
  PCFD_LAUNCH_TIMED(clk, k_check, dim3(mb), dim3(kBlock), 0, s_, op_, x_, b_, (int)CM_FINAL,
                    partial_, (const SolveState*)st_);
  PCFD_LAUNCH_CHECK();
  SolveLog* log = (log_slot_ >= 0 && log_slot_ < kLogSlots) ? log_ + log_slot_ : nullptr;
  return fold(s_, partial_, mb,
              EpiCheck{st_, (int)CM_FINAL, cfg_.tol, cfg_.maxiter, singular_, n_fluid_, log},
              clock_step(CK_FOLD, 0.0, kFoldCheck));
}

int PoissonEngine::read_log(SolveLog* out, int n) {
  n = std::min(n, kLogSlots);
  PCFD_HIP_CHECK(hipMemcpyAsync(log_host_, log_, n * sizeof(SolveLog), hipMemcpyDeviceToHost, s_));
  PCFD_HIP_CHECK(hipStreamSynchronize(s_));
  std::memcpy(out, log_host_, n * sizeof(SolveLog));
  return kOk;
}

int PoissonEngine::scatter_pressure(double* p, int ldp, const WarmHistory& wh) {
  // algorithmic bytes: x, flags; x (normalised guess), p (and the history slot)
  const LaunchClock clk =
      clock_step(CK_SCATTER, (double)n_ * (8 + 1 + 8 + 8 + (wh.slots ? 8 : 0)));
  PCFD_LAUNCH_TIMED(clk, k_scatter, tiles_for(op_.nx + 2, op_.ny + 2), dim3(kBlock), 0, s_, op_,
                    x_, (const SolveState*)st_, p, ldp, (double*)nullptr, wh);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int PoissonEngine::warm_guess(const WarmHistory& wh, int order) {
  // algorithmic bytes with the history full: order + 1 solutions read, the
  // guess written (the first calls after a restart read fewer)
  const LaunchClock clk = clock_step(CK_WARM, (double)n_ * 8 * (order + 2));
  PCFD_LAUNCH_TIMED(clk, k_warm_guess, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s_, n_,
                    x_, wh, order);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int PoissonEngine::write_interior(double* out) {
  hipLaunchKernelGGL(k_scatter, tiles_for(op_.nx + 2, op_.ny + 2), dim3(kBlock), 0, s_, op_, x_,
                     (const SolveState*)st_, (double*)nullptr, 0, out, WarmHistory{});
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int PoissonEngine::apply_operator(const double* x, double* y) {
  hipLaunchKernelGGL(k_apply_fine, tiles_for(op_.nx, op_.ny), dim3(kBlock), 0, s_, op_, x, y);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

// Iteration driver shared by the solvers: enqueue a chunk of iterations,
// then poll the device flag.  Kernels turn into no-ops once the flag is set
// (also when the start already decided the solve), so over-enqueueing only
// costs launch slots; the first chunk is the last converged solve's count
// (warm-started solves barely change), so a solve that needs as many
// iterations polls once.  A solve that ran out of iterations does not seed
// the next one's chunk (it would enqueue the whole budget).  Every poll
// drains the stream and leaves the device idle until the next launch.  PER
// is the number of iterations one body advances; MAXB bounds the bodies.
// Under a budget (a captured solve) exactly enough bodies for that many
// iterations are enqueued and nothing is polled.  HOOK runs after each chunk,
// before its poll (and after a budget's bodies): it enqueues what the state
// the host reads still waits for.
#define PCFD_ITERATE_HOOK(PER, MAXB, HOOK, ...)                               \
  do {                                                                        \
    if (budget_ > 0) {                                                        \
      const int _nb = std::min((MAXB), (budget_ + (PER) - 1) / (PER));        \
      for (int _b = 0; _b < _nb; ++_b) {                                      \
        __VA_ARGS__;                                                          \
      }                                                                       \
      if (int _hrc = (HOOK)) return _hrc;                                     \
      last_run_ = budget_;                                                    \
      break;                                                                  \
    }                                                                         \
    int _it = 0, _done = 0, _conv = 0, _bodies = 0, _rc = 0;                  \
    int _chunk = std::max(1, (predicted_ > 0 ? predicted_ : 4) / (PER));      \
    while (!_done && _bodies < (MAXB)) {                                      \
      for (int _c = 0; _c < _chunk && _bodies < (MAXB); ++_c, ++_bodies) {    \
        __VA_ARGS__;                                                          \
      }                                                                       \
      if (int _hrc = (HOOK)) return _hrc;                                     \
      _rc = poll(&_it, &_done, &_conv);                                       \
      if (_rc) return _rc;                                                    \
      _chunk = std::max(1, std::min(64, _it / (8 * (PER))));                  \
    }                                                                         \
    last_run_ = _it;                                                          \
    if (_conv) predicted_ = _it;                                              \
  } while (0)
#define PCFD_ITERATE(PER, MAXB, ...) PCFD_ITERATE_HOOK(PER, MAXB, kOk, __VA_ARGS__)

// The right-hand side pass (k_prepare) has zeroed the guess off the
// solvable set.
int PoissonEngine::start(int mode, double* r) {
  const int mb = multi_blocks(op_.nx, op_.ny);
  // algorithmic bytes: b, x, flags; b, r
  const LaunchClock clk = clock_step(CK_INIT, (double)n_ * (8 + 8 + 1 + 8 + (r ? 8 : 0)));
  PCFD_LAUNCH_TIMED(clk, k_init, dim3(mb), dim3(kBlock), 0, s_, op_, b_, (const double*)x_, r,
                    mode, (const SolveState*)st_, partial_);
  PCFD_LAUNCH_CHECK();
  return fold(s_, partial_, mb, EpiInit{st_, mode, cfg_.tol, cfg_.maxiter, n_proj_},
              clock_step(CK_FOLD, 0.0, kFoldInit));
}

// ----------------------------------------------------------------------------
// Jacobi-CG of a small grid in one workgroup (scipy.sparse.linalg.cg with
// M = 1/diag, pressure.py:366-398, the same iteration as k_jpcg_dir /
// k_jpcg_update_f and their epilogues).  On a tiny grid the multi-launch
// iteration is bound by its four dependent launches per iteration; here the
// whole solve is one launch on one CU.  1024 threads hold x and r of CPT
// cells each in registers; the direction p and the flag bytes sit in LDS,
// where the operator product reads the neighbours (q = A p is formed twice
// per iteration, for p.q and for the update, rather than held in registers).
// Each inner product is a wavefront tree plus a fixed-order sum of the 16
// wavefront totals that every thread forms itself (one barrier).  The solve
// iterates until the recursive residual meets the stopping test or maxiter,
// then writes x, r and the state.  One CU's FP64 rate bounds it: measured on
// MI355X, 5.9 us per iteration at 32 x 32 against 12.5 for the two-launch
// iteration below, but 13.8 against 12.0 at 64 x 64, so it serves grids of at
// most 2048 cells.
// ----------------------------------------------------------------------------
constexpr int kSmallCgThreads = 1024;
constexpr int kSmallCgWaves = kSmallCgThreads / kWave;
constexpr int kSmallCgCells = 2 * kSmallCgThreads;

template <int K>
__device__ __forceinline__ void small_cg_sum(double (&v)[K], double (*slot)[kSmallCgWaves]) {
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double w = wave_sum(v[k]);
    if (lane == 0) slot[k][wid] = w;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double a = slot[k][0];
    for (int w = 1; w < kSmallCgWaves; ++w) a += slot[k][w];
    v[k] = a;
  }
}

// Cell m of this thread: c = m * 1024 + tid, (i, j) stepped from cell m - 1.
struct SmallCell {
  int c, i, j;
};

template <int CPT>
__global__ void __launch_bounds__(kSmallCgThreads)
k_jpcg_small(FineOp A, double* __restrict__ x, double* __restrict__ r, int maxiter,
             SolveState* st) {
  __shared__ double P[CPT * kSmallCgThreads];
  __shared__ uint8_t F[CPT * kSmallCgThreads];
  // two slot sets: the p.q sum and the (r.r, r.z) sum of one iteration use
  // different ones, so no sum overwrites slots another thread still reads
  __shared__ double slots_pq[1][kSmallCgWaves];
  __shared__ double slots_rr[2][kSmallCgWaves];
  if (st->done) return;  // decided by the start (zero rhs, met at once, maxiter 0)
  const int n = A.nx * A.ny, ny = A.ny;
  const int step_i = kSmallCgThreads / ny, step_j = kSmallCgThreads - step_i * ny;
  // The cells' indices are formed afresh in every pass from a thread index the
  // compiler cannot see through (an empty asm on the register): hoisted out of
  // the iteration loop, the per-cell indices and neighbours of CPT cells would
  // stay live across it and spill.
  auto cell = [&](int m, SmallCell& q) {  // q holds cell m - 1 (or is fresh for m = 0)
    if (m == 0) {
      int t = (int)threadIdx.x;
      asm volatile("" : "+v"(t));
      const int i0 = t / ny;
      q = SmallCell{t, i0, t - i0 * ny};
      return;
    }
    q.c += kSmallCgThreads;
    q.i += step_i;
    q.j += step_j;
    if (q.j >= ny) {
      q.j -= ny;
      ++q.i;
    }
  };
  double xv[CPT], rv[CPT];
  {
    SmallCell q{};
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      cell(m, q);
      const bool in = q.c < n;
      xv[m] = in ? x[q.c] : 0.0;
      rv[m] = in ? r[q.c] : 0.0;
      if (in) F[q.c] = A.flags[q.c];
    }
  }
  // (A p)[c] from the LDS copies, in fine_apply's term order
  auto ap = [&](const SmallCell& q) {
    const uint8_t f = F[q.c];
    const FaceC fc = fine_faces(A, q.i, q.j);
    const double d = fine_diag(A, fc, f, q.i, q.j);
    return fine_apply<double>(fc, f, fine_neighbours(A, q.i, q.j), q.c, d,
                              [&](int k) { return P[k]; });
  };
  double rho = st->rho, rho_prev = st->rho_prev, beta = 0.0, alpha = 0.0, rr = st->rr;
  const double atol = st->atol;
  int it = st->iter, conv = 0;
  __syncthreads();
  for (bool first = true;; first = false) {
    // p = z + beta p_old, z = r / diag (k_jpcg_dir's arithmetic)
    {
      SmallCell q{};
#pragma unroll
      for (int m = 0; m < CPT; ++m) {
        cell(m, q);
        if (q.c >= n) continue;
        const double z = rv[m] / fine_diag(A, F[q.c], q.i, q.j);
        P[q.c] = first ? z : P[q.c] * beta + z;
      }
    }
    __syncthreads();
    double pq[1] = {0.0};
    {
      SmallCell q{};
#pragma unroll 4
      for (int m = 0; m < CPT; ++m) {
        cell(m, q);
        if (q.c < n) pq[0] += P[q.c] * ap(q);
      }
    }
    small_cg_sum<1>(pq, slots_pq);
    alpha = rho / pq[0];
    // x += alpha p, r -= alpha q; r.r and r.(r / diag) (k_jpcg_update_f)
    double s[2] = {0.0, 0.0};
    {
      SmallCell q{};
#pragma unroll
      for (int m = 0; m < CPT; ++m) {
        cell(m, q);
        if (q.c >= n) continue;
        xv[m] = xv[m] + alpha * P[q.c];
        rv[m] = rv[m] - alpha * ap(q);
        s[0] += rv[m] * rv[m];
        s[1] += rv[m] * (rv[m] / fine_diag(A, F[q.c], q.i, q.j));
      }
    }
    small_cg_sum<2>(s, slots_rr);
    // the update epilogue (EpiUpdate, Jacobi-CG)
    ++it;
    rr = s[0];
    rho_prev = rho;
    rho = s[1];
    beta = rho / rho_prev;
    if (sqrt(rr) < atol) {
      conv = 1;
      break;
    }
    if (it >= maxiter) break;
  }
  {
    SmallCell q{};
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
      cell(m, q);
      if (q.c >= n) continue;
      x[q.c] = xv[m];
      r[q.c] = rv[m];
    }
  }
  if (threadIdx.x == 0) {
    st->iter = it;
    st->rr = rr;
    st->rho = rho;
    st->rho_prev = rho_prev;
    st->beta = beta;
    st->alpha = alpha;
    st->done = 1;
    st->converged = conv;
  }
}

int PoissonEngine::small_cg_cells_per_thread() const {
  const long n = (long)op_.nx * op_.ny;
  const char* e = std::getenv("PCFD_SMALL_CG");  // "0": always the multi-launch iteration
  if (n > kSmallCgCells || (e && e[0] == '0')) return 0;
  int cpt = 1;
  while ((long)cpt * kSmallCgThreads < n) cpt *= 2;
  return cpt;
}

int PoissonEngine::run_cg() {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  const int nt = ntiles(op_.nx, op_.ny);
  if (int rc = start(IM_JPCG, r_)) return rc;
  if (const int cpt = small_cg_cells_per_thread()) {
    const dim3 one(1), blk(kSmallCgThreads);
    const LaunchClock clk = clock_next(CK_SMALL_CG, 0, (double)n_ * 8 * 4);  // x, r in and out
    if (cpt == 1)
      PCFD_LAUNCH_TIMED(clk, k_jpcg_small<1>, one, blk, 0, s_, op_, x_, r_, cfg_.maxiter, st_);
    else
      PCFD_LAUNCH_TIMED(clk, k_jpcg_small<2>, one, blk, 0, s_, op_, x_, r_, cfg_.maxiter, st_);
    PCFD_LAUNCH_CHECK();
    last_run_ = 0;
    return kOk;
  }
  // Two launches per iteration: the direction kernel folds the previous
  // update's partials, the update folds the direction's p.q; a chunk's last
  // update is folded by EpiUpdateJ before the host reads the state.
  const size_t uoff = (nt + 255) / 256 * 256;
  if (int rc = ensure_partials(uoff + 2 * (size_t)nt)) return rc;
  double* p_old = pa_;
  double* p_new = pb_;
  int k = 0;             // updates enqueued
  bool pending = false;  // the last one's partials not folded yet
  auto flush = [&]() -> int {
    if (!pending) return kOk;
    pending = false;
    return fold(s_, partial_ + uoff, nt, EpiUpdateJ{st_, k, cfg_.maxiter});
  };
  PCFD_ITERATE_HOOK(1, cfg_.maxiter, flush(), {
    const JpcgFold jf{partial_ + uoff, pending ? nt : 0, k, cfg_.maxiter};
    pending = false;
    hipLaunchKernelGGL(k_jpcg_dir, g, dim3(kBlock), 0, s_, op_, r_, p_old, p_new, q_, partial_,
                       st_, jf);
    hipLaunchKernelGGL(k_jpcg_update_f, g, dim3(kBlock), 0, s_, op_, x_, r_, p_new, q_,
                       (const double*)partial_, nt, k, partial_ + uoff, st_);
    PCFD_LAUNCH_CHECK();
    ++k;
    pending = true;
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
    hipLaunchKernelGGL(k_check, dim3(multi_blocks(op_.nx, op_.ny)), dim3(kBlock), 0, s_, op_, x_,
                       b_, (int)CM_RELAX, partial_, (const SolveState*)st_);
    if (int rc = fold(s_, partial_, multi_blocks(op_.nx, op_.ny),
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
    hipLaunchKernelGGL(k_check, dim3(multi_blocks(op_.nx, op_.ny)), dim3(kBlock), 0, s_, op_, cur,
                       b_, (int)CM_RELAX, partial_, (const SolveState*)st_);
    if (int rc = fold(s_, partial_, multi_blocks(op_.nx, op_.ny),
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
  D.cls32 = L.cls32;
  D.row_ptr = L.row_ptr;
  D.ent_lin = L.ent_lin;
  D.ent_di = L.ent_di;
  D.ent_dj = L.ent_dj;
  D.ent_val = (const T*)L.ent_val;
  D.cdinv = (const T*)L.cdinv;
  D.cactive = L.cactive;
  D.rr = L.rr;
  D.rdinv = (const T*)L.rdinv;
  D.ract = L.ract;
  D.fx = L.fx;
  D.fy = L.fy;
  D.cny = (L.ny + L.fy - 1) / L.fy;
  D.spec_e0 = L.spec_e0;
  D.spec_e1 = L.spec_e1;
  return D;
}

template <typename T, int OP, bool PER>
static void cls_rows_reg(hipStream_t s, int reg_lanes, const ClassLevel<T>& L, T w, const T* b,
                         const T* xin, const T* xc, T* o1, T* o2, T* o3, const SolveState* st,
                         const LaunchClock& clk) {
  const dim3 grid((unsigned)(((long)L.n * reg_lanes + kBlock - 1) / kBlock));
#define PCFD_ROWS_REG(LPR)                                                                 \
  case LPR:                                                                                \
    PCFD_LAUNCH_TIMED(clk, (k_cls_rows_reg<T, LPR, OP, PER>), grid, dim3(kBlock), 0, s, L, w, \
                      b, xin, xc, o1, o2, o3, st);                                         \
    break;
  switch (reg_lanes) {
    PCFD_ROWS_REG(1)
    PCFD_ROWS_REG(2)
    PCFD_ROWS_REG(4)
    PCFD_ROWS_REG(8)
    PCFD_ROWS_REG(16)
    PCFD_ROWS_REG(32)
    PCFD_ROWS_REG(64)
    default:
      break;
  }
#undef PCFD_ROWS_REG
}

// Launches one row operation on a coarse level: the register-resident kernel
// when the level's rows fit it (reg_lanes > 0), else the looping one with the
// lane count chosen for its widest row.
template <typename T, int OP>
static void cls_rows(hipStream_t s, int lanes, int reg_lanes, const ClassLevel<T>& L, T w,
                     const T* b, const T* xin, const T* xc, const int* agg, T* o1, T* o2, T* o3,
                     const SolveState* st, const LaunchClock& clk = LaunchClock{}) {
  if (reg_lanes > 0) {
    if (L.px | L.py)
      cls_rows_reg<T, OP, true>(s, reg_lanes, L, w, b, xin, xc, o1, o2, o3, st, clk);
    else
      cls_rows_reg<T, OP, false>(s, reg_lanes, L, w, b, xin, xc, o1, o2, o3, st, clk);
    return;
  }
  const long threads = (long)L.n * lanes;
  const dim3 grid((unsigned)((threads + kBlock - 1) / kBlock));
  switch (lanes) {
    case 1:
      PCFD_LAUNCH_TIMED(clk, (k_cls_rows<T, 1, OP>), grid, dim3(kBlock), 0, s, L, w, b, xin, xc,
                        agg, o1, o2, o3, st);
      break;
    case 4:
      PCFD_LAUNCH_TIMED(clk, (k_cls_rows<T, 4, OP>), grid, dim3(kBlock), 0, s, L, w, b, xin, xc,
                        agg, o1, o2, o3, st);
      break;
    case 16:
      PCFD_LAUNCH_TIMED(clk, (k_cls_rows<T, 16, OP>), grid, dim3(kBlock), 0, s, L, w, b, xin, xc,
                        agg, o1, o2, o3, st);
      break;
    default:
      PCFD_LAUNCH_TIMED(clk, (k_cls_rows<T, 64, OP>), grid, dim3(kBlock), 0, s, L, w, b, xin, xc,
                        agg, o1, o2, o3, st);
      break;
  }
}

template <typename T>
static ClassRows<T> class_rows(const LevelBuf& L, int px, int py) {
  ClassRows<T> D;
  D.nx = L.nx;
  D.ny = L.ny;
  D.px = px;
  D.py = py;
  D.cls = L.cls;
  D.row_ptr = L.row_ptr;
  D.ent_off = L.ent_off;
  D.ent_val = (const T*)L.ent_val;
  D.cdinv = (const T*)L.cdinv;
  D.cactive = L.cactive;
  D.hot = L.hot;
  D.hot_n = L.hot_n;
  D.rdinv = (const T*)L.rdinv;
  D.ract = L.ract;
  D.rr = L.rr;
  D.hot_e0 = L.hot_e0;
  return D;
}

template <typename T>
static BoxRows<T> box_rows(const LevelBuf& L, int px, int py) {
  BoxRows<T> D;
  D.nx = L.nx;
  D.ny = L.ny;
  D.px = px;
  D.py = py;
  D.cls = L.cls;
  D.box = (const T*)L.box;
  D.cdinv = (const T*)L.box_dinv;
  D.cactive = L.box_act;
  D.zero_class = L.ncls;
  return D;
}

template <typename T, bool PER>
static void cls_restrict_reg(hipStream_t s, int reg_lanes, const ClassLevel<T>& L, T w,
                             const T* r, const T* t, int fx, int fy, int cnx, int cny, T* bc,
                             const SolveState* st, const LaunchClock& clk) {
  const long cells = (long)cnx * cny;
  auto grid = [&](int per) { return dim3((unsigned)((cells * per + kBlock - 1) / kBlock)); };
#define PCFD_RESTRICT_REG(LPR)                                                               \
  case LPR:                                                                                  \
    PCFD_LAUNCH_TIMED(clk, (k_cls_restrict_reg<T, LPR, PER>), grid(4 * LPR), dim3(kBlock), 0, \
                      s, L, w, r, t, fx, fy, cnx, cny, bc, st);                              \
    break;
  switch (reg_lanes) {
    PCFD_RESTRICT_REG(1)
    PCFD_RESTRICT_REG(2)
    PCFD_RESTRICT_REG(4)
    PCFD_RESTRICT_REG(8)
    PCFD_RESTRICT_REG(16)
    PCFD_RESTRICT_REG(32)
    PCFD_RESTRICT_REG(64)
    default:
      break;
  }
#undef PCFD_RESTRICT_REG
}

template <typename T>
static void cls_restrict(hipStream_t s, int lanes, int reg_lanes, const ClassLevel<T>& L, T w,
                         const T* r, const T* t, int fx, int fy, int cnx, int cny, T* bc,
                         const SolveState* st, const LaunchClock& clk = LaunchClock{}) {
  const long cells = (long)cnx * cny;
  auto grid = [&](int per) { return dim3((unsigned)((cells * per + kBlock - 1) / kBlock)); };
  if (reg_lanes > 0) {
    if (L.px | L.py)
      cls_restrict_reg<T, true>(s, reg_lanes, L, w, r, t, fx, fy, cnx, cny, bc, st, clk);
    else
      cls_restrict_reg<T, false>(s, reg_lanes, L, w, r, t, fx, fy, cnx, cny, bc, st, clk);
    return;
  }
  switch (lanes) {
    case 1:
      PCFD_LAUNCH_TIMED(clk, (k_cls_restrict<T, 1>), grid(4), dim3(kBlock), 0, s, L, w, r, t, fx,
                        fy, cnx, cny, bc, st);
      break;
    case 4:
      PCFD_LAUNCH_TIMED(clk, (k_cls_restrict<T, 4>), grid(16), dim3(kBlock), 0, s, L, w, r, t, fx,
                        fy, cnx, cny, bc, st);
      break;
    case 16:
      PCFD_LAUNCH_TIMED(clk, (k_cls_restrict<T, 16>), grid(64), dim3(kBlock), 0, s, L, w, r, t,
                        fx, fy, cnx, cny, bc, st);
      break;
    default:
      PCFD_LAUNCH_TIMED(clk, (k_cls_restrict_serial<T, 64>), grid(64), dim3(kBlock), 0, s, L, w, r,
                        t, fx, fy, cnx, cny, bc, st);
      break;
  }
}

// Down leg of a stored level L (right-hand side L.b) towards C: pre-smoothed
// iterate into L.x, restricted residual into C.b.  Fused when the level's
// tile plan says so, else pass by pass.
template <typename T>
int PoissonEngine::level_down(LevelBuf& L, LevelBuf& C, const SolveState* cst, int lev) {
  const int nu = cfg_.sweeps;
  const T w = (T)L.weight;
  // algorithmic bytes for the clock: vectors at T, the class index, the
  // coarse vector; the L2-resident row dictionary counts zero
  const double n = (double)L.n, nc = (double)C.n, ts = (double)sizeof(T);
  const double cb = L.cls32 ? 4.0 : 2.0;
  if (L.box_rows) {
    return launch_class_down_stream<T>(s_, box_rows<T>(L, op_.px, op_.py), L.box_radius,
                                       L.box_rows, w, (const T*)L.b, (T*)L.x, (T*)C.b, L.fx,
                                       L.fy, C.nx, C.ny, cst,
                                       clock_next(CK_CLASS_DOWN, lev, n * (2 * ts + cb) + nc * ts));
  }
  if (L.fused) {
    const ClassRows<T> D = class_rows<T>(L, op_.px, op_.py);
    return launch_class_down<T>(s_, D, L.down0, L.rxs, L.rys, w, nu, (const T*)L.b, (T*)L.x,
                                (T*)C.b, L.fx, L.fy, C.nx, C.ny, cst,
                                clock_next(CK_CLASS_DOWN, lev, n * (2 * ts + cb) + nc * ts));
  }
  const ClassLevel<T> D = class_level<T>(L, op_.px, op_.py);
  const int la = L.lanes, rl = L.row_lanes;
  cls_rows<T, RO_SMOOTH0_RES>(s_, la, rl, D, w, (const T*)L.b, nullptr, nullptr, nullptr,
                              (T*)L.x, (T*)L.r, (T*)L.t, cst,
                              clock_next(CK_ROWS_SMOOTH0, lev, n * (4 * ts + cb)));
  if (nu > 1) {
    for (int sw = 1; sw < nu; ++sw) {
      cls_rows<T, RO_SMOOTH>(s_, la, rl, D, w, (const T*)L.b, (const T*)L.x, nullptr, nullptr,
                             (T*)L.y, nullptr, nullptr, cst,
                             clock_next(CK_ROWS_SMOOTH, lev, n * (3 * ts + cb)));
      std::swap(L.x, L.y);
    }
    cls_rows<T, RO_RES>(s_, la, rl, D, w, (const T*)L.b, (const T*)L.x, nullptr, nullptr,
                        nullptr, (T*)L.r, (T*)L.t, cst,
                        clock_next(CK_ROWS_RES, lev, n * (4 * ts + cb)));
  }
  // b_c = P^T r in one pass over the coarse cells
  cls_restrict<T>(s_, la, rl, D, w, (const T*)L.r, (const T*)L.t, L.fx, L.fy, C.nx, C.ny,
                  (T*)C.b, cst, clock_next(CK_RESTRICT, lev, n * (2 * ts + cb) + nc * ts));
  PCFD_LAUNCH_CHECK();
  return kOk;
}

// Up leg of a stored level: L.y = S^nu (L.x + P C.y).
template <typename T>
int PoissonEngine::level_up(LevelBuf& L, LevelBuf& C, const SolveState* cst, int lev) {
  const int nu = cfg_.sweeps;
  const T w = (T)L.weight;
  const double n = (double)L.n, nc = (double)C.n, ts = (double)sizeof(T);
  const double cb = L.cls32 ? 4.0 : 2.0;
  if (L.box_rows) {
    return launch_class_up_stream<T>(s_, box_rows<T>(L, op_.px, op_.py), L.box_radius,
                                     L.box_rows, w, (const T*)L.x, (const T*)L.b, (const T*)C.y,
                                     L.fx, L.fy, C.ny, (T*)L.y, cst,
                                     clock_next(CK_CLASS_UP, lev, n * (3 * ts + cb) + nc * ts));
  }
  if (L.fused) {
    const ClassRows<T> D = class_rows<T>(L, op_.px, op_.py);
    return launch_class_up<T>(s_, D, L.up, L.rxs, L.rys, w, nu, (const T*)L.x, (const T*)L.b,
                              (const T*)C.y, L.fx, L.fy, C.ny, (T*)L.y, cst,
                              clock_next(CK_CLASS_UP, lev, n * (3 * ts + cb) + nc * ts));
  }
  const ClassLevel<T> D = class_level<T>(L, op_.px, op_.py);
  const int la = L.lanes, rl = L.row_lanes;
  cls_rows<T, RO_PROLONG>(s_, la, rl, D, w, nullptr, nullptr, (const T*)C.y, L.agg, (T*)L.x,
                          nullptr, nullptr, cst,
                          clock_next(CK_ROWS_PROLONG, lev, n * (2 * ts + 4 + cb) + nc * ts));
  for (int sw = 0; sw < nu; ++sw) {
    cls_rows<T, RO_SMOOTH>(s_, la, rl, D, w, (const T*)L.b, (const T*)L.x, nullptr, nullptr,
                           (T*)L.y, nullptr, nullptr, cst,
                           clock_next(CK_ROWS_SMOOTH, lev, n * (3 * ts + cb)));
    std::swap(L.x, L.y);
  }
  std::swap(L.x, L.y);  // the corrected iterate is named y
  PCFD_LAUNCH_CHECK();
  return kOk;
}

// The stored levels from `first` down, the coarsest solve and the way back
// up: per-level legs above the cooperative chain (when the plan has one), the
// chain itself, then the per-level up legs.
template <typename T>
int PoissonEngine::coarse_legs(std::vector<LevelBuf>& lv, int first, const SolveState* cst) {
  const int nl = (int)lv.size();
  const bool cycle_precision = (cfg_.precision == 1) == (sizeof(T) == 4);
  const int cf = (cycle_precision && chain_first_ >= first && chain_first_ < nl) ? chain_first_
                                                                                : nl - 1;
  for (int l = first; l < cf; ++l)
    if (int rc = level_down<T>(lv[l], lv[l + 1], cst, l)) return rc;
  if (cf < nl - 1) {
    if (int rc = launch_chain<T>(lv, cf, cst)) return rc;
  } else {
    for (int l = cf; l < nl - 1; ++l)
      if (int rc = level_down<T>(lv[l], lv[l + 1], cst, l)) return rc;
    LevelBuf& L = lv[nl - 1];
    const double n = (double)L.n;
    if (int rc = launch_dense_apply<T>(s_, L.cinv, L.n, (const T*)L.b, (T*)L.y, cst,
                                       clock_next(CK_DENSE, nl - 1,
                                                  n * n * 8 + 2 * n * sizeof(T))))
      return rc;
  }
  for (int l = cf - 1; l >= first; --l)
    if (int rc = level_up<T>(lv[l], lv[l + 1], cst, l)) return rc;
  return kOk;
}

template <typename T>
int PoissonEngine::launch_chain(std::vector<LevelBuf>& lv, int first, const SolveState* cst) {
  const int nl = (int)lv.size();
  Chain<T> C{};
  C.n = nl - 1 - first;
  C.nu = cfg_.sweeps;
  if (C.n < 1 || C.n > kMaxChain) {
    set_last_error("cooperative chain of %d levels is out of range", C.n);
    return kErrBadArg;
  }
  for (int k = 0; k < C.n; ++k) {
    LevelBuf& L = lv[first + k];
    ChainLevel<T>& c = C.lv[k];
    c.L = class_level<T>(L, op_.px, op_.py);
    c.w = (T)L.weight;
    c.lanes = L.lanes;
    c.fx = L.fx;
    c.fy = L.fy;
    c.b = (T*)L.b;
    c.x = (T*)L.x;
    c.y = (T*)L.y;
    c.r = (T*)L.r;
    c.t = (T*)L.t;
    c.agg = L.agg;
  }
  LevelBuf& Lc = lv[nl - 1];
  C.cn = Lc.n;
  C.cinv = Lc.cinv;
  C.cb = (T*)Lc.b;
  C.cy = (T*)Lc.y;
  C.cnx = Lc.nx;
  C.cny = Lc.ny;
  void* args[] = {(void*)&C, (void*)&cst};
  PCFD_HIP_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_chain<T>),
                                            dim3(chain_blocks_), dim3(kBlock), args,
                                            (unsigned)(Lc.n * sizeof(double)), s_));
  return kOk;
}

// One V-cycle (multigrid.py:430-451) over the levels in precision T.
//
// Buffer contract per level: the down leg reads the level's right-hand side
// b, leaves the pre-smoothed iterate (level 0: t, coarser levels: x) and
// writes the next level's b; the up leg reads those and the next level's
// corrected iterate y and writes its own corrected iterate to y.  The
// coarsest level is the dense pseudo-inverse b -> y.  Level 0's right-hand
// side comes from r_ (Krylov mode; b0 = r - mean_r is kept in lv[0].b) or is
// lv[0].b itself with the iterate lv[0].x (standalone mode).  The cycle
// output lands in lv[0].y (Krylov) or back in lv[0].x (standalone; the up leg
// reads the pre-smoothed copy in lv[0].t, so nothing reads x while it is
// written) with the partial sums `reduce` asks for.
template <typename T>
int PoissonEngine::enqueue_vcycle_t(int mode, int reduce, bool defer_fold) {
  auto& lv = (sizeof(T) == 4) ? lv32_ : lv64_;
  const int nl = (int)lv.size();
  const int nu = cfg_.sweeps;
  const SolveState* cst = st_;
  const bool krylov = mode == VC_KRYLOV;
  LevelBuf& L0 = lv[0];
  (void)nu;
  if (!(krylov && L0.stream_rows && nl > 1))
    if (int rc = flush_update_fold()) return rc;  // no streaming leg to fold it
  T* out = krylov ? (T*)L0.y : (T*)L0.x;
  const int nt0 = ntiles(op_.nx, op_.ny);
  if (nl == 1) {
    if (krylov) {
      hipLaunchKernelGGL(k_project_in<T>, tiles_for(op_.nx, op_.ny), dim3(kBlock), 0, s_, op_,
                         (const double*)r_, (T*)L0.b, cst);
      PCFD_LAUNCH_CHECK();
    }
    if (int rc = launch_dense_apply<T>(s_, L0.cinv, L0.n, (const T*)L0.b, out, cst)) return rc;
    hipLaunchKernelGGL(k_cycle_sums<T>, tiles_for(op_.nx, op_.ny), dim3(kBlock), 0, s_, op_,
                       (const T*)L0.b, (const T*)out, partial_, cst);
    PCFD_LAUNCH_CHECK();
    if (defer_fold) return kOk;
    return fold(s_, partial_, nt0, EpiCycle{st_, reduce, n_proj_});
  }
  // Downward leg.  Algorithmic bytes for the clock: r (FP64) or b_in and
  // x_in, flags, the hand-over vectors and the restricted residual.
  const double n0 = (double)L0.n, ts = (double)sizeof(T);
  if (krylov && L0.stream_rows) {
    LevelBuf& C = lv[1];
    const double bytes_down = n0 * (8 + 1 + 2 * ts) + C.n * ts;
    // the streaming down leg folds a pending update's partials itself
    const UpdateFold uf{upd_partial_, upd_rows_, upd_iter_, cfg_.maxiter, n_proj_};
    upd_rows_ = 0;
    if (int rc = launch_fine_down_stream<T>(s_, op_, L0.stream_rows, (T)L0.weight,
                                            (const double*)r_, (T*)L0.b, (T*)L0.t, (T*)C.b,
                                            L0.fx, L0.fy, C.nx, C.ny, st_, uf,
                                            clock_next(CK_FINE_DOWN, 0, bytes_down)))
      return rc;
    if (int rc = coarse_legs<T>(lv, 1, cst)) return rc;
    const double bytes_up = n0 * (3 * ts + 1) + C.n * ts;
    if (int rc = launch_fine_up_stream<T>(s_, op_, L0.stream_rows, (T)L0.weight,
                                          (const T*)L0.t, (const T*)L0.b, (const T*)C.y, L0.fx,
                                          L0.fy, C.ny, out, reduce, partial_, cst,
                                          clock_next(CK_FINE_UP, 0, bytes_up)))
      return rc;
    if (defer_fold) return kOk;
    return fold(s_, partial_, L0.stream_nb, EpiCycle{st_, reduce, n_proj_},
                clock_next(CK_FOLD, 0, L0.stream_nb * 24.0));
  }
  {
    LevelBuf& C = lv[1];
    const double bytes = (krylov ? n0 * (8 + 1 + 2 * ts) : n0 * (3 * ts + 1)) + C.n * ts;
    const int rc = launch_fine_down<T>(
        s_, op_, krylov ? L0.down0 : L0.down1, (T)L0.weight, nu, krylov ? 0 : 1,
        krylov ? (const double*)r_ : nullptr, krylov ? nullptr : (const T*)L0.b,
        krylov ? nullptr : (const T*)L0.x, krylov ? 0 : 1, krylov ? (T*)L0.b : nullptr,
        (T*)L0.t, (T*)C.b, L0.fx, L0.fy, C.nx, C.ny, cst, clock_next(CK_FINE_DOWN, 0, bytes));
    if (rc) return rc;
  }
  if (int rc = coarse_legs<T>(lv, 1, cst)) return rc;
  {
    LevelBuf& C = lv[1];
    const double bytes = n0 * (3 * ts + 1) + C.n * ts;  // x_pre, b0, flags, z; x_c
    if (int rc = launch_fine_up<T>(s_, op_, L0.up, (T)L0.weight, nu, (const T*)L0.t,
                                   (const T*)L0.b, (const T*)C.y, L0.fx, L0.fy, C.ny, out,
                                   reduce, partial_, cst, clock_next(CK_FINE_UP, 0, bytes)))
      return rc;
  }
  if (defer_fold) return kOk;
  return fold(s_, partial_, L0.up.ntiles(), EpiCycle{st_, reduce, n_proj_},
              clock_next(CK_FOLD, 0, L0.up.ntiles() * 24.0));
}

int PoissonEngine::enqueue_vcycle(int mode, int reduce, bool defer_fold) {
  if (cfg_.precision == 1) return enqueue_vcycle_t<float>(mode, reduce, defer_fold);
  return enqueue_vcycle_t<double>(mode, reduce, defer_fold);
}

// Rows of partials a Krylov-mode cycle leaves (the fold's row count in
// enqueue_vcycle_t).
int PoissonEngine::cycle_rows() const {
  const std::vector<LevelBuf>& lv = (cfg_.precision == 1) ? lv32_ : lv64_;
  if (lv.size() == 1) return ntiles(op_.nx, op_.ny);
  if (lv[0].stream_rows) return lv[0].stream_nb;
  return lv[0].up.ntiles();
}

int PoissonEngine::run_mgcg() {
  const int nt = ntiles(op_.nx, op_.ny);
  const int ub = (nt + kCgTiles - 1) / kCgTiles;  // update workgroups
  // partials: the cycle's sums (folded by the direction kernel), the
  // direction's p.q (folded by the update), the update's r.r, sum(r)
  auto up256 = [](size_t n) { return (n + 255) / 256 * 256; };
  const int crows = cycle_rows();
  const size_t doff = up256(3 * (size_t)crows);
  const size_t uoff = doff + up256((size_t)dir_tile_.ntiles());
  if (int rc = ensure_partials(uoff + 2 * (size_t)ub)) return rc;
  if (int rc = start(IM_MGCG, r_)) return rc;
  double* p_old = pa_;
  double* p_new = pb_;
  const bool f32 = cfg_.precision == 1;
  const int it0 = clock_iter_;
  int n_upd = 0;  // updates enqueued in this solve (their iteration counts)
  upd_rows_ = 0;
  PCFD_ITERATE_HOOK(1, cfg_.maxiter, flush_update_fold(), {
    ++clock_iter_;
    if (int rc = enqueue_vcycle(VC_KRYLOV, SR_MGCG, true)) return rc;
    // direction (folds the cycle's sums itself): z (cycle precision), p_old,
    // flags in; p_new, q out
    const LaunchClock cd = clock_next(CK_DIR, 0, (double)n_ * ((f32 ? 4 : 8) + 3 * 8 + 1));
    const CycleFold cf{partial_, crows, n_proj_};
    const int rc = f32 ? launch_fine_dir<float>(s_, op_, dir_tile_, (const float*)lv32_[0].y,
                                                p_old, p_new, q_, partial_ + doff, st_, cf, cd)
                       : launch_fine_dir<double>(s_, op_, dir_tile_, (const double*)lv64_[0].y,
                                                 p_old, p_new, q_, partial_ + doff, st_, cf, cd);
    if (rc) return rc;
    // the update folds alpha itself and writes its partials past the direction's
    const LaunchClock cu = clock_next(CK_UPDATE, 0, (double)n_ * (6 * 8 + 1));  // x, p, r, q, flags; x, r
    PCFD_LAUNCH_TIMED(cu, k_cg_update, dim3(ub), dim3(kBlock), 0, s_, op_, x_, r_, p_new, q_,
                      (const double*)partial_ + doff, dir_tile_.ntiles(), partial_ + uoff, st_);
    // its partials are folded by the next cycle's down leg, or by
    // flush_update_fold before the host reads the state
    upd_partial_ = partial_ + uoff;
    upd_rows_ = ub;
    upd_iter_ = ++n_upd;
    std::swap(p_old, p_new);
  });
  clock_trim(it0, last_run_);
  return kOk;
}

// The last multigrid-CG update's partials folded by a launch of its own
// (EpiUpdate), when no streaming down leg follows to fold them.
int PoissonEngine::flush_update_fold() {
  if (upd_rows_ <= 0) return kOk;
  const int rows = upd_rows_;
  upd_rows_ = 0;
  return fold(s_, (double*)upd_partial_, rows,
              EpiUpdate{st_, (int)IM_MGCG, cfg_.maxiter, n_proj_},
              clock_next(CK_FOLD, 2, rows * 16.0));
}

int PoissonEngine::run_multigrid() {
  if (int rc = start(IM_MG, nullptr)) return rc;
  LevelBuf& L0 = lv64_[0];
  // the hierarchy's level-0 vectors hold b and the iterate
  const size_t bytes = (size_t)n_ * sizeof(double);
  PCFD_HIP_CHECK(hipMemcpyAsync(L0.b, b_, bytes, hipMemcpyDeviceToDevice, s_));
  PCFD_HIP_CHECK(hipMemcpyAsync(L0.x, x_, bytes, hipMemcpyDeviceToDevice, s_));
  const int it0 = clock_iter_;
  PCFD_ITERATE(1, cfg_.maxiter, {
    ++clock_iter_;
    if (int rc = enqueue_vcycle_t<double>(VC_STANDALONE, SR_MG, false)) return rc;
    hipLaunchKernelGGL(k_check, dim3(multi_blocks(op_.nx, op_.ny)), dim3(kBlock), 0, s_, op_,
                       (const double*)L0.x, b_, (int)CM_MG, partial_, (const SolveState*)st_);
    if (int rc = fold(s_, partial_, multi_blocks(op_.nx, op_.ny),
                      EpiCheck{st_, (int)CM_MG, cfg_.tol, cfg_.maxiter, singular_, n_fluid_}))
      return rc;
  });
  clock_trim(it0, last_run_);
  // The pending mean shift is applied by the final residual check and the scatter.
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
  if (int rc = enqueue_vcycle(VC_KRYLOV, SR_MGCG)) return rc;
  if (cfg_.precision == 1) {
    hipLaunchKernelGGL(k_cast_out<float>, g, dim3(kBlock), 0, s_, op_, (const float*)lv32_[0].y,
                       (const SolveState*)st_, z);
  } else {
    hipLaunchKernelGGL(k_cast_out<double>, g, dim3(kBlock), 0, s_, op_,
                       (const double*)lv64_[0].y, (const SolveState*)st_, z);
  }
  PCFD_LAUNCH_CHECK();
  PCFD_HIP_CHECK(hipStreamSynchronize(s_));
  return kOk;
}

int PoissonEngine::time_kernel(int which, int reps, double* avg_ms, double* bytes) {
  const dim3 g = tiles_for(op_.nx, op_.ny);
  const double n = (double)n_;
  const bool have_mg = lv64_.size() >= 2;
  if (which == 1 && dir_tile_.ntiles() == 0) {
    set_last_error("kernel 1 needs a multigrid-CG solver configured");
    return kErrBadArg;
  }
  if ((which == 2 || which == 3) && !have_mg) {
    set_last_error("kernel %d needs a multigrid hierarchy of at least two levels", which);
    return kErrBadArg;
  }
  SolveState* st = nullptr;
  PCFD_HIP_CHECK(hipMalloc(&st, sizeof(SolveState)));
  PCFD_HIP_CHECK(hipMemsetAsync(st, 0, sizeof(SolveState), s_));
  hipEvent_t e0, e1;
  PCFD_HIP_CHECK(hipEventCreate(&e0));
  PCFD_HIP_CHECK(hipEventCreate(&e1));
  double B = 0.0;
  auto launch = [&]() -> int {
    switch (which) {
      case 0:  // CG update (the separate-fold form): read x, p, r, q + flags, write x, r
        hipLaunchKernelGGL(k_jpcg_update, g, dim3(kBlock), 0, s_, op_, x_, r_, pa_, q_,
                           (int)IM_MGCG, partial_, (const SolveState*)st);
        B = n * (6 * 8 + 1);
        break;
      case 1:  // multigrid-CG direction + SpMV: read z, p_old + flags, write p, q
        if (int rc = launch_fine_dir<double>(s_, op_, dir_tile_, (const double*)z_, pa_, pb_, q_,
                                             partial_, st))
          return rc;
        B = n * (4 * 8 + 1);
        break;
      case 2: {  // level-0 down leg: read r + flags, write b0, x, and b_c (n/4)
        LevelBuf& L = lv64_[0];
        LevelBuf& C = lv64_[1];
        const int rc =
            L.stream_rows
                ? launch_fine_down_stream<double>(s_, op_, L.stream_rows, L.weight, r_,
                                                  (double*)L.b, (double*)L.t, (double*)C.b, L.fx,
                                                  L.fy, C.nx, C.ny, st)
                : launch_fine_down<double>(s_, op_, L.down0, L.weight, cfg_.sweeps, 0, r_,
                                           nullptr, nullptr, 0, (double*)L.b, (double*)L.t,
                                           (double*)C.b, L.fx, L.fy, C.nx, C.ny, st);
        if (rc) return rc;
        B = n * (8 + 1 + 8 + 8) + (double)C.n * 8;
        break;
      }
      case 3: {  // level-0 up leg: read x, b0, flags, x_c (n/4), write z
        LevelBuf& L = lv64_[0];
        LevelBuf& C = lv64_[1];
        const int rc =
            L.stream_rows
                ? launch_fine_up_stream<double>(s_, op_, L.stream_rows, L.weight,
                                                (const double*)L.t, (const double*)L.b,
                                                (const double*)C.y, L.fx, L.fy, C.ny,
                                                (double*)L.y, (int)SR_MGCG, partial_, st)
                : launch_fine_up<double>(s_, op_, L.up, L.weight, cfg_.sweeps, (const double*)L.t,
                                         (const double*)L.b, (const double*)C.y, L.fx, L.fy,
                                         C.ny, (double*)L.y, (int)SR_MGCG, partial_, st);
        if (rc) return rc;
        B = n * (8 + 8 + 1 + 8) + (double)C.n * 8;
        break;
      }
      default:
        set_last_error("unknown kernel %d", which);
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
