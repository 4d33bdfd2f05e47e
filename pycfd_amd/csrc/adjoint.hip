// A differentiable projection step with a soft immersed-boundary mask, and
// its hand-written adjoint.
//
// The reference forces a body by zeroing every face that touches a solid
// cell after the predictor and after the projection, and books the removed
// momentum as the reaction force (pycfd/core/solver.py:553-599).  Here the
// solid is a per-cell field m in [0, 1] (m = sigmoid(theta)), a face's weight
// is the soft "touches a solid" m_f = 1 - (1 - m_a)(1 - m_b) of its two
// cells, the forcing is u_f <- (1 - m_f) u_f and the force is sum m_f u_f
// A / dt.  With m in {0, 1} this is the reference's binary forcing exactly.
// The Poisson operator keeps the topology of a fixed binary set (the cells
// the caller marks solid), and every face touching that set is held at weight
// 1 as in the binary step, so the soft weights live on the faces between two
// operator cells (the band around the body) and the step is a smooth function
// of m there.
//
// The adjoint runs the step's operations backwards, each one transposed
// exactly: the wall conditions (line assignments run in reverse), the
// momentum right-hand side (its Jacobian, gathered per input so no atomics
// are needed), the projection I - G A^-1 D (G^T, the same symmetric Poisson
// solve, D^T), the soft forcing with its force term, and the Runge-Kutta
// combinations.  Uniform meshes, central advection, constant viscosity.
#include <cstring>
#include <vector>

#include "../../include/pycfd_hip.h"
#include "poisson.hpp"
#include "poisson_dev.hpp"

namespace pcfd {

namespace {

struct SoftGrid {
  int nx, ny, ldu, ldv;
  UpdRange R;
  const uint8_t* solid;  // cells outside the Poisson operator (the fixed binary set)
};

__device__ __forceinline__ bool solid_cell(const SoftGrid& G, int ci, int cj) {
  return ci >= 0 && ci < G.nx && cj >= 0 && cj < G.ny && G.solid[ci * G.ny + cj] != 0;
}

__device__ __forceinline__ bool u_upd(const SoftGrid& G, int m, int j) {
  return m >= G.R.u_lo && m < G.R.u_hi && j >= 1 && j <= G.ny;
}
__device__ __forceinline__ bool v_upd(const SoftGrid& G, int i, int n) {
  return i >= 1 && i <= G.nx && n >= G.R.v_lo && n < G.R.v_hi;
}

// Cell value of m with zero outside the grid (a face on a wall has one cell).
__device__ __forceinline__ double m_cell(const double* m, int nx, int ny, int ci, int cj) {
  return (ci >= 0 && ci < nx && cj >= 0 && cj < ny) ? m[ci * ny + cj] : 0.0;
}

// Face weights of the solvable faces: u face m (cells m-2, m-1 along x,
// solver.py:62-79), v face n (cells n-2, n-1 along y); zero elsewhere.  A
// face touching a cell outside the Poisson operator is held at 1, as the
// binary forcing holds it: the projection corrects every solvable face, and
// only a face between two operator cells may keep flux the operator sees.
__global__ void __launch_bounds__(kBlock)
k_soft_faces(SoftGrid G, const double* __restrict__ m, double* __restrict__ mfu,
             double* __restrict__ mfv) {
  const Tile2D t = tile_coords(G.ny + 3);
  const int i = t.i, j = t.j;
  if (i <= G.nx + 2 && j <= G.ny + 1) {
    double w = 0.0;
    if (u_upd(G, i, j)) {
      if (solid_cell(G, i - 2, j - 1) || solid_cell(G, i - 1, j - 1)) {
        w = 1.0;
      } else {
        const double a = m_cell(m, G.nx, G.ny, i - 2, j - 1);
        const double b = m_cell(m, G.nx, G.ny, i - 1, j - 1);
        w = 1.0 - (1.0 - a) * (1.0 - b);
      }
    }
    mfu[(size_t)i * G.ldu + j] = w;
  }
  if (i <= G.nx + 1 && j <= G.ny + 2) {
    double w = 0.0;
    if (v_upd(G, i, j)) {
      if (solid_cell(G, i - 1, j - 2) || solid_cell(G, i - 1, j - 1)) {
        w = 1.0;
      } else {
        const double a = m_cell(m, G.nx, G.ny, i - 1, j - 2);
        const double b = m_cell(m, G.nx, G.ny, i - 1, j - 1);
        w = 1.0 - (1.0 - a) * (1.0 - b);
      }
    }
    mfv[(size_t)i * G.ldv + j] = w;
  }
}

// Soft forcing: partials of sum m_f u_f (x) and sum m_f v_f (y) over the
// faces, then u_f <- (1 - m_f) u_f.
__global__ void __launch_bounds__(kBlock)
k_soft_mask(SoftGrid G, const double* __restrict__ mfu, const double* __restrict__ mfv,
            double* __restrict__ u, double* __restrict__ v, double* partial) {
  const Tile2D t = tile_coords(G.ny + 3);
  const int i = t.i, j = t.j;
  double s[2] = {0.0, 0.0};
  if (i <= G.nx + 2 && j <= G.ny + 1) {
    const size_t k = (size_t)i * G.ldu + j;
    const double w = mfu[k], x = u[k];
    s[0] = w * x;
    u[k] = (1.0 - w) * x;
  }
  if (i <= G.nx + 1 && j <= G.ny + 2) {
    const size_t k = (size_t)i * G.ldv + j;
    const double w = mfv[k], x = v[k];
    s[1] = w * x;
    v[k] = (1.0 - w) * x;
  }
  const int ops[2] = {kSum, kSum};
  block_partials<2>(s, ops, partial, t.tile);
}

struct EpiSoftForce {  // force = (reset ? 0 : force) + area / dt * sums
  static constexpr int K = 2;
  double* force;
  double scale;
  int reset;
  __device__ void ops(int (&o)[K]) const { o[0] = o[1] = kSum; }
  __device__ bool active() const { return true; }
  __device__ void operator()(const double (&v)[K]) const {
    const double fx = v[0] * scale, fy = v[1] * scale;
    force[0] = reset ? fx : force[0] + fx;
    force[1] = reset ? fy : force[1] + fy;
  }
};

// Transpose of the soft forcing at one application, given the state s before
// it and the adjoint after it: lam <- (1 - m_f) lam + cx m_f (u faces),
// g_f += -s lam + cx s (u faces; v faces without the force term), with cx =
// dL/dF_x * area / dt.
__global__ void __launch_bounds__(kBlock)
k_soft_mask_adj(SoftGrid G, const double* __restrict__ mfu, const double* __restrict__ mfv,
                const double* __restrict__ su, const double* __restrict__ sv, double cx,
                double cy, double* __restrict__ lu, double* __restrict__ lv,
                double* __restrict__ gu, double* __restrict__ gv) {
  const Tile2D t = tile_coords(G.ny + 3);
  const int i = t.i, j = t.j;
  if (i <= G.nx + 2 && j <= G.ny + 1) {
    const size_t k = (size_t)i * G.ldu + j;
    const double w = mfu[k], l = lu[k], x = su[k];
    lu[k] = (1.0 - w) * l + cx * w;
    gu[k] += cx * x - x * l;
  }
  if (i <= G.nx + 1 && j <= G.ny + 2) {
    const size_t k = (size_t)i * G.ldv + j;
    const double w = mfv[k], l = lv[k], x = sv[k];
    lv[k] = (1.0 - w) * l + cy * w;
    gv[k] += cy * x - x * l;
  }
}

// d m_f / d m_a = 1 - m_b: face adjoints to the cells, then
// d m / d theta = m (1 - m).
__global__ void __launch_bounds__(kBlock)
k_soft_cells_adj(SoftGrid G, const double* __restrict__ m, const double* __restrict__ gu,
                 const double* __restrict__ gv, double* __restrict__ gtheta) {
  Tile2D t;
  if (!cell_of_tile(G.nx, G.ny, t)) return;
  const int ci = t.i, cj = t.j;
  double g = 0.0;
  // u faces touching cell (ci, cj): m = ci + 1 (cell is m-2 side... the cell
  // is cell m-1 of face m = ci + 1, and cell m-2 of face m = ci + 2)
  // (faces touching a cell outside the operator have the constant weight 1)
  if (!solid_cell(G, ci, cj)) {
    const int j = cj + 1;
    int mm = ci + 1;  // cell = m - 1, other = m - 2
    if (u_upd(G, mm, j) && !solid_cell(G, ci - 1, cj))
      g += gu[(size_t)mm * G.ldu + j] * (1.0 - m_cell(m, G.nx, G.ny, ci - 1, cj));
    mm = ci + 2;      // cell = m - 2, other = m - 1
    if (u_upd(G, mm, j) && !solid_cell(G, ci + 1, cj))
      g += gu[(size_t)mm * G.ldu + j] * (1.0 - m_cell(m, G.nx, G.ny, ci + 1, cj));
    const int i = ci + 1;
    int n = cj + 1;
    if (v_upd(G, i, n) && !solid_cell(G, ci, cj - 1))
      g += gv[(size_t)i * G.ldv + n] * (1.0 - m_cell(m, G.nx, G.ny, ci, cj - 1));
    n = cj + 2;
    if (v_upd(G, i, n) && !solid_cell(G, ci, cj + 1))
      g += gv[(size_t)i * G.ldv + n] * (1.0 - m_cell(m, G.nx, G.ny, ci, cj + 1));
  }
  const double mc = m[ci * G.ny + cj];
  gtheta[ci * G.ny + cj] += g * (mc * (1.0 - mc));
}

// ---------------------------------------------------------------------------
// Jacobian-transpose of the predictor s = w + dt R(w) on the solvable faces
// (solver.py:334-465 with the central scheme), gathered per input entry:
// every output face whose stencil reads the entry contributes its partial
// times dt * lam_s of that face.  lam_w = lam_s + J^T (dt lam_s).
// ---------------------------------------------------------------------------
struct AdjRhs {
  const double* u;
  const double* v;
  const double* lu;  // lam_s
  const double* lv;
  SoftGrid G;
  double dx, dy, nu, dt;
  __device__ double U(int i, int j) const { return u[(size_t)i * G.ldu + j]; }
  __device__ double V(int i, int j) const { return v[(size_t)i * G.ldv + j]; }
  __device__ double mu_u(int m, int j) const {
    return u_upd(G, m, j) ? dt * lu[(size_t)m * G.ldu + j] : 0.0;
  }
  __device__ double mu_v(int i, int n) const {
    return v_upd(G, i, n) ? dt * lv[(size_t)i * G.ldv + n] : 0.0;
  }
};

// Partials of R_u at face (m, j) (its stencil's local quantities).
struct UPart {
  double d0, dp, dm, du, dd, dvt, dvb;  // u0, up, um, uu, ud, each v_top node, each v_bot node
};
__device__ UPart u_partials(const AdjRhs& A, int m, int j) {
  const double u0 = A.U(m, j), um = A.U(m - 1, j), up = A.U(m + 1, j);
  const double ud = A.U(m, j - 1), uu = A.U(m, j + 1);
  const double uu_r = 0.5 * (u0 + up), uu_l = 0.5 * (um + u0);
  const double u_top = 0.5 * (u0 + uu), u_bot = 0.5 * (ud + u0);
  const double v_top = 0.5 * (A.V(m - 1, j + 1) + A.V(m, j + 1));
  const double v_bot = 0.5 * (A.V(m - 1, j) + A.V(m, j));
  const double cx = A.nu / (A.dx * A.dx), cy = A.nu / (A.dy * A.dy);
  UPart P;
  P.d0 = -((uu_r - uu_l) / A.dx + 0.5 * (v_top - v_bot) / A.dy) - 2.0 * cx - 2.0 * cy;
  P.dp = -uu_r / A.dx + cx;
  P.dm = uu_l / A.dx + cx;
  P.du = -0.5 * v_top / A.dy + cy;
  P.dd = 0.5 * v_bot / A.dy + cy;
  P.dvt = -0.5 * u_top / A.dy;
  P.dvb = 0.5 * u_bot / A.dy;
  return P;
}

struct VPart {
  double d0, dr, dl, dp, dm, dur, dul;  // v0, vr, vl, vp, vm, each u_right node, each u_left node
};
__device__ VPart v_partials(const AdjRhs& A, int i, int n) {
  const double v0 = A.V(i, n), vm = A.V(i, n - 1), vp = A.V(i, n + 1);
  const double vl = A.V(i - 1, n), vr = A.V(i + 1, n);
  const double vv_t = 0.5 * (v0 + vp), vv_b = 0.5 * (vm + v0);
  const double u_right = 0.5 * (A.U(i + 1, n - 1) + A.U(i + 1, n));
  const double u_left = 0.5 * (A.U(i, n - 1) + A.U(i, n));
  const double v_right = 0.5 * (v0 + vr), v_left = 0.5 * (vl + v0);
  const double cx = A.nu / (A.dx * A.dx), cy = A.nu / (A.dy * A.dy);
  VPart P;
  P.d0 = -(0.5 * (u_right - u_left) / A.dx + (vv_t - vv_b) / A.dy) - 2.0 * cx - 2.0 * cy;
  P.dr = -0.5 * u_right / A.dx + cx;
  P.dl = 0.5 * u_left / A.dx + cx;
  P.dp = -vv_t / A.dy + cy;
  P.dm = vv_b / A.dy + cy;
  P.dur = -0.5 * v_right / A.dx;
  P.dul = 0.5 * v_left / A.dx;
  return P;
}

__global__ void __launch_bounds__(kBlock)
k_rhs_adjoint(AdjRhs A, double* __restrict__ ou, double* __restrict__ ov) {
  const SoftGrid& G = A.G;
  const Tile2D t = tile_coords(G.ny + 3);
  const int a = t.i, b = t.j;
  if (a <= G.nx + 2 && b <= G.ny + 1) {  // input U(a, b)
    double g = A.lu[(size_t)a * G.ldu + b];
    double mu;
    if ((mu = A.mu_u(a, b)) != 0.0) g += u_partials(A, a, b).d0 * mu;
    if ((mu = A.mu_u(a - 1, b)) != 0.0) g += u_partials(A, a - 1, b).dp * mu;
    if ((mu = A.mu_u(a + 1, b)) != 0.0) g += u_partials(A, a + 1, b).dm * mu;
    if ((mu = A.mu_u(a, b - 1)) != 0.0) g += u_partials(A, a, b - 1).du * mu;
    if ((mu = A.mu_u(a, b + 1)) != 0.0) g += u_partials(A, a, b + 1).dd * mu;
    // as U(i+1, n-1) / U(i+1, n) of v face (a-1, b+1) / (a-1, b), and as
    // U(i, n-1) / U(i, n) of v face (a, b+1) / (a, b)
    if ((mu = A.mu_v(a - 1, b + 1)) != 0.0) g += v_partials(A, a - 1, b + 1).dur * mu;
    if ((mu = A.mu_v(a - 1, b)) != 0.0) g += v_partials(A, a - 1, b).dur * mu;
    if ((mu = A.mu_v(a, b + 1)) != 0.0) g += v_partials(A, a, b + 1).dul * mu;
    if ((mu = A.mu_v(a, b)) != 0.0) g += v_partials(A, a, b).dul * mu;
    ou[(size_t)a * G.ldu + b] = g;
  }
  if (a <= G.nx + 1 && b <= G.ny + 2) {  // input V(a, b)
    double g = A.lv[(size_t)a * G.ldv + b];
    double mu;
    if ((mu = A.mu_v(a, b)) != 0.0) g += v_partials(A, a, b).d0 * mu;
    if ((mu = A.mu_v(a - 1, b)) != 0.0) g += v_partials(A, a - 1, b).dr * mu;
    if ((mu = A.mu_v(a + 1, b)) != 0.0) g += v_partials(A, a + 1, b).dl * mu;
    if ((mu = A.mu_v(a, b - 1)) != 0.0) g += v_partials(A, a, b - 1).dp * mu;
    if ((mu = A.mu_v(a, b + 1)) != 0.0) g += v_partials(A, a, b + 1).dm * mu;
    // as V(m-1, j+1) / V(m, j+1) of u face (a+1, b-1) / (a, b-1), and as
    // V(m-1, j) / V(m, j) of u face (a+1, b) / (a, b)
    if ((mu = A.mu_u(a + 1, b - 1)) != 0.0) g += u_partials(A, a + 1, b - 1).dvt * mu;
    if ((mu = A.mu_u(a, b - 1)) != 0.0) g += u_partials(A, a, b - 1).dvt * mu;
    if ((mu = A.mu_u(a + 1, b)) != 0.0) g += u_partials(A, a + 1, b).dvb * mu;
    if ((mu = A.mu_u(a, b)) != 0.0) g += u_partials(A, a, b).dvb * mu;
    ov[(size_t)a * G.ldv + b] = g;
  }
}

// ---------------------------------------------------------------------------
// Transposes of the projection's linear pieces.  The step computes
// s' = s - G Z A^-1 Z D s (dt cancels: b = D s / dt, s' = s - dt G p), with D
// the divergence of every fluid cell, Z zeroing the non-solvable rows, and G
// the correction on the solvable faces from the ghosted pressure (solid
// cells 0, a wall ghost the interior value or, at a pressure outlet,
// 2 p_ref - p_in with p_ref = 0).
// ---------------------------------------------------------------------------
struct ProjT {
  SoftGrid G;
  const uint8_t* flags;  // F_SOLV of each cell
  int kind[4];           // wall kinds
  double dx, dy;
};

// Weight of interior cell c in the ghosted pressure at ghosted index q of an
// axis of n cells (1..n interior, 0 and n+1 ghosts): 1 on itself, and for a
// ghost the wall rule's coefficient on its neighbour.
__device__ __forceinline__ double ghost_w(int q, int c, int n, int kind_lo, int kind_hi) {
  if (q == c) return 1.0;
  if (q == 0 && c == 1) return kind_lo == BC_PRESSURE_OUTLET ? -1.0 : 1.0;
  if (q == n + 1 && c == n) return kind_hi == BC_PRESSURE_OUTLET ? -1.0 : 1.0;
  return 0.0;
}

// g = Z G^T lam (cells, (nx, ny)).
__global__ void __launch_bounds__(kBlock)
k_grad_t(ProjT T, const double* __restrict__ lu, const double* __restrict__ lv,
         double* __restrict__ g) {
  Tile2D t;
  if (!cell_of_tile(T.G.nx, T.G.ny, t)) return;
  const int a = t.i + 1, b = t.j + 1;  // ghosted cell
  const SoftGrid& G = T.G;
  double acc = 0.0;
  if (T.flags[t.i * G.ny + t.j] & F_SOLV) {
    // u faces m whose correction (p[m] - p[m-1]) / dx reads cell a
    for (int m = a - 1; m <= a + 2; ++m) {
      if (!u_upd(G, m, b)) continue;
      const double w = ghost_w(m, a, G.nx, T.kind[0], T.kind[1]) -
                       ghost_w(m - 1, a, G.nx, T.kind[0], T.kind[1]);
      if (w != 0.0) acc += w * lu[(size_t)m * G.ldu + b] / T.dx;
    }
    for (int n = b - 1; n <= b + 2; ++n) {
      if (!v_upd(G, a, n)) continue;
      const double w = ghost_w(n, b, G.ny, T.kind[2], T.kind[3]) -
                       ghost_w(n - 1, b, G.ny, T.kind[2], T.kind[3]);
      if (w != 0.0) acc += w * lv[(size_t)a * G.ldv + n] / T.dy;
    }
  }
  g[t.i * G.ny + t.j] = acc;
}

// lam -= D^T Z q: u face m gets (q[m-1] - q[m]) / dx over the interior cells
// m-1 and m (ghosted), v faces likewise; q is zero off the solvable set.
__global__ void __launch_bounds__(kBlock)
k_div_t_sub(ProjT T, const double* __restrict__ q, double* __restrict__ lu,
            double* __restrict__ lv) {
  const SoftGrid& G = T.G;
  const Tile2D t = tile_coords(G.ny + 3);
  const int i = t.i, j = t.j;
  auto Q = [&](int a, int b) -> double {  // ghosted cell, zero outside / off the solvable set
    if (a < 1 || a > G.nx || b < 1 || b > G.ny) return 0.0;
    const int k = (a - 1) * G.ny + (b - 1);
    return (T.flags[k] & F_SOLV) ? q[k] : 0.0;
  };
  if (i >= 1 && i <= G.nx + 1 && j >= 1 && j <= G.ny)
    lu[(size_t)i * G.ldu + j] -= (Q(i - 1, j) - Q(i, j)) / T.dx;
  if (i >= 1 && i <= G.nx && j >= 1 && j <= G.ny + 1)
    lv[(size_t)i * G.ldv + j] -= (Q(i, j - 1) - Q(i, j)) / T.dy;
}

// out = a x + b y (whole arrays)
__global__ void __launch_bounds__(kBlock)
k_axpby(const double* __restrict__ x, const double* __restrict__ y, double a, double b,
        double* __restrict__ out, size_t n) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = a * x[k] + (y ? b * y[k] : 0.0);
}

}  // namespace

}  // namespace pcfd

using namespace pcfd;

// Host side: the context's pieces this file needs (capi.hip owns the
// context; these accessors expose what the differentiable step uses).
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
  double* scratch_a[2];  // stage arrays
  double* scratch_b[2];
  double* work;          // (nx*ny) cell scratch
  const uint8_t* flags;
  double* guess_backup;
  int kind[4];
  int scheme;
  const uint8_t* solid;
};

int pcfd_soft_view(pcfd_ctx* c, SoftView* out);  // capi.hip

namespace {

int usize(const SoftView& V) { return (V.nx + 3) * V.ldu; }
int vsize(const SoftView& V) { return (V.nx + 2) * V.ldv; }

SoftGrid grid_of(const SoftView& V) {
  return SoftGrid{V.nx, V.ny, V.ldu, V.ldv, V.upd, V.solid};
}

int axpby(const SoftView& V, const double* x, const double* y, double a, double b, double* out,
          size_t n) {
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     V.stream, x, y, a, b, out, n);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int copy(const SoftView& V, double* dst, const double* src, size_t n) {
  PCFD_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, V.stream));
  return kOk;
}

int soft_mask(const SoftView& V, const double* mfu, const double* mfv, double* u, double* v,
              double dt, bool reset) {
  const dim3 g((unsigned)num_tiles(V.nx + 3, V.ny + 3));
  double* partial = V.engine->partial_;
  hipLaunchKernelGGL(k_soft_mask, g, dim3(kBlock), 0, V.stream, grid_of(V), mfu, mfv, u, v,
                     partial);
  PCFD_LAUNCH_CHECK();
  const EpiSoftForce e{V.force, V.dx * V.dy / dt, reset ? 1 : 0};
  hipLaunchKernelGGL(k_fold<EpiSoftForce>, dim3(1), dim3(kBlock), 0, V.stream,
                     (const double*)partial, (int)g.x, e);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

// Tape layout of one step: per substep k (0..2) six arrays w_u, w_v, s1_u,
// s1_v, s3_u, s3_v (the substep's input, the state before its first masking,
// the state before its second).
size_t tape_doubles(const SoftView& V) { return 3 * 3 * (size_t)(usize(V) + vsize(V)); }
double* tape_at(const SoftView& V, double* tape, int k, int which, int comp) {
  const size_t uv = (size_t)usize(V) + vsize(V);
  return tape + ((size_t)k * 3 + which) * uv + (comp ? usize(V) : 0);
}

int project(const SoftView& V, double* u, double* v, double* p, double dt) {
  const StepScalars S{dt, 0.0, nullptr};
  PoissonEngine& E = *V.engine;
  if (int rc = E.prepare_from_velocity(u, v, V.ldu, V.ldv, V.dx, V.dy, S)) return rc;
  if (int rc = E.solve(nullptr, false)) return rc;
  if (int rc = E.scatter_pressure(p, V.ldp)) return rc;
  return launch_correct_ghosts(V.stream, p, V.ldp, *V.bc, V.ldu, V.ldv, V.dx, V.dy, *V.g, V.upd,
                               S, u, v);
}

BCOps ops1(int op) {
  BCOps o{};
  o.op[o.n++] = op;
  return o;
}

// One soft-masked substep from w (or ca w + cb we); states for the tape.
int soft_substep(const SoftView& V, const double* w_u, const double* w_v, const double* we_u,
                 const double* we_v, double ca, double cb, double dt, const double* mfu,
                 const double* mfv, double* su, double* sv, double* p, double* tape, int k) {
  const StepScalars S{dt, 0.0, nullptr};
  int rc;
  const double* iu = w_u;
  const double* iv = w_v;
  if (tape) {  // the substep's input, materialised (the same values the fused loads form)
    double* tu = tape_at(V, tape, k, 0, 0);
    double* tv = tape_at(V, tape, k, 0, 1);
    if (we_u) {
      if ((rc = axpby(V, w_u, we_u, ca, cb, tu, usize(V)))) return rc;
      if ((rc = axpby(V, w_v, we_v, ca, cb, tv, vsize(V)))) return rc;
    } else {
      if ((rc = copy(V, tu, w_u, usize(V))) || (rc = copy(V, tv, w_v, vsize(V)))) return rc;
    }
    iu = tu;
    iv = tv;
    we_u = we_v = nullptr;
  }
  if (we_u) rc = launch_predictor_comb(V.stream, iu, we_u, iv, we_v, ca, cb, *V.mom, V.upd, S, su, sv);
  else rc = launch_predictor(V.stream, iu, iv, *V.mom, V.upd, S, su, sv);
  if (rc) return rc;
  if ((rc = launch_boundary_pass(V.stream, *V.bc, ops1(BOP_VEL_PRED), S, su, sv))) return rc;
  if (tape && ((rc = copy(V, tape_at(V, tape, k, 1, 0), su, usize(V))) ||
               (rc = copy(V, tape_at(V, tape, k, 1, 1), sv, vsize(V)))))
    return rc;
  if ((rc = soft_mask(V, mfu, mfv, su, sv, dt, true))) return rc;
  if ((rc = project(V, su, sv, p, dt))) return rc;
  if (tape && ((rc = copy(V, tape_at(V, tape, k, 2, 0), su, usize(V))) ||
               (rc = copy(V, tape_at(V, tape, k, 2, 1), sv, vsize(V)))))
    return rc;
  if ((rc = soft_mask(V, mfu, mfv, su, sv, dt, false))) return rc;
  return launch_boundary_pass(V.stream, *V.bc, ops1(BOP_VEL_CORR), S, su, sv);
}

ProjT projt_of(const SoftView& V) {
  ProjT T{grid_of(V), V.flags, {V.kind[0], V.kind[1], V.kind[2], V.kind[3]}, V.dx, V.dy};
  return T;
}

// Transpose of one substep: lam (lu, lv) enters as the adjoint of the
// substep's output and leaves as that of its input; cx = dL/dF_x of this
// substep's force.  The pressure solves of the transposed projection start
// from zero and use the configured solver.
int soft_substep_adjoint(const SoftView& V, const double* tape, int k, double dt,
                         const double* mfu, const double* mfv, double cx, double* lu, double* lv,
                         double* gu, double* gv) {
  int rc;
  const SoftGrid G = grid_of(V);
  const dim3 ga((unsigned)num_tiles(V.nx + 3, V.ny + 3));
  const double* wu = tape_at(V, const_cast<double*>(tape), k, 0, 0);
  const double* wv = tape_at(V, const_cast<double*>(tape), k, 0, 1);
  const double* s1u = tape_at(V, const_cast<double*>(tape), k, 1, 0);
  const double* s1v = tape_at(V, const_cast<double*>(tape), k, 1, 1);
  const double* s3u = tape_at(V, const_cast<double*>(tape), k, 2, 0);
  const double* s3v = tape_at(V, const_cast<double*>(tape), k, 2, 1);
  const double area = V.dx * V.dy;
  // corrector walls
  if ((rc = launch_walls_adjoint(V.stream, *V.bc, false, lu, lv))) return rc;
  // second masking (force term)
  hipLaunchKernelGGL(k_soft_mask_adj, ga, dim3(kBlock), 0, V.stream, G, mfu, mfv, s3u, s3v,
                     cx * area / dt, 0.0, lu, lv, gu, gv);
  PCFD_LAUNCH_CHECK();
  // projection: lam -= D^T Z A^-1 Z G^T lam
  const ProjT T = projt_of(V);
  const dim3 gc((unsigned)num_tiles(V.nx, V.ny));
  hipLaunchKernelGGL(k_grad_t, gc, dim3(kBlock), 0, V.stream, T, (const double*)lu,
                     (const double*)lv, V.work);
  PCFD_LAUNCH_CHECK();
  PoissonEngine& E = *V.engine;
  const size_t bytes = (size_t)V.nx * V.ny * sizeof(double);
  PCFD_HIP_CHECK(hipMemcpyAsync(V.guess_backup, E.guess(), bytes, hipMemcpyDeviceToDevice,
                                V.stream));
  if ((rc = E.reset_guess())) return rc;
  if ((rc = E.prepare_from_rhs(V.work))) return rc;
  if ((rc = E.solve(nullptr, false))) return rc;
  if ((rc = E.write_interior(V.work))) return rc;
  PCFD_HIP_CHECK(hipMemcpyAsync(E.guess(), V.guess_backup, bytes, hipMemcpyDeviceToDevice,
                                V.stream));
  hipLaunchKernelGGL(k_div_t_sub, ga, dim3(kBlock), 0, V.stream, T, (const double*)V.work, lu, lv);
  PCFD_LAUNCH_CHECK();
  // first masking
  hipLaunchKernelGGL(k_soft_mask_adj, ga, dim3(kBlock), 0, V.stream, G, mfu, mfv, s1u, s1v,
                     cx * area / dt, 0.0, lu, lv, gu, gv);
  PCFD_LAUNCH_CHECK();
  // predictor walls, then the predictor itself
  if ((rc = launch_walls_adjoint(V.stream, *V.bc, true, lu, lv))) return rc;
  double* tu = V.scratch_b[0];
  double* tv = V.scratch_b[1];
  const AdjRhs A{wu, wv, lu, lv, G, V.dx, V.dy, V.nu, dt};
  hipLaunchKernelGGL(k_rhs_adjoint, ga, dim3(kBlock), 0, V.stream, A, tu, tv);
  PCFD_LAUNCH_CHECK();
  if ((rc = copy(V, lu, tu, usize(V))) || (rc = copy(V, lv, tv, vsize(V)))) return rc;
  return kOk;
}

int check_soft(pcfd_ctx* c, SoftView& V) {
  if (!c) {
    set_last_error("null context");
    return kErrBadArg;
  }
  if (int rc = pcfd_soft_view(c, &V)) return rc;
  if (V.scheme != PCFD_SCHEME_RK3) {
    set_last_error("the differentiable step is SSP-RK3 only");
    return kErrUnsupported;
  }
  for (int w = 0; w < 4; ++w)
    if (V.kind[w] == BC_OUTLET || V.kind[w] == BC_PERIODIC) {
      set_last_error("the differentiable step supports inlet, pressure-outlet, wall and "
                     "symmetry boundaries (no OUTLET mass balance, no periodic axis)");
      return kErrUnsupported;
    }
  if (V.g->hx || V.mom->nuc) {
    set_last_error("the differentiable step needs a uniform mesh and constant viscosity");
    return kErrUnsupported;
  }
  return kOk;
}

}  // namespace

extern "C" {

int pcfd_soft_face_masks(pcfd_ctx* c, const double* m, double* mf_u, double* mf_v) {
  SoftView V;
  if (int rc = check_soft(c, V)) return rc;
  hipLaunchKernelGGL(k_soft_faces, dim3((unsigned)num_tiles(V.nx + 3, V.ny + 3)), dim3(kBlock), 0,
                     V.stream, grid_of(V), m, mf_u, mf_v);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int pcfd_soft_tape_size(pcfd_ctx* c, int64_t* doubles) {
  SoftView V;
  if (int rc = check_soft(c, V)) return rc;
  if (doubles) *doubles = (int64_t)tape_doubles(V);
  return kOk;
}

int pcfd_soft_step(pcfd_ctx* c, const double* u0, const double* v0, double dt,
                   const double* mf_u, const double* mf_v, double* u, double* v, double* p,
                   double* tape, double* force_dev) {
  SoftView V;
  if (int rc = check_soft(c, V)) return rc;
  if (!(dt > 0.0)) {
    set_last_error("pcfd_soft_step: dt must be > 0");
    return kErrBadArg;
  }
  double* A[2] = {V.scratch_a[0], V.scratch_a[1]};
  double* B[2] = {V.scratch_b[0], V.scratch_b[1]};
  int rc;
  if ((rc = soft_substep(V, u0, v0, nullptr, nullptr, 0, 0, dt, mf_u, mf_v, A[0], A[1], p, tape, 0)))
    return rc;
  if ((rc = soft_substep(V, A[0], A[1], nullptr, nullptr, 0, 0, dt, mf_u, mf_v, B[0], B[1], p, tape,
                         1)))
    return rc;
  if ((rc = soft_substep(V, u0, v0, B[0], B[1], 0.75, 0.25, dt, mf_u, mf_v, A[0], A[1], p, tape, 2)))
    return rc;
  if ((rc = launch_combine(V.stream, u0, A[0], v0, A[1], 1.0 / 3.0, 2.0 / 3.0, V.nx, V.ny, V.ldu,
                           V.ldv, u, v)))
    return rc;
  if ((rc = launch_boundary_pass(V.stream, *V.bc, ops1(BOP_VEL_CORR), StepScalars{dt, 0.0, nullptr},
                                 u, v)))
    return rc;
  if (force_dev)
    PCFD_HIP_CHECK(hipMemcpyAsync(force_dev, V.force, 2 * sizeof(double), hipMemcpyDeviceToDevice,
                                  V.stream));
  return kOk;
}

int pcfd_soft_step_adjoint(pcfd_ctx* c, const double* tape, double dt, const double* mf_u,
                           const double* mf_v, double dL_dFx, double* lam_u, double* lam_v,
                           double* g_mf_u, double* g_mf_v) {
  SoftView V;
  if (int rc = check_soft(c, V)) return rc;
  if (!tape || !lam_u || !lam_v || !g_mf_u || !g_mf_v) {
    set_last_error("pcfd_soft_step_adjoint: null argument");
    return kErrBadArg;
  }
  int rc;
  const size_t nu = usize(V), nv = vsize(V);
  // lam_c = BC_corr^T lam_out; lam_u0 = lam_c / 3; lam_a3 = 2 lam_c / 3
  if ((rc = launch_walls_adjoint(V.stream, *V.bc, false, lam_u, lam_v))) return rc;
  double* l0u = V.scratch_a[0];  // accumulates the input's adjoint
  double* l0v = V.scratch_a[1];
  if ((rc = axpby(V, lam_u, nullptr, 1.0 / 3.0, 0.0, l0u, nu))) return rc;
  if ((rc = axpby(V, lam_v, nullptr, 1.0 / 3.0, 0.0, l0v, nv))) return rc;
  if ((rc = axpby(V, lam_u, nullptr, 2.0 / 3.0, 0.0, lam_u, nu))) return rc;
  if ((rc = axpby(V, lam_v, nullptr, 2.0 / 3.0, 0.0, lam_v, nv))) return rc;
  // third substep (its force is this step's drag): lam_bb
  if ((rc = soft_substep_adjoint(V, tape, 2, dt, mf_u, mf_v, dL_dFx, lam_u, lam_v, g_mf_u, g_mf_v)))
    return rc;
  // bb = 0.75 u0 + 0.25 a2
  if ((rc = axpby(V, l0u, lam_u, 1.0, 0.75, l0u, nu))) return rc;
  if ((rc = axpby(V, l0v, lam_v, 1.0, 0.75, l0v, nv))) return rc;
  if ((rc = axpby(V, lam_u, nullptr, 0.25, 0.0, lam_u, nu))) return rc;
  if ((rc = axpby(V, lam_v, nullptr, 0.25, 0.0, lam_v, nv))) return rc;
  if ((rc = soft_substep_adjoint(V, tape, 1, dt, mf_u, mf_v, 0.0, lam_u, lam_v, g_mf_u, g_mf_v)))
    return rc;
  if ((rc = soft_substep_adjoint(V, tape, 0, dt, mf_u, mf_v, 0.0, lam_u, lam_v, g_mf_u, g_mf_v)))
    return rc;
  if ((rc = axpby(V, l0u, lam_u, 1.0, 1.0, lam_u, nu))) return rc;
  return axpby(V, l0v, lam_v, 1.0, 1.0, lam_v, nv);
}

int pcfd_soft_cell_gradient(pcfd_ctx* c, const double* m, const double* g_mf_u,
                            const double* g_mf_v, double* g_theta) {
  SoftView V;
  if (int rc = check_soft(c, V)) return rc;
  hipLaunchKernelGGL(k_soft_cells_adj, dim3((unsigned)num_tiles(V.nx, V.ny)), dim3(kBlock), 0,
                     V.stream, grid_of(V), m, g_mf_u, g_mf_v, g_theta);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

}  // extern "C"
