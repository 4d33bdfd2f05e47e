// Fused multigrid level kernels (declarations and the tiling scheme in
// vcycle.hpp).  The arithmetic of every pass follows the reference V-cycle
// (pycfd/core/multigrid.py:424-451): damped Jacobi x += w D^-1 (b - A x),
// residual b - A x, restriction P^T r = P_tent^T (r - w A D^-1 r) and
// prolongation P e = e - w D^-1 A e with e = P_tent x_c, P being the
// Jacobi-smoothed tentative prolongator of multigrid.py:248-273.
#include "vcycle.hpp"

#include <algorithm>

#include "poisson_dev.hpp"

namespace pcfd {

namespace {

constexpr uint8_t F_DIRB = 0x20;  // region flag: solvable cell on a Dirichlet wall
constexpr uint16_t kOutside = 0xFFFF;

// Periodic image of i in [0, n).  Halos span a few cells, so the loops run
// once at most except on grids smaller than the halo.
__device__ __forceinline__ int wrap_index(int i, int n) {
  while (i < 0) i += n;
  while (i >= n) i -= n;
  return i;
}

// Global cell of a region cell; false when it lies outside a non-periodic grid.
__device__ __forceinline__ bool resolve(int gi, int gj, int nx, int ny, int px, int py, int& i,
                                        int& j) {
  i = gi;
  j = gj;
  if (px) i = wrap_index(gi, nx);
  if (py) j = wrap_index(gj, ny);
  return i >= 0 && i < nx && j >= 0 && j < ny;
}

struct TileOrigin {
  int t;         // linear tile index (partials slot)
  int ti, tj;    // tile coordinates
  int gi0, gj0;  // global (unwrapped) coordinates of region cell (0, 0)
};

__device__ __forceinline__ TileOrigin tile_origin(const FusedTile& ft) {
  TileOrigin o;
  o.t = xcd_remap(blockIdx.x, gridDim.x);
  o.ti = o.t / ft.nty;
  o.tj = o.t - o.ti * ft.nty;
  o.gi0 = o.ti * ft.tx - ft.hx;
  o.gj0 = o.tj * ft.ty - ft.hy;
  return o;
}

struct Ring {
  int a0, a1, c0, c1;
};

// Ring v of the region: the tile grown by v stencil radii.
__device__ __forceinline__ Ring ring(const FusedTile& ft, int v, int rxs, int rys) {
  return {ft.hx - v * rxs, ft.hx + ft.tx + v * rxs, ft.hy - v * rys, ft.hy + ft.ty + v * rys};
}

// Visits the cells of a ring, flattened over the block's 256 threads.
template <class Fn>
__device__ __forceinline__ void for_ring(const Ring& g, Fn&& fn) {
  const int w = g.c1 - g.c0;
  const int n = (g.a1 - g.a0) * w;
  const float inv = 1.0f / (float)w;
  for (int q = threadIdx.x; q < n; q += kBlock) {
    int a = (int)(((float)q + 0.5f) * inv);
    int c = q - a * w;
    if (c < 0) {
      --a;
      c += w;
    } else if (c >= w) {
      ++a;
      c -= w;
    }
    fn(g.a0 + a, g.c0 + c);
  }
}

// ---------------------------------------------------------------------------
// Level 0: matrix-free 5-point operator from the flag bytes.
// ---------------------------------------------------------------------------

// Region cell code: the flag byte's low five bits (solvable, four flux faces)
// plus bit 5 for a solvable cell on a Dirichlet x wall (i = 0 or nx - 1) and
// bit 6 for one on a Dirichlet y wall.  The diagonal and its inverse are a
// 128-entry table of the code, in pycfd/core/pressure.py:219-238's
// accumulation order (x faces and wall, then y faces and wall).
constexpr uint8_t C_XWALL = 0x20, C_YWALL = 0x40;

template <typename T>
__device__ __forceinline__ void fine_tables(const FineOp& A, T* td, T* ti) {
  const int f = threadIdx.x;
  if (f < 128) {
    double d = 1.0;
    if (f & F_SOLV) {
      d = 0.0;
      if (f & F_L) d -= A.cx;
      if (f & C_XWALL) d -= 2.0 * A.cx;
      if (f & F_R) d -= A.cx;
      if (f & F_B) d -= A.cy;
      if (f & C_YWALL) d -= 2.0 * A.cy;
      if (f & F_T) d -= A.cy;
    }
    td[f] = (T)d;
    ti[f] = (T)(1.0 / d);
  }
}

__device__ __forceinline__ uint8_t cell_code(const FineOp& A, uint8_t f, int i, int j) {
  uint8_t code = f & 31;
  if (f & F_SOLV) {
    if ((A.dir[0] && i == 0) || (A.dir[1] && i == A.nx - 1)) code |= C_XWALL;
    if ((A.dir[2] && j == 0) || (A.dir[3] && j == A.ny - 1)) code |= C_YWALL;
  }
  return code;
}

// Row coefficients of a region cell: its four face coefficients, diagonal
// and diagonal inverse.  Uniform mesh (S = false): the constants and the
// code-indexed tables.  Stretched mesh (S = true): formed from the metric
// arrays at the cell's grid position (gi, gj unwrapped; a stretched axis is
// never periodic and a periodic one is uniform, so wrapping only picks an
// equal value), in the accumulation order of fine_diag.  A cell outside the
// grid or not solvable has code 0: an identity row with no faces.
template <typename T>
struct RCoef {
  T l, r, b, t, d, di;
};

template <typename T, bool S>
__device__ __forceinline__ RCoef<T> rcoef(const FineOp& A, uint8_t f, int gi, int gj, const T* td,
                                          const T* ti) {
  if constexpr (!S) {
    return {(T)A.cx, (T)A.cx, (T)A.cy, (T)A.cy, td[f], ti[f]};
  } else {
    if (!(f & F_SOLV)) return {(T)0, (T)0, (T)0, (T)0, (T)1, (T)1};
    const int i = A.px ? wrap_index(gi, A.nx) : gi;
    const int j = A.py ? wrap_index(gj, A.ny) : gj;
    const FaceC c = fine_faces(A, i, j);
    const double d = fine_diag(A, c, f, i, j);
    return {(T)c.l, (T)c.r, (T)c.b, (T)c.t, (T)d, (T)(1.0 / d)};
  }
}

// (A x) at region cell idx, terms in CSR column order.
template <typename T>
__device__ __forceinline__ T rapply(const RCoef<T>& c, uint8_t f, const T* X, int idx, int ry) {
  T y = (T)0;
  if (f & F_L) y += c.l * X[idx - ry];
  if (f & F_B) y += c.b * X[idx - 1];
  y += c.d * X[idx];
  if (f & F_T) y += c.t * X[idx + 1];
  if (f & F_R) y += c.r * X[idx + ry];
  return y;
}

// LDS of a level-0 tile: the diagonal tables, three vectors and the cell codes.
template <typename T>
struct FineLds {
  T *td, *ti, *P, *Q, *S;
  uint8_t* F;
  __device__ FineLds(unsigned char* smem, int rn) {
    td = (T*)smem;
    ti = td + 128;
    P = ti + 128;
    Q = P + rn;
    S = Q + rn;
    F = (uint8_t*)(S + rn);
  }
};

template <typename T, bool S>
__global__ void __launch_bounds__(kBlock)
k_fine_down(FineOp A, FusedTile ft, T w, int nu, int nonzero, const double* __restrict__ r,
            const T* __restrict__ b_in, const T* __restrict__ x_in, int shift_in,
            T* __restrict__ b_out, T* __restrict__ x_out, T* __restrict__ bc, int fx, int fy,
            int cnx, int cny, const SolveState* __restrict__ st) {
  if (st->done) return;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ry = ft.ry;
  FineLds<T> M(smem, ft.rx * ry);
  T* B = M.P;
  T* X = M.Q;
  T* Y = M.S;
  const T* td = M.td;
  const T* ti = M.ti;
  const uint8_t* F = M.F;
  const TileOrigin o = tile_origin(ft);
  fine_tables<T>(A, M.td, M.ti);
  const double mr = r ? st->mean_r : 0.0;
  const T m = shift_in ? (T)st->mean_x : (T)0;
  __syncthreads();

  // Load; a zero start also forms x = 0 + w D^-1 (b - A 0) here.
  for_ring(Ring{0, ft.rx, 0, ry}, [&](int a, int c) {
    int i, j;
    uint8_t f = 0;
    T bv = (T)0, xv = (T)0;
    if (resolve(o.gi0 + a, o.gj0 + c, A.nx, A.ny, A.px, A.py, i, j)) {
      const int k = i * A.ny + j;
      f = cell_code(A, A.flags[k], i, j);
      if (r) bv = (f & F_SOLV) ? (T)(r[k] - mr) : (T)0;
      else bv = b_in[k];
      if (nonzero) {
        const T x0 = x_in[k];
        xv = (f & F_SOLV) ? x0 - m : x0;
      } else {
        xv = w * rcoef<T, S>(A, f, i, j, td, ti).di * bv;
      }
    }
    const int idx = a * ry + c;
    M.F[idx] = f;
    B[idx] = bv;
    X[idx] = xv;
  });
  __syncthreads();

  // Pre-smoothing; v is the number of stencil rings still valid.
  int v = ft.hx;
  for (int s = nonzero ? 0 : 1; s < nu; ++s) {
    --v;
    for_ring(ring(ft, v, 1, 1), [&](int a, int c) {
      const int idx = a * ry + c;
      const uint8_t f = F[idx];
      const RCoef<T> cf = rcoef<T, S>(A, f, o.gi0 + a, o.gj0 + c, td, ti);
      Y[idx] = X[idx] + (w * cf.di) * (B[idx] - rapply(cf, f, X, idx, ry));
    });
    __syncthreads();
    T* t = X;
    X = Y;
    Y = t;
  }

  // Residual on ring 1 into Y (v >= 2 rings of the iterate are valid here).
  for_ring(ring(ft, 1, 1, 1), [&](int a, int c) {
    const int idx = a * ry + c;
    const uint8_t f = F[idx];
    Y[idx] = B[idx] - rapply(rcoef<T, S>(A, f, o.gi0 + a, o.gj0 + c, td, ti), f, X, idx, ry);
  });
  __syncthreads();
  // Hand the tile's iterate (and right-hand side) over, then X <- D^-1 r.
  for_ring(ring(ft, 1, 1, 1), [&](int a, int c) {
    const int idx = a * ry + c;
    const int ia = a - ft.hx, jc = c - ft.hy;
    if (ia >= 0 && ia < ft.tx && jc >= 0 && jc < ft.ty) {
      const int i = o.ti * ft.tx + ia, j = o.tj * ft.ty + jc;
      if (i < A.nx && j < A.ny) {
        const int k = i * A.ny + j;
        x_out[k] = X[idx];
        if (b_out) b_out[k] = B[idx];
      }
    }
    X[idx] = rcoef<T, S>(A, F[idx], o.gi0 + a, o.gj0 + c, td, ti).di * Y[idx];
  });
  __syncthreads();
  // Restriction: one thread per coarse cell of the tile (cty is a power of two).
  const int ctx = ft.tx / fx, cty = ft.ty / fy;
  const int sh = __ffs(cty) - 1;
  for (int q = threadIdx.x; q < ctx * cty; q += kBlock) {
    const int ca = q >> sh, cc = q & (cty - 1);
    const int I = o.ti * ctx + ca, J = o.tj * cty + cc;
    if (I >= cnx || J >= cny) continue;
    T acc = (T)0;
    for (int a2 = 0; a2 < fx; ++a2) {
      const int i = I * fx + a2;
      if (i >= A.nx) break;
      for (int c2 = 0; c2 < fy; ++c2) {
        const int j = J * fy + c2;
        if (j >= A.ny) break;
        const int idx = (ft.hx + ca * fx + a2) * ry + (ft.hy + cc * fy + c2);
        const uint8_t f = F[idx];
        if (!(f & F_SOLV)) continue;
        acc += Y[idx] - w * rapply(rcoef<T, S>(A, f, i, j, td, ti), f, X, idx, ry);
      }
    }
    bc[I * cny + J] = acc;
  }
}

template <typename T, bool S>
__global__ void __launch_bounds__(kBlock)
k_fine_up(FineOp A, FusedTile ft, T w, int nu, const T* __restrict__ x_pre,
          const T* __restrict__ b0, const T* __restrict__ xc, int fx, int fy, int cny,
          T* __restrict__ z, int reduce, double* partial, const SolveState* __restrict__ st) {
  if (st->done) return;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ry = ft.ry;
  FineLds<T> M(smem, ft.rx * ry);
  T* E = M.P;
  T* X = M.Q;
  T* B = M.S;
  const T* td = M.td;
  const T* ti = M.ti;
  const uint8_t* F = M.F;
  const TileOrigin o = tile_origin(ft);
  fine_tables<T>(A, M.td, M.ti);
  __syncthreads();

  for_ring(Ring{0, ft.rx, 0, ry}, [&](int a, int c) {
    int i, j;
    uint8_t f = 0;
    T xv = (T)0, bv = (T)0, ev = (T)0;
    if (resolve(o.gi0 + a, o.gj0 + c, A.nx, A.ny, A.px, A.py, i, j)) {
      const int k = i * A.ny + j;
      f = cell_code(A, A.flags[k], i, j);
      xv = x_pre[k];
      bv = b0[k];
      if (f & F_SOLV) ev = xc[(i / fx) * cny + (j / fy)];
    }
    const int idx = a * ry + c;
    M.F[idx] = f;
    X[idx] = xv;
    B[idx] = bv;
    E[idx] = ev;
  });
  __syncthreads();
  // x += e - w D^-1 A e on ring nu.
  for_ring(ring(ft, nu, 1, 1), [&](int a, int c) {
    const int idx = a * ry + c;
    const uint8_t f = F[idx];
    const RCoef<T> cf = rcoef<T, S>(A, f, o.gi0 + a, o.gj0 + c, td, ti);
    const T ae = rapply(cf, f, E, idx, ry);
    X[idx] += E[idx] - w * cf.di * ae;
  });
  __syncthreads();
  // Post-smoothing; the last sweep writes the tile straight to z.
  double s[3] = {0.0, 0.0, 0.0};
  T* Y = E;
  for (int sw = 1; sw <= nu; ++sw) {
    const bool last = sw == nu;
    for_ring(ring(ft, nu - sw, 1, 1), [&](int a, int c) {
      const int idx = a * ry + c;
      const uint8_t f = F[idx];
      const RCoef<T> cf = rcoef<T, S>(A, f, o.gi0 + a, o.gj0 + c, td, ti);
      const T val = X[idx] + (w * cf.di) * (B[idx] - rapply(cf, f, X, idx, ry));
      if (!last) {
        Y[idx] = val;
        return;
      }
      const int gi = o.gi0 + a, gj = o.gj0 + c;
      if (gi >= A.nx || gj >= A.ny) return;  // tile cells past the grid's end
      z[gi * A.ny + gj] = val;
      if (reduce != SR_NONE && (f & F_SOLV)) {
        const double bd = (double)B[idx];
        s[0] += (double)val;
        s[1] += bd * (double)val;
        s[2] += bd;
      }
    });
    if (!last) {
      __syncthreads();
      T* t = X;
      X = Y;
      Y = t;
    }
  }
  if (reduce == SR_NONE) return;
  const int ops[3] = {kSum, kSum, kSum};
  block_partials<3>(s, ops, partial, o.t);
}

// Multigrid-CG direction on level 0 (pressure.py:577-632 through scipy's cg):
// p = zp + beta p_old with zp = z - mean_z on solvable cells (0 elsewhere),
// q = A p, and the tile's partial of p.q.  p is formed once per region cell
// in LDS (halo 1), then q on the tile.
template <typename T, bool S>
__global__ void __launch_bounds__(kBlock)
k_fine_dir(FineOp A, FusedTile ft, const T* __restrict__ z, const double* __restrict__ p_old,
           double* __restrict__ p_new, double* __restrict__ q, double* partial,
           const SolveState* __restrict__ st) {
  if (st->done) return;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ry = ft.ry, rn = ft.rx * ry;
  double* td = (double*)smem;
  double* ti = td + 128;
  double* P = ti + 128;
  uint8_t* F = (uint8_t*)(P + rn);
  const TileOrigin o = tile_origin(ft);
  fine_tables<double>(A, td, ti);
  const bool first = st->iter == 0;
  const double beta = st->beta;
  const double mz = st->mean_z;
  for_ring(Ring{0, ft.rx, 0, ry}, [&](int a, int c) {
    int i, j;
    uint8_t f = 0;
    double pv = 0.0;
    if (resolve(o.gi0 + a, o.gj0 + c, A.nx, A.ny, A.px, A.py, i, j)) {
      const int k = i * A.ny + j;
      f = cell_code(A, A.flags[k], i, j);
      const double zp = (f & F_SOLV) ? (double)z[k] - mz : 0.0;
      pv = first ? zp : p_old[k] * beta + zp;
    }
    const int idx = a * ry + c;
    F[idx] = f;
    P[idx] = pv;
  });
  __syncthreads();
  double s[1] = {0.0};
  for_ring(ring(ft, 0, 1, 1), [&](int a, int c) {
    const int gi = o.gi0 + a, gj = o.gj0 + c;
    if (gi >= A.nx || gj >= A.ny) return;
    const int idx = a * ry + c;
    const uint8_t f = F[idx];
    const double pk = P[idx];
    const double qk = rapply(rcoef<double, S>(A, f, gi, gj, td, ti), f, P, idx, ry);
    const int k = gi * A.ny + gj;
    p_new[k] = pk;
    q[k] = qk;
    s[0] += pk * qk;
  });
  const int ops[1] = {kSum};
  block_partials<1>(s, ops, partial, o.t);
}

// ---------------------------------------------------------------------------
// Levels >= 1: stored Galerkin rows (class dictionary).
// ---------------------------------------------------------------------------
// Row c of the level applied at region cell idx; four entries per trip with
// their loads issued together, summed in entry order.
template <typename T>
__device__ __forceinline__ T class_row(const ClassRows<T>& L, int c, const T* X, int idx) {
  T acc = (T)0;
  const int e1 = L.row_ptr[c + 1];
  int e = L.row_ptr[c];
  for (; e + 3 < e1; e += 4) {
    const int o0 = L.ent_off[e], o1 = L.ent_off[e + 1], o2 = L.ent_off[e + 2],
              o3 = L.ent_off[e + 3];
    const T a0 = L.ent_val[e], a1 = L.ent_val[e + 1], a2 = L.ent_val[e + 2],
            a3 = L.ent_val[e + 3];
    const T x0 = X[idx + o0], x1 = X[idx + o1], x2 = X[idx + o2], x3 = X[idx + o3];
    acc += a0 * x0;
    acc += a1 * x1;
    acc += a2 * x2;
    acc += a3 * x3;
  }
  for (; e < e1; ++e) acc += L.ent_val[e] * X[idx + L.ent_off[e]];
  return acc;
}

// LDS of a stored-level tile: the hot row's coefficients and offsets, three
// vectors and the row classes.
template <typename T>
struct ClassLds {
  T *HV, *P, *Q, *S;
  int* HO;
  uint16_t* C;
  __device__ ClassLds(unsigned char* smem, int rn, int hot_n) {
    HV = (T*)smem;
    P = HV + hot_n;
    Q = P + rn;
    S = Q + rn;
    HO = (int*)(S + rn);
    C = (uint16_t*)(HO + hot_n);
  }
};

// Row c at region cell idx, the hot row from LDS (the common case: all lanes
// of an interior wavefront share it), any other row from the global dictionary.
template <typename T>
__device__ __forceinline__ T class_row_lds(const ClassRows<T>& L, const ClassLds<T>& M, int c,
                                           const T* X, int idx) {
  if (c != L.hot) return class_row(L, c, X, idx);
  T acc = (T)0;
  const int n = L.hot_n;
  int e = 0;
  for (; e + 3 < n; e += 4) {
    const T x0 = X[idx + M.HO[e]], x1 = X[idx + M.HO[e + 1]], x2 = X[idx + M.HO[e + 2]],
            x3 = X[idx + M.HO[e + 3]];
    acc += M.HV[e] * x0;
    acc += M.HV[e + 1] * x1;
    acc += M.HV[e + 2] * x2;
    acc += M.HV[e + 3] * x3;
  }
  for (; e < n; ++e) acc += M.HV[e] * X[idx + M.HO[e]];
  return acc;
}

template <typename T>
__device__ __forceinline__ void stage_hot_row(const ClassRows<T>& L, const ClassLds<T>& M) {
  if (L.hot < 0) return;
  const int e0 = L.row_ptr[L.hot];
  for (int e = threadIdx.x; e < L.hot_n; e += kBlock) {
    M.HV[e] = L.ent_val[e0 + e];
    M.HO[e] = L.ent_off[e0 + e];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_class_down(ClassRows<T> L, FusedTile ft, int rxs, int rys, T w, int nu, const T* __restrict__ b,
             T* __restrict__ x_out, T* __restrict__ bc, int fx, int fy, int cnx, int cny,
             const SolveState* __restrict__ st) {
  if (st->done) return;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ry = ft.ry;
  ClassLds<T> M(smem, ft.rx * ry, L.hot < 0 ? 0 : L.hot_n);
  T* B = M.P;
  T* X = M.Q;
  T* Y = M.S;
  uint16_t* C = M.C;
  const TileOrigin o = tile_origin(ft);
  stage_hot_row(L, M);

  // Load b and the row classes, and start from zero: x = w D^-1 b.
  for_ring(Ring{0, ft.rx, 0, ry}, [&](int a, int c) {
    int i, j;
    uint16_t cl = kOutside;
    T bv = (T)0, xv = (T)0;
    if (resolve(o.gi0 + a, o.gj0 + c, L.nx, L.ny, L.px, L.py, i, j)) {
      const int k = i * L.ny + j;
      cl = L.cls[k];
      bv = b[k];
      xv = w * L.cdinv[cl] * bv;
    }
    const int idx = a * ry + c;
    C[idx] = cl;
    B[idx] = bv;
    X[idx] = xv;
  });
  __syncthreads();
  for (int s = 1; s < nu; ++s) {
    for_ring(ring(ft, nu + 1 - s, rxs, rys), [&](int a, int c) {
      const int idx = a * ry + c;
      const uint16_t cl = C[idx];
      if (cl == kOutside) {
        Y[idx] = (T)0;
        return;
      }
      Y[idx] = X[idx] + w * L.cdinv[cl] * (B[idx] - class_row_lds(L, M, cl, X, idx));
    });
    __syncthreads();
    T* t = X;
    X = Y;
    Y = t;
  }
  // Residual on ring 1 into Y.
  for_ring(ring(ft, 1, rxs, rys), [&](int a, int c) {
    const int idx = a * ry + c;
    const uint16_t cl = C[idx];
    Y[idx] = (cl == kOutside) ? (T)0 : B[idx] - class_row_lds(L, M, cl, X, idx);
  });
  __syncthreads();
  for_ring(ring(ft, 1, rxs, rys), [&](int a, int c) {
    const int idx = a * ry + c;
    const int ia = a - ft.hx, jc = c - ft.hy;
    if (ia >= 0 && ia < ft.tx && jc >= 0 && jc < ft.ty) {
      const int i = o.ti * ft.tx + ia, j = o.tj * ft.ty + jc;
      if (i < L.nx && j < L.ny) x_out[i * L.ny + j] = X[idx];
    }
    const uint16_t cl = C[idx];
    X[idx] = (cl == kOutside) ? (T)0 : L.cdinv[cl] * Y[idx];
  });
  __syncthreads();
  const int ctx = ft.tx / fx, cty = ft.ty / fy;
  for (int q = threadIdx.x; q < ctx * cty; q += kBlock) {
    const int ca = q / cty, cc = q - ca * cty;
    const int I = o.ti * ctx + ca, J = o.tj * cty + cc;
    if (I >= cnx || J >= cny) continue;
    T acc = (T)0;
    for (int a2 = 0; a2 < fx; ++a2) {
      const int i = I * fx + a2;
      if (i >= L.nx) break;
      for (int c2 = 0; c2 < fy; ++c2) {
        const int j = J * fy + c2;
        if (j >= L.ny) break;
        const int idx = (ft.hx + ca * fx + a2) * ry + (ft.hy + cc * fy + c2);
        const uint16_t cl = C[idx];
        acc += L.cactive[cl] ? Y[idx] - w * class_row_lds(L, M, cl, X, idx) : (T)0;
      }
    }
    bc[I * cny + J] = acc;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_class_up(ClassRows<T> L, FusedTile ft, int rxs, int rys, T w, int nu, const T* __restrict__ x_pre,
           const T* __restrict__ b, const T* __restrict__ xc, int fx, int fy, int cny,
           T* __restrict__ x_out, const SolveState* __restrict__ st) {
  if (st->done) return;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ry = ft.ry;
  ClassLds<T> M(smem, ft.rx * ry, L.hot < 0 ? 0 : L.hot_n);
  T* E = M.P;
  T* X = M.Q;
  T* B = M.S;
  uint16_t* C = M.C;
  const TileOrigin o = tile_origin(ft);
  stage_hot_row(L, M);

  for_ring(Ring{0, ft.rx, 0, ry}, [&](int a, int c) {
    int i, j;
    uint16_t cl = kOutside;
    T xv = (T)0, bv = (T)0, ev = (T)0;
    if (resolve(o.gi0 + a, o.gj0 + c, L.nx, L.ny, L.px, L.py, i, j)) {
      const int k = i * L.ny + j;
      cl = L.cls[k];
      xv = x_pre[k];
      bv = b[k];
      if (L.cactive[cl]) ev = xc[(i / fx) * cny + (j / fy)];
    }
    const int idx = a * ry + c;
    C[idx] = cl;
    X[idx] = xv;
    B[idx] = bv;
    E[idx] = ev;
  });
  __syncthreads();
  for_ring(ring(ft, nu, rxs, rys), [&](int a, int c) {
    const int idx = a * ry + c;
    const uint16_t cl = C[idx];
    if (cl == kOutside) return;
    X[idx] = X[idx] + E[idx] - w * L.cdinv[cl] * class_row_lds(L, M, cl, E, idx);
  });
  __syncthreads();
  T* Y = E;
  for (int sw = 1; sw <= nu; ++sw) {
    const bool last = sw == nu;
    for_ring(ring(ft, nu - sw, rxs, rys), [&](int a, int c) {
      const int idx = a * ry + c;
      const uint16_t cl = C[idx];
      T val = (T)0;
      if (cl != kOutside) val = X[idx] + w * L.cdinv[cl] * (B[idx] - class_row_lds(L, M, cl, X, idx));
      if (!last) {
        Y[idx] = val;
        return;
      }
      const int i = o.gi0 + a, j = o.gj0 + c;
      if (i < L.nx && j < L.ny) x_out[i * L.ny + j] = val;
    });
    if (!last) {
      __syncthreads();
      T* t = X;
      X = Y;
      Y = t;
    }
  }
}

// ---------------------------------------------------------------------------
// Dense coarsest-level operator, one wavefront per row.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_dense_apply(const double* __restrict__ C, int n, const T* __restrict__ b, T* __restrict__ x,
              const SolveState* __restrict__ st) {
  if (st->done) return;
  extern __shared__ double sb[];
  for (int k = threadIdx.x; k < n; k += kBlock) sb[k] = (double)b[k];
  __syncthreads();
  const int lane = threadIdx.x % kWave;
  const int wv = threadIdx.x / kWave;
  constexpr int kWaves = kBlock / kWave;
  for (int row = blockIdx.x * kWaves + wv; row < n; row += gridDim.x * kWaves) {
    const double* c = C + (size_t)row * n;
    double acc = 0.0;
#pragma unroll 8
    for (int k = lane; k < n; k += kWave) acc += c[k] * sb[k];
    acc = wave_sum(acc);
    if (lane == 0) x[row] = (T)acc;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
template <typename T>
int launch_fine_down(hipStream_t s, const FineOp& A, const FusedTile& ft, T w, int nu,
                     int nonzero, const double* r, const T* b_in, const T* x_in, int shift_in,
                     T* b_out, T* x_out, T* bc, int fx, int fy, int cnx, int cny,
                     const SolveState* st, const LaunchClock& clk) {
  if (A.hxc)
    PCFD_LAUNCH_TIMED(clk, (k_fine_down<T, true>), dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, A,
                      ft, w, nu, nonzero, r, b_in, x_in, shift_in, b_out, x_out, bc, fx, fy, cnx,
                      cny, st);
  else
    PCFD_LAUNCH_TIMED(clk, (k_fine_down<T, false>), dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, A,
                      ft, w, nu, nonzero, r, b_in, x_in, shift_in, b_out, x_out, bc, fx, fy, cnx,
                      cny, st);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int launch_fine_up(hipStream_t s, const FineOp& A, const FusedTile& ft, T w, int nu,
                   const T* x_pre, const T* b0, const T* xc, int fx, int fy, int cny, T* z,
                   int reduce, double* partial, const SolveState* st, const LaunchClock& clk) {
  if (A.hxc)
    PCFD_LAUNCH_TIMED(clk, (k_fine_up<T, true>), dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, A,
                      ft, w, nu, x_pre, b0, xc, fx, fy, cny, z, reduce, partial, st);
  else
    PCFD_LAUNCH_TIMED(clk, (k_fine_up<T, false>), dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, A,
                      ft, w, nu, x_pre, b0, xc, fx, fy, cny, z, reduce, partial, st);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int launch_class_down(hipStream_t s, const ClassRows<T>& L, const FusedTile& ft, int rxs,
                      int rys, T w, int nu, const T* b, T* x_out, T* bc, int fx, int fy, int cnx,
                      int cny, const SolveState* st) {
  hipLaunchKernelGGL(k_class_down<T>, dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, L, ft, rxs, rys,
                     w, nu, b, x_out, bc, fx, fy, cnx, cny, st);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int launch_class_up(hipStream_t s, const ClassRows<T>& L, const FusedTile& ft, int rxs, int rys,
                    T w, int nu, const T* x_pre, const T* b, const T* xc, int fx, int fy, int cny,
                    T* x_out, const SolveState* st) {
  hipLaunchKernelGGL(k_class_up<T>, dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, L, ft, rxs, rys, w,
                     nu, x_pre, b, xc, fx, fy, cny, x_out, st);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int launch_fine_dir(hipStream_t s, const FineOp& A, const FusedTile& ft, const T* z,
                    const double* p_old, double* p_new, double* q, double* partial,
                    const SolveState* st, const LaunchClock& clk) {
  if (A.hxc)
    PCFD_LAUNCH_TIMED(clk, (k_fine_dir<T, true>), dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, A,
                      ft, z, p_old, p_new, q, partial, st);
  else
    PCFD_LAUNCH_TIMED(clk, (k_fine_dir<T, false>), dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, A,
                      ft, z, p_old, p_new, q, partial, st);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int launch_dense_apply(hipStream_t s, const double* C, int n, const T* b, T* x,
                       const SolveState* st) {
  constexpr int kWaves = kBlock / kWave;
  const int blocks = std::max(1, std::min((n + kWaves - 1) / kWaves, 1024));
  hipLaunchKernelGGL(k_dense_apply<T>, dim3(blocks), dim3(kBlock), (size_t)n * sizeof(double), s,
                     C, n, b, x, st);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

std::vector<FusedTile> tile_candidates(int nx, int ny, int rxs, int rys, int halo_mult, int fx,
                                       int fy, unsigned bytes_per_cell, unsigned fixed_bytes,
                                       unsigned lds_cap, int force_tx, int force_ty) {
  static const int kTx[] = {64, 32, 16, 8, 4, 2};
  static const int kTy[] = {256, 128, 64, 32, 16, 8, 4, 2};
  std::vector<FusedTile> out;
  for (int tx : kTx) {
    if (force_tx > 0 && tx != force_tx) continue;
    if (tx % fx) continue;
    if (force_tx <= 0 && tx > 2 * nx && tx > fx) continue;
    for (int ty : kTy) {
      if (force_ty > 0 && ty != force_ty) continue;
      if (ty % fy) continue;
      if (force_ty <= 0 && ty > 2 * ny && ty > fy) continue;
      // tiny tiles only where the level itself is tiny
      if (force_tx <= 0 && force_ty <= 0 && tx * ty < 32 && (long)nx * ny >= 4096) continue;
      FusedTile t;
      t.tx = tx;
      t.ty = ty;
      t.hx = halo_mult * rxs;
      t.hy = halo_mult * rys;
      t.rx = tx + 2 * t.hx;
      t.ry = ty + 2 * t.hy;
      t.ntx = (nx + tx - 1) / tx;
      t.nty = (ny + ty - 1) / ty;
      const size_t lds = (size_t)fixed_bytes + (size_t)t.rx * t.ry * bytes_per_cell;
      if (lds > lds_cap) continue;
      t.lds = (unsigned)((lds + 15) / 16 * 16);
      out.push_back(t);
    }
  }
  return out;
}

double tile_cost(const FusedTile& t) {
  // Work-conserving estimate: region cells per CU (the halo is the redundant
  // part), plus a per-tile overhead; 256 CUs.
  const double tiles_per_cu = std::max(1.0, (double)t.ntiles() / 256.0);
  return tiles_per_cu * ((double)t.rx * t.ry + 256.0);
}

bool choose_tile(int nx, int ny, int rxs, int rys, int halo_mult, int fx, int fy,
                 unsigned bytes_per_cell, unsigned fixed_bytes, unsigned lds_cap, int force_tx,
                 int force_ty, FusedTile* out) {
  const std::vector<FusedTile> c = tile_candidates(nx, ny, rxs, rys, halo_mult, fx, fy,
                                                   bytes_per_cell, fixed_bytes, lds_cap,
                                                   force_tx, force_ty);
  if (c.empty()) return false;
  *out = *std::min_element(c.begin(), c.end(), [](const FusedTile& a, const FusedTile& b) {
    return tile_cost(a) < tile_cost(b);
  });
  return true;
}

// Explicit instantiations for the two cycle precisions.
#define PCFD_VCYCLE_INST(T)                                                                    \
  template int launch_fine_down<T>(hipStream_t, const FineOp&, const FusedTile&, T, int, int,   \
                                   const double*, const T*, const T*, int, T*, T*, T*, int,   \
                                   int, int, int, const SolveState*, const LaunchClock&);     \
  template int launch_fine_up<T>(hipStream_t, const FineOp&, const FusedTile&, T, int,         \
                                 const T*, const T*, const T*, int, int, int, T*, int, double*, \
                                 const SolveState*, const LaunchClock&);                      \
  template int launch_class_down<T>(hipStream_t, const ClassRows<T>&, const FusedTile&, int,   \
                                    int, T, int, const T*, T*, T*, int, int, int, int,        \
                                    const SolveState*);                                       \
  template int launch_class_up<T>(hipStream_t, const ClassRows<T>&, const FusedTile&, int, int, \
                                  T, int, const T*, const T*, const T*, int, int, int, T*,    \
                                  const SolveState*);                                         \
  template int launch_dense_apply<T>(hipStream_t, const double*, int, const T*, T*,            \
                                     const SolveState*);                                      \
  template int launch_fine_dir<T>(hipStream_t, const FineOp&, const FusedTile&, const T*,      \
                                  const double*, double*, double*, double*, const SolveState*, \
                                  const LaunchClock&);
PCFD_VCYCLE_INST(float)
PCFD_VCYCLE_INST(double)
#undef PCFD_VCYCLE_INST

}  // namespace pcfd
