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

// Visits every cell of a whole region (ring 0) in batches of KB cells per
// thread: `ld(a, c)` issues a cell's global loads and returns them, and all
// KB loads of a batch are in flight before `st(a, c, idx, v)` consumes them
// (one load round trip per batch instead of one per cell).
template <int KB, class Ld, class St>
__device__ __forceinline__ void for_region_batched(int rx, int ry, Ld&& ld, St&& st) {
  using V = decltype(ld(0, 0));
  const int n = rx * ry;
  const float inv = 1.0f / (float)ry;
  auto split = [&](int q, int& a, int& c) {
    a = (int)(((float)q + 0.5f) * inv);
    c = q - a * ry;
    if (c < 0) {
      --a;
      c += ry;
    } else if (c >= ry) {
      ++a;
      c -= ry;
    }
  };
  for (int base = threadIdx.x; base < n; base += KB * kBlock) {
    V vals[KB];
#pragma unroll
    for (int u = 0; u < KB; ++u) {
      const int q = base + u * kBlock;
      if (q < n) {
        int a, c;
        split(q, a, c);
        vals[u] = ld(a, c);
      }
    }
#pragma unroll
    for (int u = 0; u < KB; ++u) {
      const int q = base + u * kBlock;
      if (q < n) {
        int a, c;
        split(q, a, c);
        st(a, c, q, vals[u]);
      }
    }
  }
}

constexpr int kLoadBatch = 4;

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
  struct DownLd {
    int i, j;
    bool ok;
    uint8_t fl;
    double rv;
    T bi, xi;
  };
  for_region_batched<kLoadBatch>(
      ft.rx, ry,
      [&](int a, int c) {
        DownLd v;
        v.ok = resolve(o.gi0 + a, o.gj0 + c, A.nx, A.ny, A.px, A.py, v.i, v.j);
        v.fl = 0;
        v.rv = 0.0;
        v.bi = (T)0;
        v.xi = (T)0;
        if (v.ok) {
          const int k = v.i * A.ny + v.j;
          v.fl = A.flags[k];
          if (r) v.rv = r[k];
          else v.bi = b_in[k];
          if (nonzero) v.xi = x_in[k];
        }
        return v;
      },
      [&](int a, int c, int idx, const DownLd& v) {
        uint8_t f = 0;
        T bv = (T)0, xv = (T)0;
        if (v.ok) {
          f = cell_code(A, v.fl, v.i, v.j);
          if (r) bv = (f & F_SOLV) ? (T)(v.rv - mr) : (T)0;
          else bv = v.bi;
          if (nonzero) {
            xv = (f & F_SOLV) ? v.xi - m : v.xi;
          } else {
            xv = w * rcoef<T, S>(A, f, v.i, v.j, td, ti).di * bv;
          }
        }
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

  struct UpLd {
    int i, j;
    bool ok;
    uint8_t fl;
    T xv, bv, ev;
  };
  for_region_batched<kLoadBatch>(
      ft.rx, ry,
      [&](int a, int c) {
        UpLd v;
        v.ok = resolve(o.gi0 + a, o.gj0 + c, A.nx, A.ny, A.px, A.py, v.i, v.j);
        v.fl = 0;
        v.xv = v.bv = v.ev = (T)0;
        if (v.ok) {
          const int k = v.i * A.ny + v.j;
          v.fl = A.flags[k];
          v.xv = x_pre[k];
          v.bv = b0[k];
          // the coarse value is loaded for every grid cell (always in range)
          // and discarded below on unsolvable ones, so no load waits on flags
          v.ev = xc[(v.i / fx) * cny + (v.j / fy)];
        }
        return v;
      },
      [&](int a, int c, int idx, const UpLd& v) {
        uint8_t f = 0;
        if (v.ok) f = cell_code(A, v.fl, v.i, v.j);
        M.F[idx] = f;
        X[idx] = v.xv;
        B[idx] = v.bv;
        E[idx] = (f & F_SOLV) ? v.ev : (T)0;
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
//
// With cf.rows > 0 every workgroup first folds the V-cycle's sums itself, in
// k_fold's order (the same bits as the fold launch it replaces): mean_z,
// rho = r.z_proj and beta = rho / rho_old, with rho_old from the state's
// rho history (written one launch earlier, so no workgroup reads what this
// launch writes).  Workgroup 0 records mean_z, rho, rho_prev, beta and the
// history entry of this iteration.
template <typename T, bool S>
__global__ void __launch_bounds__(kBlock)
k_fine_dir(FineOp A, FusedTile ft, const T* __restrict__ z, const double* __restrict__ p_old,
           double* __restrict__ p_new, double* __restrict__ q, double* partial,
           SolveState* st, CycleFold cf) {
  if (st->done) return;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ry = ft.ry, rn = ft.rx * ry;
  double* td = (double*)smem;
  double* ti = td + 128;
  double* P = ti + 128;
  uint8_t* F = (uint8_t*)(P + rn);
  const TileOrigin o = tile_origin(ft);
  fine_tables<double>(A, td, ti);
  const int it = st->iter;
  const bool first = it == 0;
  double beta, mz;
  if (cf.rows > 0) {
    double v[3];
    const int op3[3] = {kSum, kSum, kSum};
    fold_partials<3>(v, op3, cf.partial, cf.rows);
    __shared__ double s_cyc[2];
    if (threadIdx.x == 0) {
      // EpiCycle's arithmetic (r.z_proj = sum(b0 z) - mean_z sum(b0))
      const double m = (cf.n_proj > 0) ? v[0] / (double)cf.n_proj : 0.0;
      const double rho = v[1] - m * v[2];
      double rho_old = 1.0;
      if (!first) rho_old = st->rho_hist[(it + 1) & 1];
      else if (blockIdx.x == 0) rho_old = st->rho;
      const double b = rho / rho_old;
      s_cyc[0] = m;
      s_cyc[1] = b;
      if (blockIdx.x == 0) {
        st->mean_z = m;
        st->rho_prev = rho_old;
        st->rho = rho;
        st->beta = b;
        st->rho_hist[it & 1] = rho;
      }
    }
    __syncthreads();
    mz = s_cyc[0];
    beta = s_cyc[1];
  } else {
    beta = st->beta;
    mz = st->mean_z;
  }
  struct DirLd {
    int i, j;
    bool ok;
    uint8_t fl;
    T zv;
    double po;
  };
  for_region_batched<kLoadBatch>(
      ft.rx, ry,
      [&](int a, int c) {
        DirLd v;
        v.ok = resolve(o.gi0 + a, o.gj0 + c, A.nx, A.ny, A.px, A.py, v.i, v.j);
        v.fl = 0;
        v.zv = (T)0;
        v.po = 0.0;
        if (v.ok) {
          const int k = v.i * A.ny + v.j;
          v.fl = A.flags[k];
          v.zv = z[k];
          if (!first) v.po = p_old[k];
        }
        return v;
      },
      [&](int a, int c, int idx, const DirLd& v) {
        uint8_t f = 0;
        double pv = 0.0;
        if (v.ok) {
          f = cell_code(A, v.fl, v.i, v.j);
          const double zp = (f & F_SOLV) ? (double)v.zv - mz : 0.0;
          pv = first ? zp : v.po * beta + zp;
        }
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
// Level 0, streaming form: uniform mesh, one sweep each way, zero start (the
// multigrid-CG preconditioner's cycle).  One wavefront owns a strip of 64
// columns (the middle 60 produce output; two halo lanes each side) and walks
// down L rows plus a four-row halo, keeping the last rows of every
// intermediate in registers: a row's x-neighbours are the rows before and
// after it in the walk, its y-neighbours the adjacent lanes (cross-lane
// shuffles).  No LDS tile and no barriers.  Cells that are not solvable
// hold zero at every stage, so a solvable row of the operator is
// d u + cx (u_l + u_r) + cy (u_b + u_t) without face tests: a face carries
// flux exactly when the neighbour is solvable (two fluid cells always share
// one), and every other neighbour contributes zero.  Global loads run
// kStreamPrefetch rows ahead of their use.
// ---------------------------------------------------------------------------
constexpr int kStreamHalo = 2;
constexpr int kStreamOut = kWave - 2 * kStreamHalo;
constexpr int kStreamPrefetch = 4;

struct StreamPos {
  int lane, i0, j, jr;
  bool jin, out_col, active;
};

__device__ __forceinline__ StreamPos stream_pos(const FineOp& A, int strips, int L) {
  const int wave = (int)(blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave);
  StreamPos p;
  p.lane = (int)(threadIdx.x % kWave);
  const int strip = wave % strips;
  p.i0 = (wave / strips) * L;
  p.active = p.i0 < A.nx;
  p.j = strip * kStreamOut - kStreamHalo + p.lane;
  p.jr = p.j;
  p.jin = true;
  if (A.py) p.jr = wrap_index(p.j, A.ny);
  else p.jin = p.j >= 0 && p.j < A.ny;
  p.out_col = p.lane >= kStreamHalo && p.lane < kStreamHalo + kStreamOut && p.j < A.ny;
  return p;
}

// Grid row of walk row p: wrapped on a periodic axis; false outside the grid.
__device__ __forceinline__ bool stream_row(const FineOp& A, int p, int& ir) {
  ir = p;
  if (A.px) {
    ir = wrap_index(p, A.nx);
    return true;
  }
  return p >= 0 && p < A.nx;
}

template <typename T>
__device__ __forceinline__ T lane_up(T v) {  // value of lane - 1
  return __shfl_up(v, 1, kWave);
}
template <typename T>
__device__ __forceinline__ T lane_down(T v) {  // value of lane + 1
  return __shfl_down(v, 1, kWave);
}

// Down leg: b0 = r - mean_r on solvable cells, x = w D^-1 b0 (the zero-start
// sweep), res = b0 - A x, t = D^-1 res, and the restriction
// b_c = P_tent^T (res - w A t) over FX x FY blocks; writes b0, x and b_c.
template <typename T, int FX, int FY>
__global__ void __launch_bounds__(kBlock)
k_fine_down_stream(FineOp A, int strips, int L, T w, const double* __restrict__ r,
                   T* __restrict__ b_out, T* __restrict__ x_out, T* __restrict__ bc, int cnx,
                   int cny, SolveState* __restrict__ st, UpdateFold uf) {
  if (st->done) return;
  __shared__ T td[128], ti[128];
  fine_tables<T>(A, td, ti);
  const StreamPos P = stream_pos(A, strips, L);
  const T cx = (T)A.cx, cy = (T)A.cy;
  const int p_begin = P.i0 - 2, p_end = P.i0 + L + 2;  // L % 4 == 0: a whole number of batches
  // prefetch slots
  uint8_t pf_f[kStreamPrefetch];
  double pf_r[kStreamPrefetch];
  int pf_i[kStreamPrefetch];
  auto issue = [&](int slot, int p) {
    int ir = 0;
    pf_f[slot] = 0;
    pf_r[slot] = 0.0;
    pf_i[slot] = 0;
    if (p < p_end && stream_row(A, p, ir) && P.jin) {
      const int k = ir * A.ny + P.jr;
      pf_f[slot] = A.flags[k];
      pf_r[slot] = r[k];
      pf_i[slot] = ir;
    }
  };
  if (P.active) {
#pragma unroll
    for (int u = 0; u < kStreamPrefetch; ++u) issue(u, p_begin + u);
  }
  // With uf.rows > 0 every workgroup folds the previous update's partials
  // itself, in k_fold's fixed order, while its first loads are in flight:
  // the update epilogue's arithmetic (iteration count, r.r, mean of r, the
  // stopping test), bit for bit, without the fold launch.  Workgroup 0
  // records it in the state.  The count comes in uf.iter; the one field read
  // above that this launch may write is `done`, and a workgroup that sees it
  // set here would have stopped after its own fold anyway.
  double mr;
  if (uf.rows > 0) {
    double v[2];
    const int op2[2] = {kSum, kSum};
    fold_partials<2>(v, op2, uf.partial, uf.rows);
    __shared__ double s_mr;
    __shared__ int s_done;
    if (threadIdx.x == 0) {
      const double m = (uf.n_proj > 0) ? v[1] / (double)uf.n_proj : 0.0;
      int done = 0, conv = 0;
      if (sqrt(v[0]) < st->atol) {
        done = 1;
        conv = 1;
      } else if (uf.iter >= uf.maxiter) {
        done = 1;
      }
      s_mr = m;
      s_done = done;
      if (blockIdx.x == 0) {
        st->iter = uf.iter;
        st->rr = v[0];
        st->mean_r = m;
        if (done) {
          st->done = 1;
          st->converged = conv;
        }
      }
    }
    __syncthreads();
    if (s_done) return;
    mr = s_mr;
  } else {
    __syncthreads();
    mr = st->mean_r;
  }
  if (!P.active) return;  // whole wavefront: no shuffles are left half-done
  // pipeline registers: rows p-3 .. p
  T x0 = 0, x1 = 0, b1 = 0, d0 = 1, d1 = 1, di1 = 1, res0 = 0, t0 = 0, t1 = 0, acc = 0;
  bool s0 = false, s1 = false;
  for (int pb = p_begin; pb < p_end; pb += kStreamPrefetch) {
#pragma unroll
    for (int u = 0; u < kStreamPrefetch; ++u) {
      const int p = pb + u;
      // stage A, row p
      uint8_t f = 0;
      if (pf_f[u]) f = cell_code(A, pf_f[u], pf_i[u], P.jr);
      const double rv = pf_r[u];
      issue(u, p + kStreamPrefetch);
      const bool s2 = (f & F_SOLV) != 0;
      const T b2 = s2 ? (T)(rv - mr) : (T)0;
      const T d2 = td[f], di2 = ti[f];
      const T x2 = w * di2 * b2;
      if (p >= P.i0 && p < P.i0 + L && p < A.nx && P.out_col) {
        const size_t k = (size_t)p * A.ny + P.j;
        x_out[k] = x2;
        b_out[k] = b2;
      }
      // stage B, row p-1: residual and t = D^-1 res
      const T xl = lane_up(x1), xr = lane_down(x1);
      const T res1 = s1 ? b1 - (d1 * x1 + cx * (x0 + x2) + cy * (xl + xr)) : (T)0;
      const T t2 = di1 * res1;
      // stage C, row p-2: y = res - w A t, restricted
      const T tl = lane_up(t1), tr = lane_down(t1);
      const int sr = p - 2;
      const bool own = sr >= P.i0 && sr < A.nx && P.out_col;
      T y = (s0 && own) ? res0 - w * (d0 * t1 + cx * (t0 + t2) + cy * (tl + tr)) : (T)0;
      if (FY == 2) y += lane_down(y);
      if (sr >= P.i0) {
        const bool first_row = FX == 1 || ((sr - P.i0) & 1) == 0;
        acc = first_row ? y : acc + y;
        const bool last_row = FX == 1 || ((sr - P.i0) & 1) == 1;
        const bool lane_ok = FY == 1 || ((P.lane - kStreamHalo) & 1) == 0;
        if (last_row && lane_ok && P.lane >= kStreamHalo && P.lane < kStreamHalo + kStreamOut) {
          const int I = sr / FX, J = P.j / FY;
          if (I < cnx && J < cny && P.j < A.ny) bc[(size_t)I * cny + J] = acc;
        }
      }
      // rotate the pipeline
      x0 = x1;
      x1 = x2;
      b1 = b2;
      d0 = d1;
      s0 = s1;
      d1 = d2;
      di1 = di2;
      s1 = s2;
      res0 = res1;
      t0 = t1;
      t1 = t2;
    }
  }
}

// Up leg: e = P_tent x_c on solvable cells, xs = x_pre + e - w D^-1 A e, then
// one post-sweep z = xs + w D^-1 (b0 - A xs); writes z and, for the
// preconditioned CG, the block partials of sum z, sum b0 z, sum b0.
template <typename T, int FX, int FY>
__global__ void __launch_bounds__(kBlock)
k_fine_up_stream(FineOp A, int strips, int L, T w, const T* __restrict__ x_pre,
                 const T* __restrict__ b0, const T* __restrict__ xc, int cny, T* __restrict__ z,
                 int reduce, double* partial, const SolveState* __restrict__ st) {
  if (st->done) return;
  __shared__ T td[128], ti[128];
  fine_tables<T>(A, td, ti);
  __syncthreads();
  const StreamPos P = stream_pos(A, strips, L);
  double sums[3] = {0.0, 0.0, 0.0};
  if (P.active) {
    const T cx = (T)A.cx, cy = (T)A.cy;
    const int p_begin = P.i0 - 2, p_end = P.i0 + L + 2;
    uint8_t pf_f[kStreamPrefetch];
    int pf_i[kStreamPrefetch];
    T pf_x[kStreamPrefetch], pf_b[kStreamPrefetch], pf_e[kStreamPrefetch];
    auto issue = [&](int slot, int p) {
      int ir = 0;
      pf_f[slot] = 0;
      pf_i[slot] = 0;
      pf_x[slot] = pf_b[slot] = pf_e[slot] = (T)0;
      if (p < p_end && stream_row(A, p, ir) && P.jin) {
        const int k = ir * A.ny + P.jr;
        pf_f[slot] = A.flags[k];
        pf_i[slot] = ir;
        pf_x[slot] = x_pre[k];
        pf_b[slot] = b0[k];
        pf_e[slot] = xc[(ir / FX) * cny + P.jr / FY];  // in range for every grid cell
      }
    };
#pragma unroll
    for (int u = 0; u < kStreamPrefetch; ++u) issue(u, p_begin + u);
    // rows p-3 .. p
    T e0 = 0, e1 = 0, x1 = 0, b0r = 0, b1 = 0, d0 = 1, d1 = 1, di0 = 1, di1 = 1;
    T xs0 = 0, xs1 = 0;
    bool s0 = false, s1 = false;
    for (int pb = p_begin; pb < p_end; pb += kStreamPrefetch) {
#pragma unroll
      for (int u = 0; u < kStreamPrefetch; ++u) {
        const int p = pb + u;
        uint8_t f = 0;
        if (pf_f[u]) f = cell_code(A, pf_f[u], pf_i[u], P.jr);
        const bool s2 = (f & F_SOLV) != 0;
        const T x2 = s2 ? pf_x[u] : (T)0, b2 = s2 ? pf_b[u] : (T)0, e2 = s2 ? pf_e[u] : (T)0;
        issue(u, p + kStreamPrefetch);
        const T d2 = td[f], di2 = ti[f];
        // stage B, row p-1: prolongation with the smoothed prolongator
        const T el = lane_up(e1), er = lane_down(e1);
        const T xs2 =
            s1 ? x1 + e1 - w * di1 * (d1 * e1 + cx * (e0 + e2) + cy * (el + er)) : (T)0;
        // stage C, row p-2: post-sweep
        const T ql = lane_up(xs1), qr = lane_down(xs1);
        const int sr = p - 2;
        if (sr >= P.i0 && sr < A.nx && P.out_col) {
          const T zv =
              s0 ? xs1 + w * di0 * (b0r - (d0 * xs1 + cx * (xs0 + xs2) + cy * (ql + qr))) : (T)0;
          z[(size_t)sr * A.ny + P.j] = zv;
          if (reduce != SR_NONE && s0) {
            const double bd = (double)b0r;
            sums[0] += (double)zv;
            sums[1] += bd * (double)zv;
            sums[2] += bd;
          }
        }
        e0 = e1;
        e1 = e2;
        x1 = x2;
        b0r = b1;
        b1 = b2;
        d0 = d1;
        di0 = di1;
        s0 = s1;
        d1 = d2;
        di1 = di2;
        s1 = s2;
        xs0 = xs1;
        xs1 = xs2;
      }
    }
  }
  if (reduce == SR_NONE) return;
  const int ops[3] = {kSum, kSum, kSum};
  block_partials<3>(sums, ops, partial, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Levels >= 1: stored Galerkin rows (class dictionary).
// ---------------------------------------------------------------------------
// Row c of the level applied at region cell idx; four entries per trip with
// their loads issued together, summed in entry order.
// Row (entry range rr) at region cell idx from the global dictionary, four
// partial sums to keep the multiply-add chain short.
template <typename T>
__device__ __forceinline__ T class_row(const ClassRows<T>& L, int2 rr, const T* X, int idx) {
  T s0 = (T)0, s1 = (T)0, s2 = (T)0, s3 = (T)0;
  int e = rr.x;
  for (; e + 3 < rr.y; e += 4) {
    const int o0 = L.ent_off[e], o1 = L.ent_off[e + 1], o2 = L.ent_off[e + 2],
              o3 = L.ent_off[e + 3];
    const T a0 = L.ent_val[e], a1 = L.ent_val[e + 1], a2 = L.ent_val[e + 2],
            a3 = L.ent_val[e + 3];
    s0 += a0 * X[idx + o0];
    s1 += a1 * X[idx + o1];
    s2 += a2 * X[idx + o2];
    s3 += a3 * X[idx + o3];
  }
  for (; e < rr.y; ++e) s0 += L.ent_val[e] * X[idx + L.ent_off[e]];
  return (s0 + s1) + (s2 + s3);
}

// LDS of a stored-level tile: the hot row's coefficients and offsets, three
// vectors, per cell 1/diag, entry range and active flag.
template <typename T>
struct ClassLds {
  T *HV, *P, *Q, *S, *D;
  int2* RR;
  int* HO;
  uint8_t* A;
  __device__ ClassLds(unsigned char* smem, int rn, int hot_n) {
    HV = (T*)smem;
    P = HV + hot_n;
    Q = P + rn;
    S = Q + rn;
    D = S + rn;
    size_t off = (size_t)((unsigned char*)(D + rn) - smem);
    off = (off + 7) & ~(size_t)7;
    RR = (int2*)(smem + off);  // (-1, -1) outside the grid
    HO = (int*)(RR + rn);
    A = (uint8_t*)(HO + hot_n);
  }
};

// (A X) at every cell of ring g, handed to out(a, c, idx, rr, ax) (rr.x < 0:
// a cell outside the grid, ax = 0).  A thread takes the cells q = tid,
// tid + 256, ... in chunks of kRingRows and runs the hot row on all of a
// chunk in lockstep: one broadcast LDS read of each hot coefficient and
// offset serves every row, and the rows' X reads are independent.  Rows of
// another class then take the global dictionary.  Every X read stays inside
// the region (the ring lies at least one stencil radius inside it), so the
// hot sums run unpredicated and are used only by hot rows.
constexpr int kRingRows = 1;

template <typename T, class Out>
__device__ __forceinline__ void ring_rows(const ClassRows<T>& L, const ClassLds<T>& M,
                                          const Ring& g, int ry, const T* X, Out&& out) {
  const int w = g.c1 - g.c0;
  const int n = (g.a1 - g.a0) * w;
  if (n <= 0) return;
  const float inv = 1.0f / (float)w;
  const int hn = L.hot < 0 ? 0 : L.hot_n;
  for (int q0 = 0; q0 < n; q0 += kRingRows * kBlock) {
    int ra[kRingRows], rc[kRingRows], idx[kRingRows];
    int2 rr[kRingRows];
    T s[kRingRows][4];
#pragma unroll
    for (int r = 0; r < kRingRows; ++r) {
      const int q = min(q0 + r * kBlock + (int)threadIdx.x, n - 1);
      int a = (int)(((float)q + 0.5f) * inv);
      int c = q - a * w;
      if (c < 0) {
        --a;
        c += w;
      } else if (c >= w) {
        ++a;
        c -= w;
      }
      ra[r] = g.a0 + a;
      rc[r] = g.c0 + c;
      idx[r] = ra[r] * ry + rc[r];
      rr[r] = M.RR[idx[r]];
#pragma unroll
      for (int p = 0; p < 4; ++p) s[r][p] = (T)0;
    }
    int e = 0;
    for (; e + 3 < hn; e += 4) {
      const int o0 = M.HO[e], o1 = M.HO[e + 1], o2 = M.HO[e + 2], o3 = M.HO[e + 3];
      const T v0 = M.HV[e], v1 = M.HV[e + 1], v2 = M.HV[e + 2], v3 = M.HV[e + 3];
#pragma unroll
      for (int r = 0; r < kRingRows; ++r) {
        s[r][0] += v0 * X[idx[r] + o0];
        s[r][1] += v1 * X[idx[r] + o1];
        s[r][2] += v2 * X[idx[r] + o2];
        s[r][3] += v3 * X[idx[r] + o3];
      }
    }
    for (; e < hn; ++e) {
      const int o = M.HO[e];
      const T v = M.HV[e];
#pragma unroll
      for (int r = 0; r < kRingRows; ++r) s[r][0] += v * X[idx[r] + o];
    }
#pragma unroll
    for (int r = 0; r < kRingRows; ++r) {
      if (q0 + r * kBlock + (int)threadIdx.x >= n) break;
      T ax = (T)0;
      if (rr[r].x >= 0)
        ax = (hn > 0 && rr[r].x == L.hot_e0) ? (s[r][0] + s[r][1]) + (s[r][2] + s[r][3])
                                             : class_row(L, rr[r], X, idx[r]);
      out(ra[r], rc[r], idx[r], rr[r], ax);
    }
  }
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
  int2* RR = M.RR;
  const TileOrigin o = tile_origin(ft);
  stage_hot_row(L, M);

  // Load b and the row classes, and start from zero: x = w D^-1 b.
  T* D = M.D;
  uint8_t* A = M.A;
  struct ClsLd {
    bool ok;
    int2 rr;
    uint8_t act;
    T bv, dv;
  };
  for_region_batched<kLoadBatch>(
      ft.rx, ry,
      [&](int a, int c) {
        ClsLd v;
        int i, j;
        v.ok = resolve(o.gi0 + a, o.gj0 + c, L.nx, L.ny, L.px, L.py, i, j);
        v.rr = make_int2(-1, -1);
        v.act = 0;
        v.bv = v.dv = (T)0;
        if (v.ok) {
          const int k = i * L.ny + j;
          v.rr = L.rr[k];
          v.bv = b[k];
          v.dv = L.rdinv[k];
          v.act = L.ract[k];
        }
        return v;
      },
      [&](int a, int c, int idx, const ClsLd& v) {
        RR[idx] = v.rr;
        B[idx] = v.bv;
        D[idx] = v.dv;
        A[idx] = v.act;
        X[idx] = w * v.dv * v.bv;
      });
  __syncthreads();
  for (int s = 1; s < nu; ++s) {
    ring_rows(L, M, ring(ft, nu + 1 - s, rxs, rys), ry, X,
              [&](int, int, int idx, int2 rr, T ax) {
                Y[idx] = (rr.x < 0) ? (T)0 : X[idx] + w * D[idx] * (B[idx] - ax);
              });
    __syncthreads();
    T* t = X;
    X = Y;
    Y = t;
  }
  // Residual on ring 1 into Y.
  ring_rows(L, M, ring(ft, 1, rxs, rys), ry, X, [&](int, int, int idx, int2 rr, T ax) {
    Y[idx] = (rr.x < 0) ? (T)0 : B[idx] - ax;
  });
  __syncthreads();
  for_ring(ring(ft, 1, rxs, rys), [&](int a, int c) {
    const int idx = a * ry + c;
    const int ia = a - ft.hx, jc = c - ft.hy;
    if (ia >= 0 && ia < ft.tx && jc >= 0 && jc < ft.ty) {
      const int i = o.ti * ft.tx + ia, j = o.tj * ft.ty + jc;
      if (i < L.nx && j < L.ny) x_out[i * L.ny + j] = X[idx];
    }
    X[idx] = D[idx] * Y[idx];  // zero outside the grid (D = 0 there)
  });
  __syncthreads();
  // (r - w A t) on the tile's rows into B (free now), then each coarse cell
  // sums its block's rows in x-major order.
  ring_rows(L, M, ring(ft, 0, rxs, rys), ry, X, [&](int, int, int idx, int2, T ax) {
    B[idx] = A[idx] ? Y[idx] - w * ax : (T)0;
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
        acc += B[(ft.hx + ca * fx + a2) * ry + (ft.hy + cc * fy + c2)];
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
  int2* RR = M.RR;
  const TileOrigin o = tile_origin(ft);
  stage_hot_row(L, M);

  T* D = M.D;
  struct ClsUpLd {
    bool ok;
    int2 rr;
    uint8_t act;
    T xv, bv, ev, dv;
  };
  for_region_batched<kLoadBatch>(
      ft.rx, ry,
      [&](int a, int c) {
        ClsUpLd v;
        int i, j;
        v.ok = resolve(o.gi0 + a, o.gj0 + c, L.nx, L.ny, L.px, L.py, i, j);
        v.rr = make_int2(-1, -1);
        v.act = 0;
        v.xv = v.bv = v.ev = v.dv = (T)0;
        if (v.ok) {
          const int k = i * L.ny + j;
          v.rr = L.rr[k];
          v.xv = x_pre[k];
          v.bv = b[k];
          v.dv = L.rdinv[k];
          v.act = L.ract[k];
          v.ev = xc[(i / fx) * cny + (j / fy)];  // in range for every grid cell
        }
        return v;
      },
      [&](int a, int c, int idx, const ClsUpLd& v) {
        RR[idx] = v.rr;
        X[idx] = v.xv;
        B[idx] = v.bv;
        D[idx] = v.dv;
        E[idx] = v.act ? v.ev : (T)0;
      });
  __syncthreads();
  ring_rows(L, M, ring(ft, nu, rxs, rys), ry, E, [&](int, int, int idx, int2 rr, T ax) {
    if (rr.x >= 0) X[idx] = X[idx] + E[idx] - w * D[idx] * ax;
  });
  __syncthreads();
  T* Y = E;
  for (int sw = 1; sw <= nu; ++sw) {
    const bool last = sw == nu;
    ring_rows(L, M, ring(ft, nu - sw, rxs, rys), ry, X, [&](int a, int c, int idx, int2 rr, T ax) {
      const T val = (rr.x >= 0) ? X[idx] + w * D[idx] * (B[idx] - ax) : (T)0;
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
// Dense coarsest-level operator, one workgroup per row: each thread loads its
// kDensePer entries of the row (and of b, an L1/L2-resident vector) in one
// batch, so a row costs one memory round trip for rows up to
// kBlock * kDensePer entries.
// ---------------------------------------------------------------------------
constexpr int kDensePer = 8;

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_dense_apply(const double* __restrict__ C, int n, const T* __restrict__ b, T* __restrict__ x,
              const SolveState* __restrict__ st) {
  if (st->done) return;
  __shared__ double part[kBlock / kWave];
  const int lane = threadIdx.x % kWave;
  const int wv = threadIdx.x / kWave;
  for (int row = blockIdx.x; row < n; row += gridDim.x) {
    const double* c = C + (size_t)row * n;
    double acc = 0.0;
    for (int k0 = 0; k0 < n; k0 += kBlock * kDensePer) {
      double cv[kDensePer], bv[kDensePer];
#pragma unroll
      for (int t = 0; t < kDensePer; ++t) {
        const int k = min(k0 + t * kBlock + (int)threadIdx.x, n - 1);
        cv[t] = c[k];
        bv[t] = (double)b[k];
      }
#pragma unroll
      for (int t = 0; t < kDensePer; ++t)
        if (k0 + t * kBlock + (int)threadIdx.x < n) acc += cv[t] * bv[t];
    }
    acc = wave_sum(acc);
    if (lane == 0) part[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot = part[0];
      for (int q = 1; q < kBlock / kWave; ++q) tot += part[q];
      x[row] = (T)tot;
    }
    __syncthreads();
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

// ---------------------------------------------------------------------------
// Stored levels, streaming form.  The same walk as level 0's streaming legs
// with a (2R+1)-row register window and lane shifts up to R: one sweep,
// zero start, rows from the box dictionary (a cell's class selects a dense
// (2R+1)^2 row; cells outside the grid take the all-zero class, so every
// value outside stays zero without tests).  Valid lanes shrink by R per
// operator application, hence 2R halo lanes and rows each side.
// ---------------------------------------------------------------------------
template <typename T, int R>
__device__ __forceinline__ T box_apply(const BoxRows<T>& L, int c, const T (&win)[2 * R + 1],
                                       int lane) {
  constexpr int W = 2 * R + 1;
  const T* row = L.box + (size_t)c * (W * W);
  T acc = (T)0;
#pragma unroll
  for (int a = 0; a < W; ++a) {
#pragma unroll
    for (int d = -R; d <= R; ++d) {
      const T v = d == 0 ? win[a] : __shfl(win[a], lane + d, kWave);
      acc += row[a * W + d + R] * v;
    }
  }
  return acc;
}

template <typename T, int N>
__device__ __forceinline__ void win_push(T (&w)[N], T v) {
#pragma unroll
  for (int k = 0; k + 1 < N; ++k) w[k] = w[k + 1];
  w[N - 1] = v;
}

template <typename T, int R>
__global__ void __launch_bounds__(kBlock)
k_class_down_stream(BoxRows<T> L, int strips, int rows, T w, const T* __restrict__ b,
                    T* __restrict__ x_out, T* __restrict__ bc, int fx, int fy, int cnx, int cny,
                    const SolveState* __restrict__ st) {
  if (st->done) return;
  constexpr int H = 2 * R, OUT = kWave - 2 * H, W = 2 * R + 1;
  const int wave = (int)(blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave);
  const int lane = (int)(threadIdx.x % kWave);
  const int strip = wave % strips;
  const int i0 = (wave / strips) * rows;
  if (i0 >= L.nx) return;  // whole wavefront
  const int j = strip * OUT - H + lane;
  int jr = j;
  bool jin = true;
  if (L.py) jr = wrap_index(j, L.ny);
  else jin = j >= 0 && j < L.ny;
  const bool out_lane = lane >= H && lane < H + OUT;
  const bool out_col = out_lane && j < L.ny;
  T xw[W], tw[W], rw[R + 1], bw[R + 1], dw[W];
  int cw[W];
#pragma unroll
  for (int k = 0; k < W; ++k) {
    xw[k] = tw[k] = dw[k] = (T)0;
    cw[k] = L.zero_class;
  }
#pragma unroll
  for (int k = 0; k <= R; ++k) rw[k] = bw[k] = (T)0;
  T acc = (T)0;
  for (int p = i0 - H; p < i0 + rows + H; ++p) {
    // stage A, row p: class, b, x = w D^-1 b
    int ir = p;
    bool in = jin;
    if (L.px) ir = wrap_index(p, L.nx);
    else in = in && p >= 0 && p < L.nx;
    int c = L.zero_class;
    T bv = (T)0;
    if (in) {
      const size_t k = (size_t)ir * L.ny + jr;
      c = L.cls[k];
      bv = b[k];
    }
    const T dinv = L.cdinv[c];
    const T xv = w * dinv * bv;
    if (p >= i0 && p < i0 + rows && p < L.nx && out_col) x_out[(size_t)p * L.ny + j] = xv;
    win_push(xw, xv);
    win_push(cw, c);
    win_push(dw, dinv);
    win_push(bw, bv);
    // stage B, row p - R: res = b - A x, t = D^-1 res
    const int cq = cw[R];
    const T res = bw[0] - box_apply<T, R>(L, cq, xw, lane);
    win_push(rw, res);
    win_push(tw, dw[R] * res);
    // stage C, row p - 2R: y = res - w A t on active rows, restricted
    const int s = p - H;
    const int cs = cw[0];
    const T at = box_apply<T, R>(L, cs, tw, lane);
    const bool own = s >= i0 && s < L.nx && out_col && L.cactive[cs];
    T y = own ? rw[0] - w * at : (T)0;
    if (fy == 2) y += __shfl_down(y, 1, kWave);
    if (s >= i0) {
      const bool first = fx == 1 || ((s - i0) & 1) == 0;
      const bool last = fx == 1 || ((s - i0) & 1) == 1;
      acc = first ? y : acc + y;
      const bool lane_ok = out_lane && (fy == 1 || ((lane - H) & 1) == 0);
      if (last && lane_ok && j < L.ny) {
        const int I = s / fx, J = j / fy;
        if (I < cnx && J < cny) bc[(size_t)I * cny + J] = acc;
      }
    }
  }
}

template <typename T, int R>
__global__ void __launch_bounds__(kBlock)
k_class_up_stream(BoxRows<T> L, int strips, int rows, T w, const T* __restrict__ x_pre,
                  const T* __restrict__ b, const T* __restrict__ xc, int fx, int fy, int cny,
                  T* __restrict__ x_out, const SolveState* __restrict__ st) {
  if (st->done) return;
  constexpr int H = 2 * R, OUT = kWave - 2 * H, W = 2 * R + 1;
  const int wave = (int)(blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave);
  const int lane = (int)(threadIdx.x % kWave);
  const int strip = wave % strips;
  const int i0 = (wave / strips) * rows;
  if (i0 >= L.nx) return;
  const int j = strip * OUT - H + lane;
  int jr = j;
  bool jin = true;
  if (L.py) jr = wrap_index(j, L.ny);
  else jin = j >= 0 && j < L.ny;
  const bool out_col = lane >= H && lane < H + OUT && j < L.ny;
  T ew[W], qw[W], dw[W], bw[W], xw[R + 1];
  int cw[W];
#pragma unroll
  for (int k = 0; k < W; ++k) {
    ew[k] = qw[k] = dw[k] = bw[k] = (T)0;
    cw[k] = L.zero_class;
  }
#pragma unroll
  for (int k = 0; k <= R; ++k) xw[k] = (T)0;
  for (int p = i0 - H; p < i0 + rows + H; ++p) {
    // stage A, row p: class, x_pre, b, e = P_tent x_c on active rows
    int ir = p;
    bool in = jin;
    if (L.px) ir = wrap_index(p, L.nx);
    else in = in && p >= 0 && p < L.nx;
    int c = L.zero_class;
    T xv = (T)0, bv = (T)0, ev = (T)0;
    if (in) {
      const size_t k = (size_t)ir * L.ny + jr;
      c = L.cls[k];
      xv = x_pre[k];
      bv = b[k];
      ev = xc[(size_t)(ir / fx) * cny + jr / fy];  // in range for every grid cell
    }
    if (!L.cactive[c]) ev = (T)0;
    win_push(ew, ev);
    win_push(cw, c);
    win_push(dw, L.cdinv[c]);
    win_push(bw, bv);
    win_push(xw, xv);
    // stage B, row p - R: xs = x + e - w D^-1 A e
    const T xs = xw[0] + ew[R] - w * dw[R] * box_apply<T, R>(L, cw[R], ew, lane);
    win_push(qw, xs);
    // stage C, row p - 2R: z = xs + w D^-1 (b - A xs)
    const int s = p - H;
    const T az = box_apply<T, R>(L, cw[0], qw, lane);
    if (s >= i0 && s < L.nx && out_col)
      x_out[(size_t)s * L.ny + j] = qw[R] + w * dw[0] * (bw[0] - az);
  }
}

int box_strips(int ny, int radius) {
  const int out = kWave - 4 * radius;
  return (ny + out - 1) / out;
}

int box_blocks(int nx, int ny, int radius, int rows) {
  const long waves = (long)box_strips(ny, radius) * ((nx + rows - 1) / rows);
  return (int)((waves + kBlock / kWave - 1) / (kBlock / kWave));
}

#define PCFD_BOX_DISPATCH(KERNEL, ...)                                                       \
  do {                                                                                     \
    switch (radius) {                                                                      \
      case 1:                                                                              \
        PCFD_LAUNCH_TIMED(clk, (KERNEL<T, 1>), grid, dim3(kBlock), 0, s, __VA_ARGS__);     \
        break;                                                                             \
      case 2:                                                                              \
        PCFD_LAUNCH_TIMED(clk, (KERNEL<T, 2>), grid, dim3(kBlock), 0, s, __VA_ARGS__);     \
        break;                                                                             \
      default:                                                                             \
        PCFD_LAUNCH_TIMED(clk, (KERNEL<T, 3>), grid, dim3(kBlock), 0, s, __VA_ARGS__);     \
        break;                                                                             \
    }                                                                                      \
  } while (0)

template <typename T>
int launch_class_down_stream(hipStream_t s, const BoxRows<T>& L, int radius, int rows, T w,
                             const T* b, T* x_out, T* bc, int fx, int fy, int cnx, int cny,
                             const SolveState* st, const LaunchClock& clk) {
  if (radius < 1 || radius > kBoxMaxRadius || rows < 2 || rows % 2) {
    set_last_error("streaming stored level: radius %d, %d rows (radius 1-%d, even rows)", radius,
                   rows, kBoxMaxRadius);
    return kErrBadArg;
  }
  const dim3 grid(box_blocks(L.nx, L.ny, radius, rows));
  const int strips = box_strips(L.ny, radius);
  PCFD_BOX_DISPATCH(k_class_down_stream, L, strips, rows, w, b, x_out, bc, fx, fy, cnx, cny, st);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int launch_class_up_stream(hipStream_t s, const BoxRows<T>& L, int radius, int rows, T w,
                           const T* x_pre, const T* b, const T* xc, int fx, int fy, int cny,
                           T* x_out, const SolveState* st, const LaunchClock& clk) {
  if (radius < 1 || radius > kBoxMaxRadius || rows < 2 || rows % 2) {
    set_last_error("streaming stored level: radius %d, %d rows (radius 1-%d, even rows)", radius,
                   rows, kBoxMaxRadius);
    return kErrBadArg;
  }
  const dim3 grid(box_blocks(L.nx, L.ny, radius, rows));
  const int strips = box_strips(L.ny, radius);
  PCFD_BOX_DISPATCH(k_class_up_stream, L, strips, rows, w, x_pre, b, xc, fx, fy, cny, x_out, st);
  PCFD_LAUNCH_CHECK();
  return kOk;
}
#undef PCFD_BOX_DISPATCH

int stream_strips(int ny) { return (ny + kStreamOut - 1) / kStreamOut; }

int stream_blocks(int nx, int ny, int rows) {
  const long waves = (long)stream_strips(ny) * ((nx + rows - 1) / rows);
  return (int)((waves + kBlock / kWave - 1) / (kBlock / kWave));
}

bool stream_supported(const FineOp& A, int nu, int fx, int fy) {
  return !A.hxc && nu == 1 && (fx == 1 || fx == 2) && (fy == 1 || fy == 2);
}

#define PCFD_STREAM_DISPATCH(KERNEL, ...)                                                    \
  do {                                                                                     \
    if (fx == 2 && fy == 2)                                                                \
      PCFD_LAUNCH_TIMED(clk, (KERNEL<T, 2, 2>), grid, dim3(kBlock), 0, s, __VA_ARGS__);    \
    else if (fx == 2)                                                                      \
      PCFD_LAUNCH_TIMED(clk, (KERNEL<T, 2, 1>), grid, dim3(kBlock), 0, s, __VA_ARGS__);    \
    else if (fy == 2)                                                                      \
      PCFD_LAUNCH_TIMED(clk, (KERNEL<T, 1, 2>), grid, dim3(kBlock), 0, s, __VA_ARGS__);    \
    else                                                                                   \
      PCFD_LAUNCH_TIMED(clk, (KERNEL<T, 1, 1>), grid, dim3(kBlock), 0, s, __VA_ARGS__);    \
  } while (0)

template <typename T>
int launch_fine_down_stream(hipStream_t s, const FineOp& A, int rows, T w, const double* r,
                            T* b_out, T* x_out, T* bc, int fx, int fy, int cnx, int cny,
                            SolveState* st, const UpdateFold& uf, const LaunchClock& clk) {
  if (rows < 4 || rows % 4 || (fx == 2 && rows % 2)) {
    set_last_error("streaming down leg: %d rows per wavefront (need a multiple of 4)", rows);
    return kErrBadArg;
  }
  const dim3 grid(stream_blocks(A.nx, A.ny, rows));
  const int strips = stream_strips(A.ny);
  PCFD_STREAM_DISPATCH(k_fine_down_stream, A, strips, rows, w, r, b_out, x_out, bc, cnx, cny, st,
                       uf);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int launch_fine_up_stream(hipStream_t s, const FineOp& A, int rows, T w, const T* x_pre,
                          const T* b0, const T* xc, int fx, int fy, int cny, T* z, int reduce,
                          double* partial, const SolveState* st, const LaunchClock& clk) {
  if (rows < 4 || rows % 4) {
    set_last_error("streaming up leg: %d rows per wavefront (need a multiple of 4)", rows);
    return kErrBadArg;
  }
  const dim3 grid(stream_blocks(A.nx, A.ny, rows));
  const int strips = stream_strips(A.ny);
  PCFD_STREAM_DISPATCH(k_fine_up_stream, A, strips, rows, w, x_pre, b0, xc, cny, z, reduce,
                       partial, st);
  PCFD_LAUNCH_CHECK();
  return kOk;
}
#undef PCFD_STREAM_DISPATCH

template <typename T>
int launch_class_down(hipStream_t s, const ClassRows<T>& L, const FusedTile& ft, int rxs,
                      int rys, T w, int nu, const T* b, T* x_out, T* bc, int fx, int fy, int cnx,
                      int cny, const SolveState* st, const LaunchClock& clk) {
  PCFD_LAUNCH_TIMED(clk, k_class_down<T>, dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, L, ft, rxs,
                    rys, w, nu, b, x_out, bc, fx, fy, cnx, cny, st);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int launch_class_up(hipStream_t s, const ClassRows<T>& L, const FusedTile& ft, int rxs, int rys,
                    T w, int nu, const T* x_pre, const T* b, const T* xc, int fx, int fy, int cny,
                    T* x_out, const SolveState* st, const LaunchClock& clk) {
  PCFD_LAUNCH_TIMED(clk, k_class_up<T>, dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, L, ft, rxs,
                    rys, w, nu, x_pre, b, xc, fx, fy, cny, x_out, st);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int launch_fine_dir(hipStream_t s, const FineOp& A, const FusedTile& ft, const T* z,
                    const double* p_old, double* p_new, double* q, double* partial,
                    SolveState* st, const CycleFold& cf, const LaunchClock& clk) {
  if (A.hxc)
    PCFD_LAUNCH_TIMED(clk, (k_fine_dir<T, true>), dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, A,
                      ft, z, p_old, p_new, q, partial, st, cf);
  else
    PCFD_LAUNCH_TIMED(clk, (k_fine_dir<T, false>), dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, A,
                      ft, z, p_old, p_new, q, partial, st, cf);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int launch_dense_apply(hipStream_t s, const double* C, int n, const T* b, T* x,
                       const SolveState* st, const LaunchClock& clk) {
  const int blocks = std::max(1, std::min(n, 8192));
  PCFD_LAUNCH_TIMED(clk, k_dense_apply<T>, dim3(blocks), dim3(kBlock), 0, s, C, n, b, x, st);
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
                                    const SolveState*, const LaunchClock&);                   \
  template int launch_class_up<T>(hipStream_t, const ClassRows<T>&, const FusedTile&, int, int, \
                                  T, int, const T*, const T*, const T*, int, int, int, T*,    \
                                  const SolveState*, const LaunchClock&);                     \
  template int launch_dense_apply<T>(hipStream_t, const double*, int, const T*, T*,            \
                                     const SolveState*, const LaunchClock&);                  \
  template int launch_fine_down_stream<T>(hipStream_t, const FineOp&, int, T, const double*,   \
                                          T*, T*, T*, int, int, int, int, SolveState*,        \
                                          const UpdateFold&, const LaunchClock&);             \
  template int launch_fine_up_stream<T>(hipStream_t, const FineOp&, int, T, const T*, const T*, \
                                        const T*, int, int, int, T*, int, double*,            \
                                        const SolveState*, const LaunchClock&);               \
  template int launch_class_down_stream<T>(hipStream_t, const BoxRows<T>&, int, int, T,        \
                                           const T*, T*, T*, int, int, int, int,              \
                                           const SolveState*, const LaunchClock&);            \
  template int launch_class_up_stream<T>(hipStream_t, const BoxRows<T>&, int, int, T, const T*, \
                                         const T*, const T*, int, int, int, T*,               \
                                         const SolveState*, const LaunchClock&);              \
  template int launch_fine_dir<T>(hipStream_t, const FineOp&, const FusedTile&, const T*,      \
                                  const double*, double*, double*, double*, SolveState*,       \
                                  const CycleFold&, const LaunchClock&);
PCFD_VCYCLE_INST(float)
PCFD_VCYCLE_INST(double)
#undef PCFD_VCYCLE_INST

}  // namespace pcfd
