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

__device__ __forceinline__ int wrap_index(int i, int n) {
  const int m = i % n;
  return m < 0 ? m + n : m;
}

// Global cell of a region cell; false when it lies outside a non-periodic grid.
__device__ __forceinline__ bool resolve(int gi, int gj, int nx, int ny, int px, int py, int& i,
                                        int& j) {
  i = px ? wrap_index(gi, nx) : gi;
  j = py ? wrap_index(gj, ny) : gj;
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
__device__ __forceinline__ bool on_dirichlet(const FineOp& A, int i, int j) {
  return (A.dir[0] && i == 0) || (A.dir[1] && i == A.nx - 1) || (A.dir[2] && j == 0) ||
         (A.dir[3] && j == A.ny - 1);
}

// Diagonal and its inverse for the 32 face-flag combinations of cells away
// from Dirichlet walls (pycfd/core/pressure.py:219-232 order: x faces, then y).
template <typename T>
__device__ __forceinline__ void fine_tables(const FineOp& A, T* td, T* ti) {
  const int f = threadIdx.x;
  if (f < 32) {
    double d = 1.0;
    if (f & F_SOLV) {
      d = 0.0;
      if (f & F_L) d -= A.cx;
      if (f & F_R) d -= A.cx;
      if (f & F_B) d -= A.cy;
      if (f & F_T) d -= A.cy;
    }
    td[f] = (T)d;
    ti[f] = (T)(1.0 / d);
  }
}

template <typename T>
__device__ __forceinline__ T rdiag(const FineOp& A, const T* td, uint8_t f, int gi, int gj) {
  return (f & F_DIRB) ? (T)fine_diag(A, f & 31, gi, gj) : td[f & 31];
}

template <typename T>
__device__ __forceinline__ T rdinv(const FineOp& A, const T* ti, uint8_t f, int gi, int gj) {
  return (f & F_DIRB) ? (T)(1.0 / fine_diag(A, f & 31, gi, gj)) : ti[f & 31];
}

// (A x) at region cell idx, terms in CSR column order.
template <typename T>
__device__ __forceinline__ T rapply(const FineOp& A, uint8_t f, const T* X, int idx, int ry, T d) {
  T y = (T)0;
  if (f & F_L) y += (T)A.cx * X[idx - ry];
  if (f & F_B) y += (T)A.cy * X[idx - 1];
  y += d * X[idx];
  if (f & F_T) y += (T)A.cy * X[idx + 1];
  if (f & F_R) y += (T)A.cx * X[idx + ry];
  return y;
}

template <typename T>
struct FineLds {
  T *td, *ti, *P, *Q, *S;
  uint8_t* F;
  __device__ FineLds(unsigned char* smem, int rn) {
    td = (T*)smem;
    ti = td + 32;
    P = ti + 32;
    Q = P + rn;
    S = Q + rn;
    F = (uint8_t*)(S + rn);
  }
};

template <typename T>
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
  uint8_t* F = M.F;
  const TileOrigin o = tile_origin(ft);
  fine_tables<T>(A, M.td, M.ti);
  const double mr = r ? st->mean_r : 0.0;
  const T m = shift_in ? (T)st->mean_x : (T)0;

  for_ring(Ring{0, ft.rx, 0, ry}, [&](int a, int c) {
    int i, j;
    uint8_t f = 0;
    T bv = (T)0, xv = (T)0;
    if (resolve(o.gi0 + a, o.gj0 + c, A.nx, A.ny, A.px, A.py, i, j)) {
      const int k = i * A.ny + j;
      f = A.flags[k];
      if (r) bv = (f & F_SOLV) ? (T)(r[k] - mr) : (T)0;
      else bv = b_in[k];
      if (nonzero) {
        const T x0 = x_in[k];
        xv = (f & F_SOLV) ? x0 - m : x0;
      }
      if ((f & F_SOLV) && on_dirichlet(A, i, j)) f |= F_DIRB;
    }
    const int idx = a * ry + c;
    F[idx] = f;
    B[idx] = bv;
    X[idx] = xv;
  });
  __syncthreads();

  // Pre-smoothing; v is the number of stencil rings still valid.
  int v = ft.hx;
  int s0 = 0;
  if (!nonzero) {  // x = 0 + w D^-1 (b - A 0)
    for_ring(Ring{0, ft.rx, 0, ry}, [&](int a, int c) {
      const int idx = a * ry + c;
      X[idx] = w * rdinv(A, M.ti, F[idx], o.gi0 + a, o.gj0 + c) * B[idx];
    });
    __syncthreads();
    s0 = 1;
  }
  for (int s = s0; s < nu; ++s) {
    --v;
    for_ring(ring(ft, v, 1, 1), [&](int a, int c) {
      const int idx = a * ry + c;
      const uint8_t f = F[idx];
      const int gi = o.gi0 + a, gj = o.gj0 + c;
      const T d = rdiag(A, M.td, f, gi, gj);
      Y[idx] = X[idx] + (w * rdinv(A, M.ti, f, gi, gj)) * (B[idx] - rapply(A, f, X, idx, ry, d));
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
    Y[idx] = B[idx] - rapply(A, f, X, idx, ry, rdiag(A, M.td, f, o.gi0 + a, o.gj0 + c));
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
    X[idx] = rdinv(A, M.ti, F[idx], o.gi0 + a, o.gj0 + c) * Y[idx];
  });
  __syncthreads();
  // Restriction: one thread per coarse cell of the tile.
  const int ctx = ft.tx / fx, cty = ft.ty / fy;
  for (int q = threadIdx.x; q < ctx * cty; q += kBlock) {
    const int ca = q / cty, cc = q - ca * cty;
    const int I = o.ti * ctx + ca, J = o.tj * cty + cc;
    if (I >= cnx || J >= cny) continue;
    T acc = (T)0;
    for (int a2 = 0; a2 < fx; ++a2) {
      const int i = I * fx + a2;
      if (i >= A.nx) break;
      for (int c2 = 0; c2 < fy; ++c2) {
        const int j = J * fy + c2;
        if (j >= A.ny) break;
        const int a = ft.hx + ca * fx + a2, c = ft.hy + cc * fy + c2;
        const int idx = a * ry + c;
        const uint8_t f = F[idx];
        if (!(f & F_SOLV)) continue;
        const T atr = rapply(A, f, X, idx, ry, rdiag(A, M.td, f, o.gi0 + a, o.gj0 + c));
        acc += Y[idx] - w * atr;
      }
    }
    bc[I * cny + J] = acc;
  }
}

template <typename T>
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
  uint8_t* F = M.F;
  const TileOrigin o = tile_origin(ft);
  fine_tables<T>(A, M.td, M.ti);

  for_ring(Ring{0, ft.rx, 0, ry}, [&](int a, int c) {
    int i, j;
    uint8_t f = 0;
    T xv = (T)0, bv = (T)0, ev = (T)0;
    if (resolve(o.gi0 + a, o.gj0 + c, A.nx, A.ny, A.px, A.py, i, j)) {
      const int k = i * A.ny + j;
      f = A.flags[k];
      xv = x_pre[k];
      bv = b0[k];
      if (f & F_SOLV) ev = xc[(i / fx) * cny + (j / fy)];
      if ((f & F_SOLV) && on_dirichlet(A, i, j)) f |= F_DIRB;
    }
    const int idx = a * ry + c;
    F[idx] = f;
    X[idx] = xv;
    B[idx] = bv;
    E[idx] = ev;
  });
  __syncthreads();
  // x += e - w D^-1 A e on ring nu.
  for_ring(ring(ft, nu, 1, 1), [&](int a, int c) {
    const int idx = a * ry + c;
    const uint8_t f = F[idx];
    const int gi = o.gi0 + a, gj = o.gj0 + c;
    const T d = rdiag(A, M.td, f, gi, gj);
    const T ae = rapply(A, f, E, idx, ry, d);
    X[idx] += E[idx] - w * rdinv(A, M.ti, f, gi, gj) * ae;
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
      const int gi = o.gi0 + a, gj = o.gj0 + c;
      const T d = rdiag(A, M.td, f, gi, gj);
      const T val =
          X[idx] + (w * rdinv(A, M.ti, f, gi, gj)) * (B[idx] - rapply(A, f, X, idx, ry, d));
      if (!last) {
        Y[idx] = val;
        return;
      }
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

// ---------------------------------------------------------------------------
// Levels >= 1: stored Galerkin rows (class dictionary).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T class_row(const ClassRows<T>& L, int c, const T* X, int idx) {
  T acc = (T)0;
  const int e1 = L.row_ptr[c + 1];
  for (int e = L.row_ptr[c]; e < e1; ++e) acc += L.ent_val[e] * X[idx + L.ent_off[e]];
  return acc;
}

template <typename T>
struct ClassLds {
  T *P, *Q, *S;
  uint16_t* C;
  __device__ ClassLds(unsigned char* smem, int rn) {
    P = (T*)smem;
    Q = P + rn;
    S = Q + rn;
    C = (uint16_t*)(S + rn);
  }
};

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_class_down(ClassRows<T> L, FusedTile ft, int rxs, int rys, T w, int nu, const T* __restrict__ b,
             T* __restrict__ x_out, T* __restrict__ bc, int fx, int fy, int cnx, int cny,
             const SolveState* __restrict__ st) {
  if (st->done) return;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ry = ft.ry;
  ClassLds<T> M(smem, ft.rx * ry);
  T* B = M.P;
  T* X = M.Q;
  T* Y = M.S;
  uint16_t* C = M.C;
  const TileOrigin o = tile_origin(ft);

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
      Y[idx] = X[idx] + w * L.cdinv[cl] * (B[idx] - class_row(L, cl, X, idx));
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
    Y[idx] = (cl == kOutside) ? (T)0 : B[idx] - class_row(L, cl, X, idx);
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
        acc += L.cactive[cl] ? Y[idx] - w * class_row(L, cl, X, idx) : (T)0;
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
  ClassLds<T> M(smem, ft.rx * ry);
  T* E = M.P;
  T* X = M.Q;
  T* B = M.S;
  uint16_t* C = M.C;
  const TileOrigin o = tile_origin(ft);

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
    X[idx] = X[idx] + E[idx] - w * L.cdinv[cl] * class_row(L, cl, E, idx);
  });
  __syncthreads();
  T* Y = E;
  for (int sw = 1; sw <= nu; ++sw) {
    const bool last = sw == nu;
    for_ring(ring(ft, nu - sw, rxs, rys), [&](int a, int c) {
      const int idx = a * ry + c;
      const uint16_t cl = C[idx];
      T val = (T)0;
      if (cl != kOutside) val = X[idx] + w * L.cdinv[cl] * (B[idx] - class_row(L, cl, X, idx));
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
                     const SolveState* st) {
  hipLaunchKernelGGL(k_fine_down<T>, dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, A, ft, w, nu,
                     nonzero, r, b_in, x_in, shift_in, b_out, x_out, bc, fx, fy, cnx, cny, st);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

template <typename T>
int launch_fine_up(hipStream_t s, const FineOp& A, const FusedTile& ft, T w, int nu,
                   const T* x_pre, const T* b0, const T* xc, int fx, int fy, int cny, T* z,
                   int reduce, double* partial, const SolveState* st) {
  hipLaunchKernelGGL(k_fine_up<T>, dim3(ft.ntiles()), dim3(kBlock), ft.lds, s, A, ft, w, nu, x_pre,
                     b0, xc, fx, fy, cny, z, reduce, partial, st);
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
                                   int, int, int, const SolveState*);                         \
  template int launch_fine_up<T>(hipStream_t, const FineOp&, const FusedTile&, T, int,         \
                                 const T*, const T*, const T*, int, int, int, T*, int, double*, \
                                 const SolveState*);                                          \
  template int launch_class_down<T>(hipStream_t, const ClassRows<T>&, const FusedTile&, int,   \
                                    int, T, int, const T*, T*, T*, int, int, int, int,        \
                                    const SolveState*);                                       \
  template int launch_class_up<T>(hipStream_t, const ClassRows<T>&, const FusedTile&, int, int, \
                                  T, int, const T*, const T*, const T*, int, int, int, T*,    \
                                  const SolveState*);                                         \
  template int launch_dense_apply<T>(hipStream_t, const double*, int, const T*, T*,            \
                                     const SolveState*);
PCFD_VCYCLE_INST(float)
PCFD_VCYCLE_INST(double)
#undef PCFD_VCYCLE_INST

}  // namespace pcfd
