// Boundary work of the projection step: wall conditions, periodic wrap,
// global mass balance, immersed-boundary direct forcing with its reaction
// force, and the pressure ghost layer.
//
// All of it touches O(nx + ny) values plus the O(perimeter) immersed-boundary
// face list, so one 256-thread workgroup runs a whole "pass" (an ordered list
// of operations) with a workgroup barrier between dependent line writes.  The
// order of operations, including which ghost corners are written last,
// follows pycfd/core/boundary.py:141-164, 375-420, 423-487 and
// pycfd/core/solver.py:553-607, so every written value is bit-identical to
// the reference (the reductions aside, which are deterministic tree sums).
#include "internal.hpp"

#include <algorithm>

namespace pcfd {

namespace {

struct WallIdx {
  int axis, n_face, n_in, n_ghost, t_ghost, t_in, p_ghost, p_in, outward;
};

// pycfd/core/boundary.py:69-80
__device__ __forceinline__ WallIdx wall_index(int w, int nx, int ny) {
  switch (w) {
    case 0: return {0, 1, 2, 0, 0, 1, 0, 1, -1};
    case 1: return {0, nx + 1, nx, nx + 2, nx + 1, nx, nx + 1, nx, +1};
    case 2: return {1, 1, 2, 0, 0, 1, 0, 1, -1};
    default: return {1, ny + 1, ny, ny + 2, ny + 1, ny, ny + 1, ny, +1};
  }
}

// A row (axis 0) or column (axis 1) of a ghosted array.
struct Line {
  double* base;
  size_t stride;
  int len;
  __device__ __forceinline__ double& operator[](int k) const { return base[(size_t)k * stride]; }
};

__device__ __forceinline__ Line line_of(double* arr, int ld, int rows, int cols, int axis,
                                        int idx) {
  if (axis == 0) return Line{arr + (size_t)idx * ld, 1, cols};
  return Line{arr + idx, (size_t)ld, rows};
}

// Lines run in chunks of kSlots entries per thread, every read of a chunk
// issued before any of its writes: one memory round trip per chunk instead
// of one per entry.  A thread visits its entries k = tid, tid + blockDim, ...
// in increasing order, as a plain strided loop does, and the reads and writes
// of one call touch different lines (or the same entry of one line), so
// every written value and every partial sum is the plain loop's.
constexpr int kSlots = 8;

// put(k, get(k)) for k in [0, n).
template <class Get, class Put>
__device__ __forceinline__ void line_map(int n, const Get& get, const Put& put) {
  if (n <= 0) return;
  for (int base = 0; base < n; base += kSlots * (int)blockDim.x) {
    double val[kSlots];
#pragma unroll
    for (int t = 0; t < kSlots; ++t)
      val[t] = get(min(base + t * (int)blockDim.x + (int)threadIdx.x, n - 1));
#pragma unroll
    for (int t = 0; t < kSlots; ++t) {
      const int k = base + t * (int)blockDim.x + (int)threadIdx.x;
      if (k < n) put(k, val[t]);
    }
  }
}

// This thread's sum of get(k) over its entries of [0, n) with take(k).
template <class Get, class Take>
__device__ __forceinline__ double line_part(int n, const Get& get, const Take& take) {
  double part = 0.0;
  if (n <= 0) return part;
  for (int base = 0; base < n; base += kSlots * (int)blockDim.x) {
    double val[kSlots];
    bool use[kSlots];
#pragma unroll
    for (int t = 0; t < kSlots; ++t) {
      const int k = base + t * (int)blockDim.x + (int)threadIdx.x;
      val[t] = get(min(k, n - 1));
      use[t] = k < n && take(min(k, n - 1));
    }
#pragma unroll
    for (int t = 0; t < kSlots; ++t)
      if (use[t]) part += val[t];
  }
  return part;
}

__device__ double block_sum(double x) {
  double v[1] = {x};
  const int ops[1] = {kSum};
  block_reduce<1>(v, ops);
  __shared__ double out;
  if (threadIdx.x == 0) out = v[0];
  __syncthreads();
  const double r = out;
  __syncthreads();
  return r;
}

// The five lines one wall's velocity condition touches and what it does
// (BoundaryCondition.apply, boundary.py:141-164): the normal face line nf
// (written when `face`), its ghost line ng from nf and the interior line nin,
// the tangential ghost line tg from the interior tangential line tin.  Each
// entry k of a line depends only on entry k of the others.
struct WallLines {
  Line nf, nin, ng, tg, tin;
  int kind, w;
  bool face;
};

__device__ __forceinline__ WallLines wall_lines(const BCParams& P, int w, bool predictor,
                                                double* u, double* v) {
  const int kind = P.kind[w];
  const WallIdx wi = wall_index(w, P.nx, P.ny);
  const int ur = P.nx + 3, uc = P.ny + 2, vr = P.nx + 2, vc = P.ny + 3;
  WallLines L;
  if (wi.axis == 0) {
    L.nf = line_of(u, P.ldu, ur, uc, 0, wi.n_face);
    L.nin = line_of(u, P.ldu, ur, uc, 0, wi.n_in);
    L.ng = line_of(u, P.ldu, ur, uc, 0, wi.n_ghost);
    L.tg = line_of(v, P.ldv, vr, vc, 0, wi.t_ghost);
    L.tin = line_of(v, P.ldv, vr, vc, 0, wi.t_in);
  } else {
    L.nf = line_of(v, P.ldv, vr, vc, 1, wi.n_face);
    L.nin = line_of(v, P.ldv, vr, vc, 1, wi.n_in);
    L.ng = line_of(v, P.ldv, vr, vc, 1, wi.n_ghost);
    L.tg = line_of(u, P.ldu, ur, uc, 1, wi.t_ghost);
    L.tin = line_of(u, P.ldu, ur, uc, 1, wi.t_in);
  }
  L.kind = kind;
  L.w = w;
  const bool prescribes = kind != BC_PRESSURE_OUTLET;
  const bool derives = kind == BC_OUTLET;
  L.face = prescribes && (predictor || !derives);
  return L;
}

// INLET: the inlet line; OUTLET: the interior line; any other kind: zero.  The
// one load goes through a selected pointer (the interior line when the wall
// has no inlet array), so no branch surrounds it.
__device__ __forceinline__ double face_value(const BCParams& P, const WallLines& L, int k) {
  const bool inl = L.kind == BC_INLET;
  const double* src = inl ? P.inlet[L.w] : L.nin.base;
  const size_t stride = inl ? 1 : L.nin.stride;
  const double val = src[(size_t)k * stride];
  return (inl || L.kind == BC_OUTLET) ? val : 0.0;
}

__device__ __forceinline__ double ghost_value(const WallLines& L, int k) {
  return (L.kind == BC_PRESSURE_OUTLET) ? L.nf[k] : 2.0 * L.nf[k] - L.nin[k];
}

__device__ __forceinline__ double tangential_value(const BCParams& P, const WallLines& L, int k) {
  const double t = L.tin[k];
  if (L.kind == BC_MOVING_WALL) return 2.0 * P.vel[L.w] - t;
  if (L.kind == BC_NO_SLIP || L.kind == BC_INLET) return -t;
  return t;  // OUTLET, PRESSURE_OUTLET, SYMMETRY: zero gradient
}

// One wall's velocity condition over the workgroup.
__device__ void apply_wall(const BCParams& P, int w, bool predictor, double* u, double* v) {
  if (P.kind[w] == BC_PERIODIC) return;
  const WallLines L = wall_lines(P, w, predictor, u, v);
  if (L.face) {
    line_map(L.nf.len, [&](int k) { return face_value(P, L, k); },
             [&](int k, double val) { L.nf[k] = val; });
    __syncthreads();
  }
  line_map(L.ng.len, [&](int k) { return ghost_value(L, k); },
           [&](int k, double val) { L.ng[k] = val; });
  line_map(L.tg.len, [&](int k) { return tangential_value(P, L, k); },
           [&](int k, double val) { L.tg[k] = val; });
  __syncthreads();
}

// Transpose of the four walls' velocity conditions (apply_wall in wall
// order): each condition is a sequence of line assignments target = a *
// source + constant, so its transpose runs them in reverse, adding a times
// the target's adjoint into the source's and zeroing the target's (a
// constant assignment, the prescribed normal face, only zeroes).  lu, lv
// are adjoints in the ghosted layouts of u and v.  Walls without a periodic
// axis and without a derived (OUTLET) normal face only.
__device__ void apply_wall_adjoint(const BCParams& P, int w, bool predictor, double* lu,
                                   double* lv) {
  if (P.kind[w] == BC_PERIODIC) return;
  const WallLines L = wall_lines(P, w, predictor, lu, lv);
  const bool po = L.kind == BC_PRESSURE_OUTLET;
  const double tb = (L.kind == BC_MOVING_WALL || L.kind == BC_NO_SLIP || L.kind == BC_INLET)
                        ? -1.0 : 1.0;
  // ng = 2 nf - nin (PO: ng = nf); tg = tb * tin (+ constant)
  for (int k = threadIdx.x; k < L.ng.len; k += blockDim.x) {
    const double g = L.ng[k];
    L.ng[k] = 0.0;
    L.nf[k] += po ? g : 2.0 * g;
    if (!po) L.nin[k] -= g;
  }
  for (int k = threadIdx.x; k < L.tg.len; k += blockDim.x) {
    const double g = L.tg[k];
    L.tg[k] = 0.0;
    L.tin[k] += tb * g;
  }
  __syncthreads();
  if (L.face) {
    for (int k = threadIdx.x; k < L.nf.len; k += blockDim.x) L.nf[k] = 0.0;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlock)
k_walls_adjoint(BCParams P, int predictor, double* lu, double* lv) {
  for (int w = 3; w >= 0; --w) apply_wall_adjoint(P, w, predictor != 0, lu, lv);
}

// BoundaryManager._wrap_velocity (boundary.py:388-410).
__device__ void wrap_velocity(const BCParams& P, double* u, double* v) {
  const int nx = P.nx, ny = P.ny;
  // Copies between distinct lines (the sources are interior lines, the
  // destinations ghost lines), so all reads may go first.
  auto copy = [&](double* arr, size_t ld, int n, int axis, int dst, int src) {
    const size_t sd = axis == 0 ? 1 : ld;
    double* d = axis == 0 ? arr + (size_t)dst * ld : arr + dst;
    const double* s = axis == 0 ? arr + (size_t)src * ld : arr + src;
    line_map(n, [&](int k) { return s[(size_t)k * sd]; },
             [&](int k, double val) { d[(size_t)k * sd] = val; });
  };
  if (P.periodic_x) {
    copy(u, P.ldu, ny + 2, 0, nx + 1, 1);
    copy(u, P.ldu, ny + 2, 0, 0, nx);
    copy(u, P.ldu, ny + 2, 0, nx + 2, 2);
    copy(v, P.ldv, ny + 3, 0, 0, nx);
    copy(v, P.ldv, ny + 3, 0, nx + 1, 1);
    __syncthreads();
  }
  if (P.periodic_y) {
    copy(v, P.ldv, nx + 2, 1, ny + 1, 1);
    copy(v, P.ldv, nx + 2, 1, 0, ny);
    copy(v, P.ldv, nx + 2, 1, ny + 2, 2);
    copy(u, P.ldu, nx + 3, 1, 0, ny);
    copy(u, P.ldu, nx + 3, 1, ny + 1, 1);
    __syncthreads();
  }
}

// ProjectionSolver._wrap_periodic_faces (solver.py:601-607).
__device__ void wrap_faces(const BCParams& P, double* u, double* v) {
  const int nx = P.nx, ny = P.ny;
  if (P.periodic_x) {
    double* d = u + (size_t)(nx + 1) * P.ldu;
    const double* s = u + (size_t)1 * P.ldu;
    line_map(ny + 2, [&](int k) { return s[k]; }, [&](int k, double val) { d[k] = val; });
  }
  if (P.periodic_y) {
    const size_t ld = P.ldv;
    line_map(
        nx + 2, [&](int k) { return v[(size_t)k * ld + 1]; },
        [&](int k, double val) { v[(size_t)k * ld + ny + 1] = val; });
  }
  __syncthreads();
}

// BoundaryManager.enforce_global_mass_balance (boundary.py:423-487).
__device__ void mass_balance(const BCParams& P, double* u, double* v) {
  double influx = 0.0, outflux = 0.0, total_area = 0.0;
  int n_out = 0;
  // The four walls' face sums: each thread's partial of a wall is the plain
  // strided loop's (kSlots entries a chunk, increasing k), all walls' loads
  // issued together when every line fits one chunk, then one block
  // reduction (the same tree per wall as a reduction of its own).
  double part[4];
  const int chunk = kSlots * (int)blockDim.x;
  if (P.nx <= chunk && P.ny <= chunk) {
    double val[4][kSlots];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const WallIdx wi = wall_index(w, P.nx, P.ny);
      Line f = (wi.axis == 0) ? line_of(u, P.ldu, P.nx + 3, P.ny + 2, 0, wi.n_face)
                              : line_of(v, P.ldv, P.nx + 2, P.ny + 3, 1, wi.n_face);
      const int n = (wi.axis == 0) ? P.ny : P.nx;
      const double* wk = (wi.axis == 0) ? P.hy : P.hx;
      // in bounds on every wall (ghosted lines); the weight test is uniform
      // and stays outside the slots, so no load sits in a branch of its own
      if (wk) {
#pragma unroll
        for (int t = 0; t < kSlots; ++t) {
          const int k = min(t * (int)blockDim.x + (int)threadIdx.x, n - 1);
          val[w][t] = f[1 + k] * wk[k];
        }
      } else {
#pragma unroll
        for (int t = 0; t < kSlots; ++t) {
          const int k = min(t * (int)blockDim.x + (int)threadIdx.x, n - 1);
          val[w][t] = f[1 + k];
        }
      }
    }
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const int n = (w < 2) ? P.ny : P.nx;
      const bool on = P.kind[w] != BC_PERIODIC;
      part[w] = 0.0;
#pragma unroll
      for (int t = 0; t < kSlots; ++t)
        if (on && t * (int)blockDim.x + (int)threadIdx.x < n) part[w] += val[w][t];
    }
  } else {
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const WallIdx wi = wall_index(w, P.nx, P.ny);
      Line f = (wi.axis == 0) ? line_of(u, P.ldu, P.nx + 3, P.ny + 2, 0, wi.n_face)
                              : line_of(v, P.ldv, P.nx + 2, P.ny + 3, 1, wi.n_face);
      const int n = (wi.axis == 0) ? P.ny : P.nx;
      // stretched mesh: each face weighted by its own height / width
      const double* wk = (wi.axis == 0) ? P.hy : P.hx;
      auto all = [](int) { return true; };
      part[w] = (P.kind[w] == BC_PERIODIC) ? 0.0
                : wk ? line_part(n, [&](int k) { return f[1 + k] * wk[k]; }, all)
                     : line_part(n, [&](int k) { return f[1 + k]; }, all);
    }
  }
  const int ops4[4] = {kSum, kSum, kSum, kSum};
  block_reduce<4>(part, ops4);
  __shared__ double wall_sum[4];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < 4; ++w) wall_sum[w] = part[w];
  }
  __syncthreads();
  for (int w = 0; w < 4; ++w) {
    const int kind = P.kind[w];
    if (kind == BC_PERIODIC) continue;
    const WallIdx wi = wall_index(w, P.nx, P.ny);
    const double width = (wi.axis == 0) ? P.dy : P.dx;
    const double* wk = (wi.axis == 0) ? P.hy : P.hx;
    const double s = wall_sum[w];
    const double flux_out = (double)wi.outward * (wk ? s : s * width);
    if (kind == BC_OUTLET) {
      outflux += flux_out;
      ++n_out;
    } else {
      influx -= flux_out;
    }
  }
  if (n_out == 0) return;
  if (outflux > 1.0e-12) {
    const double scale = influx / outflux;
    for (int w = 0; w < 4; ++w) {
      if (P.kind[w] != BC_OUTLET) continue;
      const WallIdx wi = wall_index(w, P.nx, P.ny);
      Line f = (wi.axis == 0) ? line_of(u, P.ldu, P.nx + 3, P.ny + 2, 0, wi.n_face)
                              : line_of(v, P.ldv, P.nx + 2, P.ny + 3, 1, wi.n_face);
      const int n = (wi.axis == 0) ? P.ny : P.nx;
      line_map(n, [&](int k) { return f[1 + k] * scale; },
               [&](int k, double val) { f[1 + k] = val; });
    }
  } else {
    for (int w = 0; w < 4; ++w) {
      if (P.kind[w] != BC_OUTLET) continue;
      const WallIdx wi = wall_index(w, P.nx, P.ny);
      const int n = (wi.axis == 0) ? P.ny : P.nx;
      const double width = (wi.axis == 0) ? P.dy : P.dx;
      const double* wk = (wi.axis == 0) ? P.hy : P.hx;
      if (wk) {
        const double part = line_part(n, [&](int k) { return wk[k]; }, [](int) { return true; });
        total_area = total_area + block_sum(part);
      } else {
        total_area = total_area + (double)n * width;
      }
    }
    const double deficit = influx - outflux;
    for (int w = 0; w < 4; ++w) {
      if (P.kind[w] != BC_OUTLET) continue;
      const WallIdx wi = wall_index(w, P.nx, P.ny);
      Line f = (wi.axis == 0) ? line_of(u, P.ldu, P.nx + 3, P.ny + 2, 0, wi.n_face)
                              : line_of(v, P.ldv, P.nx + 2, P.ny + 3, 1, wi.n_face);
      const int n = (wi.axis == 0) ? P.ny : P.nx;
      const double add = (double)wi.outward * deficit / total_area;
      line_map(n, [&](int k) { return f[1 + k] + add; },
               [&](int k, double val) { f[1 + k] = val; });
    }
  }
  __syncthreads();
}

// ProjectionSolver._mask_obstacle (solver.py:553-599): momentum removed at the
// faces touching a solid cell, then those faces zeroed.
__device__ void mask_obstacle(const BCParams& P, int mode, double dt, double* u, double* v) {
  if (!P.has_obstacle) return;
  if (mode != BOP_MASK_ZERO) {
    // Uniform mesh: one area after the sum; stretched: each face weighted
    // by its control volume before it (solver.py:197-207).
    const bool wu = P.uf_area != nullptr, wv = P.vf_area != nullptr;
    // One chunk of faces: field[idx[k]] (times area[k] when there is an area
    // table; x * 1.0 == x exactly otherwise) and the body of each, at clamped
    // faces.  The tests are uniform and sit around whole slot loops, so no
    // load is in a branch of its own.
    // First the index, body and area loads of u's and v's faces, then the
    // field loads that need the indices, the area products after both: two
    // memory round trips a chunk.
    auto gather_idx = [&](int n, const int* idx, const double* area, const int8_t* body,
                          int base, int (&id)[kSlots], double (&ar)[kSlots],
                          int (&bod)[kSlots]) {
      if (n <= 0) return;
#pragma unroll
      for (int t = 0; t < kSlots; ++t) {
        const int k = min(base + t * (int)blockDim.x + (int)threadIdx.x, n - 1);
        id[t] = idx[k];
        bod[t] = body[k];
      }
      if (area) {
#pragma unroll
        for (int t = 0; t < kSlots; ++t)
          ar[t] = area[min(base + t * (int)blockDim.x + (int)threadIdx.x, n - 1)];
      }
    };
    auto gather_val = [&](int n, const double* field, const int (&id)[kSlots],
                          double (&val)[kSlots]) {
      if (n <= 0) return;
#pragma unroll
      for (int t = 0; t < kSlots; ++t) val[t] = field[id[t]];
    };
    auto scaled = [&](double sum, bool weighted) -> double {
      return weighted ? sum / dt : sum * P.area / dt;
    };
    // The total and kMaskBodies bodies' sums of u and v from one read of the
    // faces per chunk; each thread adds its faces in increasing order (the
    // plain strided loop's partials), then one block reduction.
    constexpr int kMaskBodies = 3;
    constexpr int K = 2 + 2 * kMaskBodies;
    const int chunk = kSlots * (int)blockDim.x;
    const int nmax = max(P.n_uf, P.n_vf);
    for (int b0 = 0; b0 == 0 || b0 < P.n_bodies; b0 += kMaskBodies) {
      double acc[K];
#pragma unroll
      for (int c = 0; c < K; ++c) acc[c] = 0.0;
      for (int base = 0; base < nmax; base += chunk) {
        double uval[kSlots], vval[kSlots], uar[kSlots], var[kSlots];
        int uid[kSlots], vid[kSlots], ubody[kSlots], vbody[kSlots];
#pragma unroll
        for (int t = 0; t < kSlots; ++t) {
          uval[t] = vval[t] = 0.0;
          uar[t] = var[t] = 1.0;
          uid[t] = vid[t] = 0;
          ubody[t] = vbody[t] = -1;
        }
        gather_idx(P.n_uf, P.uf, P.uf_area, P.uf_body, base, uid, uar, ubody);
        gather_idx(P.n_vf, P.vf, P.vf_area, P.vf_body, base, vid, var, vbody);
        gather_val(P.n_uf, u, uid, uval);
        gather_val(P.n_vf, v, vid, vval);
#pragma unroll
        for (int t = 0; t < kSlots; ++t) {  // after both, so the two rounds overlap
          uval[t] *= uar[t];
          vval[t] *= var[t];
        }
#pragma unroll
        for (int t = 0; t < kSlots; ++t) {
          const int k = base + t * (int)blockDim.x + (int)threadIdx.x;
          if (k < P.n_uf) {
            acc[0] += uval[t];
#pragma unroll
            for (int c = 0; c < kMaskBodies; ++c)
              if (ubody[t] == b0 + c) acc[2 + 2 * c] += uval[t];
          }
          if (k < P.n_vf) {
            acc[1] += vval[t];
#pragma unroll
            for (int c = 0; c < kMaskBodies; ++c)
              if (vbody[t] == b0 + c) acc[3 + 2 * c] += vval[t];
          }
        }
      }
      int ops[K];
#pragma unroll
      for (int c = 0; c < K; ++c) ops[c] = kSum;
      block_reduce<K>(acc, ops);
      if (threadIdx.x == 0) {
        const bool reset = mode == BOP_MASK_RESET;
        if (b0 == 0) {
          const double fx = scaled(acc[0], wu), fy = scaled(acc[1], wv);
          P.force[0] = reset ? fx : P.force[0] + fx;
          P.force[1] = reset ? fy : P.force[1] + fy;
        }
        for (int c = 0; c < kMaskBodies && b0 + c < P.n_bodies; ++c) {
          const int b = b0 + c;
          const double bx = scaled(acc[2 + 2 * c], wu), by = scaled(acc[3 + 2 * c], wv);
          P.force[2 + 2 * b] = reset ? bx : P.force[2 + 2 * b] + bx;
          P.force[3 + 2 * b] = reset ? by : P.force[3 + 2 * b] + by;
        }
      }
    }
  }
  __syncthreads();
  // the faces zeroed: a chunk's u and v face indices loaded together, then
  // the stores (one memory round trip a chunk)
  for (int base = 0; base < max(P.n_uf, P.n_vf); base += kSlots * (int)blockDim.x) {
    int uid[kSlots], vid[kSlots];
#pragma unroll
    for (int t = 0; t < kSlots; ++t) {
      const int k = base + t * (int)blockDim.x + (int)threadIdx.x;
      uid[t] = k < P.n_uf ? P.uf[k] : -1;
      vid[t] = k < P.n_vf ? P.vf[k] : -1;
    }
#pragma unroll
    for (int t = 0; t < kSlots; ++t) {
      if (uid[t] >= 0) u[uid[t]] = 0.0;
      if (vid[t] >= 0) v[vid[t]] = 0.0;
    }
  }
  __syncthreads();
}

}  // namespace

// The velocity condition of all four walls (no periodic axis) over many
// workgroups.  Walls apply in order 0..3 and meet only at the corners: a
// later wall overwrites, or reads, entries an earlier one wrote within the
// first or last kWallCorner entries of its lines.  Those entries run in
// workgroup (0, 0) wall by wall, a barrier between walls, one thread per
// entry (per entry the face value, then the ghost value that reads it);
// every other entry k of a wall touches entry k of that wall's lines only,
// nothing any other wall reads or writes, and runs on its own thread
// (blockIdx.y = wall).
constexpr int kWallCorner = 4;

// With mask_op set, workgroup (0, 4) runs that immersed-boundary mask beside
// the walls (the faces lie apart from every wall line: BCParams::mask_apart).
__global__ void __launch_bounds__(kBlock)
k_walls(BCParams P, int predictor, double* u, double* v, int mask_op, StepScalars T) {
  if (blockIdx.y == 4) {
    if (blockIdx.x == 0) mask_obstacle(P, mask_op, T.get_dt(), u, v);
    return;
  }
  auto corner = [](int k, int len) { return k < kWallCorner || k >= len - kWallCorner; };
  const int w = blockIdx.y;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  const WallLines L = wall_lines(P, w, predictor != 0, u, v);
  if (L.face && k < L.nf.len && !corner(k, L.nf.len)) L.nf[k] = face_value(P, L, k);
  if (k < L.ng.len && !corner(k, L.ng.len)) L.ng[k] = ghost_value(L, k);
  if (k < L.tg.len && !corner(k, L.tg.len)) L.tg[k] = tangential_value(P, L, k);
  if (blockIdx.x != 0 || blockIdx.y != 0) return;  // whole workgroup: barriers below
  // thread t < 2C: corner entry t of the normal lines (face and ghost have
  // one length), 2C <= t < 4C: of the tangential line; entry index q and
  // whether it exists (a line shorter than 2C has fewer corner entries)
  auto corner_q = [](int t, int len, int& q) {
    q = t < kWallCorner ? t : len - 2 * kWallCorner + t;
    return t < kWallCorner ? q < len : q >= max(len - kWallCorner, kWallCorner) && q < len;
  };
  const int t = threadIdx.x;
  for (int ww = 0; ww < 4; ++ww) {
    const WallLines C = wall_lines(P, ww, predictor != 0, u, v);
    int q = 0;
    if (t < 2 * kWallCorner) {
      if (corner_q(t, C.nf.len, q)) {
        if (C.face) C.nf[q] = face_value(P, C, q);
        C.ng[q] = ghost_value(C, q);
      }
    } else if (t < 4 * kWallCorner) {
      if (corner_q(t - 2 * kWallCorner, C.tg.len, q)) C.tg[q] = tangential_value(P, C, q);
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlock)
k_boundary_pass(BCParams P, BCOps ops, StepScalars T, double* u, double* v) {
  const double dt = T.get_dt();
  for (int q = 0; q < ops.n; ++q) {
    switch (ops.op[q]) {
      case BOP_VEL_PRED:
      case BOP_VEL_CORR: {
        const bool pred = ops.op[q] == BOP_VEL_PRED;
        for (int w = 0; w < 4; ++w) apply_wall(P, w, pred, u, v);
        wrap_velocity(P, u, v);
        break;
      }
      case BOP_MASS: mass_balance(P, u, v); break;
      case BOP_MASK_RESET:
      case BOP_MASK_ACC:
      case BOP_MASK_ZERO: mask_obstacle(P, ops.op[q], dt, u, v); break;
      case BOP_WRAP_FACES: wrap_faces(P, u, v); break;
      default: break;
    }
  }
}

// Pressure ghost layer (BoundaryManager.apply_pressure, boundary.py:166-169,
// 293-303, 381-385, 412-420).
__global__ void __launch_bounds__(kBlock)
k_pressure_ghosts(BCParams P, double* p, int ldp) {
  const int nx = P.nx, ny = P.ny;
  for (int w = 0; w < 4; ++w) {
    const int kind = P.kind[w];
    if (kind == BC_PERIODIC) continue;
    const WallIdx wi = wall_index(w, nx, ny);
    Line g = line_of(p, ldp, nx + 2, ny + 2, wi.axis, wi.p_ghost);
    Line in = line_of(p, ldp, nx + 2, ny + 2, wi.axis, wi.p_in);
    const double pref = P.pref[w];
    line_map(
        g.len, [&](int k) { return (kind == BC_PRESSURE_OUTLET) ? 2.0 * pref - in[k] : in[k]; },
        [&](int k, double val) { g[k] = val; });
    __syncthreads();
  }
  if (P.periodic_x) {
    double* r0 = p;
    double* r1 = p + (size_t)(nx + 1) * ldp;
    const double* sn = p + (size_t)nx * ldp;
    const double* s1 = p + (size_t)ldp;
    line_map(ny + 2, [&](int k) { return sn[k]; }, [&](int k, double val) { r0[k] = val; });
    line_map(ny + 2, [&](int k) { return s1[k]; }, [&](int k, double val) { r1[k] = val; });
    __syncthreads();
  }
  if (P.periodic_y) {
    const size_t ld = ldp;
    line_map(
        nx + 2, [&](int k) { return p[(size_t)k * ld + ny]; },
        [&](int k, double val) { p[(size_t)k * ld] = val; });
    line_map(
        nx + 2, [&](int k) { return p[(size_t)k * ld + 1]; },
        [&](int k, double val) { p[(size_t)k * ld + ny + 1] = val; });
    __syncthreads();
  }
}

// Without a periodic axis (and on a grid wide enough that opposite walls
// touch disjoint lines) the wall conditions run on the whole device
// (k_walls); the rest of the pass (O(perimeter) lines and reductions) stays
// on one workgroup, in the pass's order.  A mask whose faces lie apart from
// the wall lines commutes with the walls and the mass balance, so it moves
// into the pass's walls launch (one workgroup beside the wall entries).
int launch_boundary_pass(hipStream_t s, const BCParams& P, const BCOps& ops,
                         const StepScalars& T, double* u, double* v, const ClockSource& clk) {
  if (ops.n == 0) return kOk;
  const bool walls_apart = !P.periodic_x && !P.periodic_y && P.nx >= 2 * kWallCorner &&
                           P.ny >= 2 * kWallCorner;
  auto is_mask = [](int op) {
    return op == BOP_MASK_RESET || op == BOP_MASK_ACC || op == BOP_MASK_ZERO;
  };
  int n_walls = 0, n_mask = 0, mask_op = 0;
  for (int q = 0; q < ops.n; ++q) {
    n_walls += ops.op[q] == BOP_VEL_PRED || ops.op[q] == BOP_VEL_CORR;
    if (is_mask(ops.op[q])) {
      ++n_mask;
      mask_op = ops.op[q];
    }
  }
  const bool fuse_mask = walls_apart && P.mask_apart && n_walls == 1 && n_mask == 1;
  BCOps rest{};
  auto flush = [&]() {
    if (rest.n == 0) return;
    const LaunchClock c = clk.get(kClockBoundaryPass, 0.0);
    PCFD_LAUNCH_TIMED(c, k_boundary_pass, dim3(1), dim3(kBlock), 0, s, P, rest, T, u, v);
    rest.n = 0;
  };
  for (int q = 0; q < ops.n; ++q) {
    const int op = ops.op[q];
    if (walls_apart && (op == BOP_VEL_PRED || op == BOP_VEL_CORR)) {
      flush();
      const int len = std::max(P.nx, P.ny) + 3;
      const LaunchClock c = clk.get(kClockWalls, 0.0);
      PCFD_LAUNCH_TIMED(c, k_walls, dim3((len + kBlock - 1) / kBlock, fuse_mask ? 5 : 4),
                        dim3(kBlock), 0, s, P, op == BOP_VEL_PRED ? 1 : 0, u, v,
                        fuse_mask ? mask_op : 0, T);
    } else if (fuse_mask && is_mask(op)) {
      continue;  // runs in the walls launch
    } else if (op != BOP_WRAP_FACES || !walls_apart) {  // face wrap: periodic axes only
      rest.op[rest.n++] = op;
    }
  }
  flush();
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int launch_walls_adjoint(hipStream_t s, const BCParams& P, bool predictor, double* lu,
                         double* lv) {
  hipLaunchKernelGGL(k_walls_adjoint, dim3(1), dim3(kBlock), 0, s, P, predictor ? 1 : 0, lu, lv);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

__global__ void k_set_step_scalars(double* dst, double dt, double blend) {
  if (threadIdx.x == 0) {
    dst[0] = dt;
    dst[1] = blend;
  }
}

int launch_set_step_scalars(hipStream_t s, double* dst, double dt, double blend) {
  hipLaunchKernelGGL(k_set_step_scalars, dim3(1), dim3(64), 0, s, dst, dt, blend);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int launch_pressure_ghosts(hipStream_t s, const BCParams& P, double* p, int ldp) {
  hipLaunchKernelGGL(k_pressure_ghosts, dim3(1), dim3(kBlock), 0, s, P, p, ldp);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

}  // namespace pcfd
