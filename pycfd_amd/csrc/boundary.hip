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

// One wall's velocity condition (BoundaryCondition.apply, boundary.py:141-164).
__device__ void apply_wall(const BCParams& P, int w, bool predictor, double* u, double* v) {
  const int kind = P.kind[w];
  if (kind == BC_PERIODIC) return;
  const WallIdx wi = wall_index(w, P.nx, P.ny);
  const int ur = P.nx + 3, uc = P.ny + 2, vr = P.nx + 2, vc = P.ny + 3;
  Line nf, nin, ng, tg, tin;
  if (wi.axis == 0) {
    nf = line_of(u, P.ldu, ur, uc, 0, wi.n_face);
    nin = line_of(u, P.ldu, ur, uc, 0, wi.n_in);
    ng = line_of(u, P.ldu, ur, uc, 0, wi.n_ghost);
    tg = line_of(v, P.ldv, vr, vc, 0, wi.t_ghost);
    tin = line_of(v, P.ldv, vr, vc, 0, wi.t_in);
  } else {
    nf = line_of(v, P.ldv, vr, vc, 1, wi.n_face);
    nin = line_of(v, P.ldv, vr, vc, 1, wi.n_in);
    ng = line_of(v, P.ldv, vr, vc, 1, wi.n_ghost);
    tg = line_of(u, P.ldu, ur, uc, 1, wi.t_ghost);
    tin = line_of(u, P.ldu, ur, uc, 1, wi.t_in);
  }
  const bool prescribes = kind != BC_PRESSURE_OUTLET;
  const bool derives = kind == BC_OUTLET;
  if (prescribes && (predictor || !derives)) {
    for (int k = threadIdx.x; k < nf.len; k += blockDim.x) {
      double val = 0.0;
      if (kind == BC_INLET) val = P.inlet[w][k];
      else if (kind == BC_OUTLET) val = nin[k];
      nf[k] = val;
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < ng.len; k += blockDim.x)
    ng[k] = (kind == BC_PRESSURE_OUTLET) ? nf[k] : 2.0 * nf[k] - nin[k];
  for (int k = threadIdx.x; k < tg.len; k += blockDim.x) {
    const double t = tin[k];
    double val;
    if (kind == BC_MOVING_WALL) val = 2.0 * P.vel[w] - t;
    else if (kind == BC_NO_SLIP || kind == BC_INLET) val = -t;
    else val = t;  // OUTLET, PRESSURE_OUTLET, SYMMETRY: zero gradient
    tg[k] = val;
  }
  __syncthreads();
}

// BoundaryManager._wrap_velocity (boundary.py:388-410).
__device__ void wrap_velocity(const BCParams& P, double* u, double* v) {
  const int nx = P.nx, ny = P.ny;
  if (P.periodic_x) {
    for (int k = threadIdx.x; k < ny + 2; k += blockDim.x) {
      u[(size_t)(nx + 1) * P.ldu + k] = u[(size_t)1 * P.ldu + k];
      u[(size_t)0 * P.ldu + k] = u[(size_t)nx * P.ldu + k];
      u[(size_t)(nx + 2) * P.ldu + k] = u[(size_t)2 * P.ldu + k];
    }
    for (int k = threadIdx.x; k < ny + 3; k += blockDim.x) {
      v[(size_t)0 * P.ldv + k] = v[(size_t)nx * P.ldv + k];
      v[(size_t)(nx + 1) * P.ldv + k] = v[(size_t)1 * P.ldv + k];
    }
    __syncthreads();
  }
  if (P.periodic_y) {
    for (int k = threadIdx.x; k < nx + 2; k += blockDim.x) {
      double* row = v + (size_t)k * P.ldv;
      row[ny + 1] = row[1];
      row[0] = row[ny];
      row[ny + 2] = row[2];
    }
    for (int k = threadIdx.x; k < nx + 3; k += blockDim.x) {
      double* row = u + (size_t)k * P.ldu;
      row[0] = row[ny];
      row[ny + 1] = row[1];
    }
    __syncthreads();
  }
}

// ProjectionSolver._wrap_periodic_faces (solver.py:601-607).
__device__ void wrap_faces(const BCParams& P, double* u, double* v) {
  const int nx = P.nx, ny = P.ny;
  if (P.periodic_x)
    for (int k = threadIdx.x; k < ny + 2; k += blockDim.x)
      u[(size_t)(nx + 1) * P.ldu + k] = u[(size_t)1 * P.ldu + k];
  if (P.periodic_y)
    for (int k = threadIdx.x; k < nx + 2; k += blockDim.x)
      v[(size_t)k * P.ldv + ny + 1] = v[(size_t)k * P.ldv + 1];
  __syncthreads();
}

// BoundaryManager.enforce_global_mass_balance (boundary.py:423-487).
__device__ void mass_balance(const BCParams& P, double* u, double* v) {
  double influx = 0.0, outflux = 0.0, total_area = 0.0;
  int n_out = 0;
  for (int w = 0; w < 4; ++w) {
    const int kind = P.kind[w];
    if (kind == BC_PERIODIC) continue;
    const WallIdx wi = wall_index(w, P.nx, P.ny);
    Line f = (wi.axis == 0) ? line_of(u, P.ldu, P.nx + 3, P.ny + 2, 0, wi.n_face)
                            : line_of(v, P.ldv, P.nx + 2, P.ny + 3, 1, wi.n_face);
    const int n = (wi.axis == 0) ? P.ny : P.nx;
    const double width = (wi.axis == 0) ? P.dy : P.dx;
    // stretched mesh: each face weighted by its own height / width
    const double* wk = (wi.axis == 0) ? P.hy : P.hx;
    double part = 0.0;
    if (wk) {
      for (int k = threadIdx.x; k < n; k += blockDim.x) part += f[1 + k] * wk[k];
    } else {
      for (int k = threadIdx.x; k < n; k += blockDim.x) part += f[1 + k];
    }
    const double s = block_sum(part);
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
      for (int k = threadIdx.x; k < n; k += blockDim.x) f[1 + k] *= scale;
    }
  } else {
    for (int w = 0; w < 4; ++w) {
      if (P.kind[w] != BC_OUTLET) continue;
      const WallIdx wi = wall_index(w, P.nx, P.ny);
      const int n = (wi.axis == 0) ? P.ny : P.nx;
      const double width = (wi.axis == 0) ? P.dy : P.dx;
      const double* wk = (wi.axis == 0) ? P.hy : P.hx;
      if (wk) {
        double part = 0.0;
        for (int k = threadIdx.x; k < n; k += blockDim.x) part += wk[k];
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
      for (int k = threadIdx.x; k < n; k += blockDim.x) f[1 + k] += add;
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
    auto ua = [&](int k) -> double { return P.uf_area ? u[P.uf[k]] * P.uf_area[k] : u[P.uf[k]]; };
    auto va = [&](int k) -> double { return P.vf_area ? v[P.vf[k]] * P.vf_area[k] : v[P.vf[k]]; };
    auto scaled = [&](double sum, bool weighted) -> double {
      return weighted ? sum / dt : sum * P.area / dt;
    };
    double su = 0.0, sv = 0.0;
    for (int k = threadIdx.x; k < P.n_uf; k += blockDim.x) su += ua(k);
    for (int k = threadIdx.x; k < P.n_vf; k += blockDim.x) sv += va(k);
    const double fx = scaled(block_sum(su), P.uf_area != nullptr);
    const double fy = scaled(block_sum(sv), P.vf_area != nullptr);
    if (threadIdx.x == 0) {
      if (mode == BOP_MASK_RESET) {
        P.force[0] = fx;
        P.force[1] = fy;
      } else {
        P.force[0] = P.force[0] + fx;
        P.force[1] = P.force[1] + fy;
      }
    }
    for (int b = 0; b < P.n_bodies; ++b) {
      double bu = 0.0, bv = 0.0;
      for (int k = threadIdx.x; k < P.n_uf; k += blockDim.x)
        if (P.uf_body[k] == b) bu += ua(k);
      for (int k = threadIdx.x; k < P.n_vf; k += blockDim.x)
        if (P.vf_body[k] == b) bv += va(k);
      const double bx = scaled(block_sum(bu), P.uf_area != nullptr);
      const double by = scaled(block_sum(bv), P.vf_area != nullptr);
      if (threadIdx.x == 0) {
        if (mode == BOP_MASK_RESET) {
          P.force[2 + 2 * b] = bx;
          P.force[3 + 2 * b] = by;
        } else {
          P.force[2 + 2 * b] = P.force[2 + 2 * b] + bx;
          P.force[3 + 2 * b] = P.force[3 + 2 * b] + by;
        }
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < P.n_uf; k += blockDim.x) u[P.uf[k]] = 0.0;
  for (int k = threadIdx.x; k < P.n_vf; k += blockDim.x) v[P.vf[k]] = 0.0;
  __syncthreads();
}

}  // namespace

__global__ void __launch_bounds__(kBlock)
k_boundary_pass(BCParams P, BCOps ops, double dt, double* u, double* v) {
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
    for (int k = threadIdx.x; k < g.len; k += blockDim.x)
      g[k] = (kind == BC_PRESSURE_OUTLET) ? 2.0 * pref - in[k] : in[k];
    __syncthreads();
  }
  if (P.periodic_x) {
    for (int k = threadIdx.x; k < ny + 2; k += blockDim.x) {
      p[k] = p[(size_t)nx * ldp + k];
      p[(size_t)(nx + 1) * ldp + k] = p[(size_t)ldp + k];
    }
    __syncthreads();
  }
  if (P.periodic_y) {
    for (int k = threadIdx.x; k < nx + 2; k += blockDim.x) {
      double* row = p + (size_t)k * ldp;
      row[0] = row[ny];
      row[ny + 1] = row[1];
    }
    __syncthreads();
  }
}

int launch_boundary_pass(hipStream_t s, const BCParams& P, const BCOps& ops, double dt, double* u,
                         double* v) {
  if (ops.n == 0) return kOk;
  hipLaunchKernelGGL(k_boundary_pass, dim3(1), dim3(kBlock), 0, s, P, ops, dt, u, v);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

int launch_pressure_ghosts(hipStream_t s, const BCParams& P, double* p, int ldp) {
  hipLaunchKernelGGL(k_pressure_ghosts, dim3(1), dim3(kBlock), 0, s, P, p, ldp);
  PCFD_LAUNCH_CHECK();
  return kOk;
}

}  // namespace pcfd
