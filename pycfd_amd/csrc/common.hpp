// Shared device/host helpers for the gfx950 projection-step kernels.
//
// Conventions used by every kernel in this library:
//   * Fields keep the reference's ghosted C-order layout (x slow, y fast):
//       u (nx+3, ny+2), v (nx+2, ny+3), p (nx+2, ny+2).
//     Each array is addressed through a row pitch (`ld`, in elements) so a
//     padded allocation needs no kernel change.
//   * Poisson vectors are the interior cells flattened as k = i*ny + j
//     (pycfd/core/pressure.py:178), length nx*ny, contiguous.
//   * Workgroups are 256 threads = 4 wavefronts of 64.  Tiles are 64 cells
//     along y (one wavefront per row segment, fully coalesced FP64 loads)
//     by 4 rows along x.
//   * Reductions are deterministic: every block writes its partials to a
//     slot indexed by its tile number and a one-workgroup fold launch sums
//     the slots in a fixed order.  No floating-point atomics are used.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cmath>

namespace pcfd {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kTileY = 64;
constexpr int kTileX = 4;
constexpr int kNumXcd = 8;

// Status codes shared with include/pycfd_hip.h.
enum Status : int {
  kOk = 0,
  kErrHip = 1,
  kErrNotConverged = 2,
  kErrStalled = 3,
  kErrNonFinite = 4,
  kErrBadArg = 5,
  kErrUnsupported = 6,
  kErrNoDevice = 7,
};

void set_last_error(const char* fmt, ...);

#define PCFD_HIP_CHECK(expr)                                                  \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::pcfd::set_last_error("%s failed: %s (%s:%d)", #expr,                  \
                             hipGetErrorString(_e), __FILE__, __LINE__);      \
      return ::pcfd::kErrHip;                                                 \
    }                                                                         \
  } while (0)

#define PCFD_LAUNCH_CHECK()                                                   \
  do {                                                                        \
    hipError_t _e = hipGetLastError();                                        \
    if (_e != hipSuccess) {                                                   \
      ::pcfd::set_last_error("kernel launch failed: %s (%s:%d)",              \
                             hipGetErrorString(_e), __FILE__, __LINE__);      \
      return ::pcfd::kErrHip;                                                 \
    }                                                                         \
  } while (0)

// Start/stop events around one kernel launch, for timing a kernel inside a
// real solve; both null for a plain launch.
struct LaunchClock {
  hipEvent_t start = nullptr, stop = nullptr;
};

// hipLaunchKernelGGL, or hipExtLaunchKernelGGL recording the clock's events
// at the kernel's own start and end when the clock is set.
#define PCFD_LAUNCH_TIMED(clk, kernel, grid, block, shm, stream, ...)                     \
  do {                                                                                   \
    if ((clk).start)                                                                     \
      hipExtLaunchKernelGGL(kernel, grid, block, shm, stream, (clk).start, (clk).stop, 0, \
                            __VA_ARGS__);                                                \
    else                                                                                 \
      hipLaunchKernelGGL(kernel, grid, block, shm, stream, __VA_ARGS__);                 \
  } while (0)

// ---------------------------------------------------------------------------
// Tile bookkeeping.  Workgroups are dispatched round-robin over the 8 XCDs;
// remapping the linear workgroup id so that each XCD owns a contiguous run of
// tiles keeps the halo rows two neighbouring tiles share inside one XCD's L2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  const int xcd = bid % kNumXcd;
  const int k = bid / kNumXcd;
  const int q = nblk / kNumXcd;
  const int r = nblk % kNumXcd;
  return (xcd < r) ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

struct Tile2D {
  int tile;   // linear tile number after the XCD remap
  int i;      // x index (slow axis) of this thread's cell
  int j;      // y index (fast axis)
};

// Grid of tiles covering [0, nxr) x [0, nyr); launched as a 1D grid.
__host__ __device__ inline int tiles_y(int nyr) { return (nyr + kTileY - 1) / kTileY; }
__host__ __device__ inline int tiles_x(int nxr) { return (nxr + kTileX - 1) / kTileX; }
inline int num_tiles(int nxr, int nyr) { return tiles_x(nxr) * tiles_y(nyr); }

__device__ __forceinline__ Tile2D tile_coords(int nyr) {
  Tile2D t;
  const int nblk = gridDim.x;
  t.tile = xcd_remap(blockIdx.x, nblk);
  const int ty = tiles_y(nyr);
  const int tx_ = t.tile / ty;
  const int ty_ = t.tile % ty;
  t.i = tx_ * kTileX + (threadIdx.x / kTileY);
  t.j = ty_ * kTileY + (threadIdx.x % kTileY);
  return t;
}

// ---------------------------------------------------------------------------
// Wavefront / block reductions (64-wide wavefronts, 4 waves per block).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, kWave));
  return v;
}

enum RedOp { kSum = 0, kMax = 1 };

// Reduce K values across a 256-thread block; thread 0 receives the result.
// The combination order is fixed (lane tree, then wave 0..3 in order).
template <int K>
__device__ __forceinline__ void block_reduce(double (&v)[K], const int (&ops)[K]) {
  __shared__ double sh[K][kBlock / kWave];
  const int lane = threadIdx.x % kWave;
  const int wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double x = (ops[k] == kSum) ? wave_sum(v[k]) : wave_max(v[k]);
    if (lane == 0) sh[k][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double a = sh[k][0];
      for (int w = 1; w < kBlock / kWave; ++w)
        a = (ops[k] == kSum) ? a + sh[k][w] : fmax(a, sh[k][w]);
      v[k] = a;
    }
  }
  __syncthreads();
}

// Grid-wide deterministic reductions are split over two launches.  The
// producing kernel folds its values inside each workgroup and writes one row
// of K partials per tile with plain stores (block_partials); a one-workgroup
// fold kernel launched right after it (k_fold) sums the rows in tile order
// and runs the reduction's epilogue.  The kernel boundary makes the partials
// visible across XCDs, so no workgroup pays for a device-scope fence.
template <int K>
__device__ __forceinline__ void block_partials(double (&v)[K], const int (&ops)[K],
                                               double* partial, int tile) {
  block_reduce<K>(v, ops);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partial[(size_t)tile * K + k] = v[k];
  }
}

// Fixed-order fold of `ntiles` rows of K partials in one workgroup; thread 0
// receives the totals.
template <int K>
__device__ __forceinline__ void fold_partials(double (&v)[K], const int (&ops)[K],
                                              const double* partial, int ntiles) {
  constexpr int kBatch = 8;  // rows loaded before any is combined: loads overlap
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = (ops[k] == kSum) ? 0.0 : -INFINITY;
  for (int t0 = threadIdx.x; t0 < ntiles; t0 += kBlock * kBatch) {
    double x[kBatch][K];
    // unconditional loads of a clamped row, the identity selected after:
    // no load sits in a branch of its own
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const size_t row = (size_t)min(t0 + u * kBlock, ntiles - 1) * K;
#pragma unroll
      for (int k = 0; k < K; ++k) x[u][k] = partial[row + k];
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (t0 + u * kBlock >= ntiles) x[u][k] = (ops[k] == kSum) ? 0.0 : -INFINITY;
    }
    // Same order as a plain strided loop: rows t0, t0 + kBlock, ...
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = (ops[k] == kSum) ? v[k] + x[u][k] : fmax(v[k], x[u][k]);
    }
  }
  block_reduce<K>(v, ops);
}

// Epi provides: static constexpr int K; ops(int (&)[K]) const; bool active()
// const (false skips the epilogue, e.g. once a solve has finished); and
// operator()(const double (&)[K]) const, run by thread 0 with the totals.
template <class Epi>
__global__ void __launch_bounds__(kBlock) k_fold(const double* __restrict__ partial, int ntiles,
                                                 Epi epi) {
  if (!epi.active()) return;
  constexpr int K = Epi::K;
  int ops[K];
  epi.ops(ops);
  double v[K];
  fold_partials<K>(v, ops, partial, ntiles);
  if (threadIdx.x == 0) epi(v);
}

}  // namespace pcfd
