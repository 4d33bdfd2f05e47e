// Fused multigrid level kernels: one launch per level and direction.
//
// A V-cycle level does, going down, `nu` damped-Jacobi sweeps from zero (or
// from a given iterate), the residual, and the restriction P^T r; going up,
// the prolongation x += P x_c and `nu` sweeps.  Each of those is a stencil
// pass, so a level costs 2*nu + 2 grid-wide dependencies when every pass is
// its own launch.  Here each workgroup instead owns a rectangular tile of the
// level, loads the tile plus a halo of (nu + 1) stencil radii into LDS once,
// and runs all passes of one direction in LDS, recomputing the halo rings it
// needs (each pass shrinks the valid region by one radius).  The halo work is
// redundant but cheap next to the HBM traffic and the launch latency it
// replaces.  Results are the same arithmetic as the pass-by-pass kernels.
#pragma once

#include <vector>

#include "poisson.hpp"
#include "vcycle_tile.hpp"

namespace pcfd {

// A stored coarse level in the form the fused kernels take.  `ent_off` is the
// entry's neighbour offset inside this level's LDS region (di * ry + dj).
template <typename T>
struct ClassRows {
  int nx, ny, px, py;
  const uint16_t* cls;
  const int* row_ptr;
  const int* ent_off;
  const T* ent_val;
  const T* cdinv;
  const uint8_t* cactive;
  int hot;    // the level's most frequent class (interior rows), staged in LDS; -1: none
  int hot_n;  // its entry count
  const T* rdinv;        // per cell 1/diag
  const uint8_t* ract;   // per cell active flag
  const int2* rr;        // per cell entry range of its class row
  int hot_e0;            // first entry of the hot class
};

// Largest row the fused class kernels stage in LDS as the level's hot class.
constexpr int kHotMax = 256;

// A stored level in streaming form: the row dictionary with every class's
// row laid out as a dense (2R+1) x (2R+1) box around the diagonal (zeros
// where the row has no entry), plus one all-zero class `zero_class` for
// cells outside the grid.
template <typename T>
struct BoxRows {
  int nx, ny, px, py;
  const uint16_t* cls;
  const T* box;
  const T* cdinv;  // zero_class + 1 entries
  const uint8_t* cactive;
  int zero_class;
};

// Largest stencil radius the streaming stored-level legs take.
constexpr int kBoxMaxRadius = 3;

// Level-0 (matrix-free 5-point) down leg.  Right-hand side either from the
// Krylov residual r (b0 = r - mean_r on solvable cells, written to b_out) or
// given in b_in; iterate from zero, or from x_in minus a pending mean shift.
// Writes the pre-smoothed iterate (x_out) and the restricted residual (bc).
template <typename T>
int launch_fine_down(hipStream_t s, const FineOp& A, const FusedTile& ft, T w, int nu,
                     int nonzero, const double* r, const T* b_in, const T* x_in, int shift_in,
                     T* b_out, T* x_out, T* bc, int fx, int fy, int cnx, int cny,
                     const SolveState* st, const LaunchClock& clk = LaunchClock{});

// Level-0 up leg: z = S^nu (x_pre + P x_c); `reduce` selects the partial sums
// of the cycle output written per tile (SmoothReduce).
template <typename T>
int launch_fine_up(hipStream_t s, const FineOp& A, const FusedTile& ft, T w, int nu,
                   const T* x_pre, const T* b0, const T* xc, int fx, int fy, int cny, T* z,
                   int reduce, double* partial, const SolveState* st,
                   const LaunchClock& clk = LaunchClock{});

// Level-0 legs in streaming form (uniform mesh, one sweep, zero start): one
// wavefront per strip of 60 output columns and `rows` rows (a multiple of 4);
// the up leg writes one row of partials per workgroup (stream_blocks).
bool stream_supported(const FineOp& A, int nu, int fx, int fy);
int stream_blocks(int nx, int ny, int rows);
// The multigrid-CG update's partials (r.r, sum r) of the previous iteration,
// left for the next streaming down leg to fold (rows = 0: the state is
// current).  `iter` is the iteration count once that update is counted.
struct UpdateFold {
  const double* partial = nullptr;
  int rows = 0;
  int iter = 0;
  int maxiter = 0;
  int n_proj = 0;
};

template <typename T>
int launch_fine_down_stream(hipStream_t s, const FineOp& A, int rows, T w, const double* r,
                            T* b_out, T* x_out, T* bc, int fx, int fy, int cnx, int cny,
                            SolveState* st, const UpdateFold& uf = UpdateFold{},
                            const LaunchClock& clk = LaunchClock{});
template <typename T>
int launch_fine_up_stream(hipStream_t s, const FineOp& A, int rows, T w, const T* x_pre,
                          const T* b0, const T* xc, int fx, int fy, int cny, T* z, int reduce,
                          double* partial, const SolveState* st,
                          const LaunchClock& clk = LaunchClock{});

// Stored-level legs in streaming form (one sweep each way, zero start), for
// stencils of radius R <= kBoxMaxRadius: one wavefront per strip of
// 64 - 4R output columns and `rows` rows.
int box_strips(int ny, int radius);
int box_blocks(int nx, int ny, int radius, int rows);
template <typename T>
int launch_class_down_stream(hipStream_t s, const BoxRows<T>& L, int radius, int rows, T w,
                             const T* b, T* x_out, T* bc, int fx, int fy, int cnx, int cny,
                             const SolveState* st, const LaunchClock& clk = LaunchClock{});
template <typename T>
int launch_class_up_stream(hipStream_t s, const BoxRows<T>& L, int radius, int rows, T w,
                           const T* x_pre, const T* b, const T* xc, int fx, int fy, int cny,
                           T* x_out, const SolveState* st, const LaunchClock& clk = LaunchClock{});

template <typename T>
int launch_class_down(hipStream_t s, const ClassRows<T>& L, const FusedTile& ft, int rxs,
                      int rys, T w, int nu, const T* b, T* x_out, T* bc, int fx, int fy, int cnx,
                      int cny, const SolveState* st, const LaunchClock& clk = LaunchClock{});

template <typename T>
int launch_class_up(hipStream_t s, const ClassRows<T>& L, const FusedTile& ft, int rxs, int rys,
                    T w, int nu, const T* x_pre, const T* b, const T* xc, int fx, int fy, int cny,
                    T* x_out, const SolveState* st, const LaunchClock& clk = LaunchClock{});

// The V-cycle's output sums (rows of sum z, sum b0 z, sum b0) left for the
// direction kernel to fold; rows = 0: mean_z and beta are already in the state.
struct CycleFold {
  const double* partial = nullptr;
  int rows = 0;
  int n_proj = 0;
};

// Multigrid-CG direction and operator product on level 0 (halo-1 tiling):
// p_new = (z - mean_z) + beta p_old, q = A p_new, per-tile partials of p.q.
template <typename T>
int launch_fine_dir(hipStream_t s, const FineOp& A, const FusedTile& ft, const T* z,
                    const double* p_old, double* p_new, double* q, double* partial,
                    SolveState* st, const CycleFold& cf = CycleFold{},
                    const LaunchClock& clk = LaunchClock{});

// x = C b with C an n x n dense FP64 matrix, spread over many workgroups.
template <typename T>
int launch_dense_apply(hipStream_t s, const double* C, int n, const T* b, T* x,
                       const SolveState* st, const LaunchClock& clk = LaunchClock{});

// Tiles (power-of-two sides, multiples of the aggregation factors) whose LDS
// region fits `lds_cap` bytes; `force_*` > 0 pins a side.
std::vector<FusedTile> tile_candidates(int nx, int ny, int rxs, int rys, int halo_mult, int fx,
                                       int fy, unsigned bytes_per_cell, unsigned fixed_bytes,
                                       unsigned lds_cap, int force_tx, int force_ty);

// Static cost estimate of a tiling (used when tiles are not measured).
double tile_cost(const FusedTile& t);

// The candidate of least static cost; false when none fits.
bool choose_tile(int nx, int ny, int rxs, int rys, int halo_mult, int fx, int fy,
                 unsigned bytes_per_cell, unsigned fixed_bytes, unsigned lds_cap, int force_tx,
                 int force_ty, FusedTile* out);

}  // namespace pcfd
