// Tiling descriptor of the fused multigrid level kernels.
#pragma once

namespace pcfd {

// Tiling of one fused level kernel.
struct FusedTile {
  int tx = 0, ty = 0;    // tile size in cells along x (slow) and y (fast)
  int hx = 0, hy = 0;    // halo: (nu + 1 [+ 1 for a nonzero start]) * stencil radius
  int rx = 0, ry = 0;    // LDS region = tile + 2 * halo
  int ntx = 0, nty = 0;  // tiles per axis
  unsigned lds = 0;      // dynamic LDS bytes of the launch
  int ntiles() const { return ntx * nty; }
};

}  // namespace pcfd
