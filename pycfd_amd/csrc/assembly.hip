// The fine Poisson operator in compressed rows, assembled on the host.
//
// pycfd/core/pressure.py:110-270: the 5-point div(grad) over the faces the
// projection corrects.  A face couples two fluid cells (a wall or a body
// face couples nothing); each coupled face adds its coefficient off the
// diagonal and subtracts it from the diagonal; a wall with a pressure value
// (Dirichlet) subtracts twice its coefficient from the diagonal of the cells
// along it and moves twice coefficient x value to the right-hand side.
// Cells that end with a zero diagonal (fluid cells walled in on every side)
// and solid cells get the identity row.  The diagonal accumulates the four
// faces in the order left, right, bottom, top, as the reference does, so the
// values agree bit for bit.

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/pycfd_hip.h"
#include "common.hpp"

namespace pcfd {
namespace {

struct Entry {
  int64_t col;
  double val;
};

// Row k of the operator: up to five entries, columns ascending (a periodic
// axis of one or two cells makes a face reach the cell itself or the same
// neighbour twice: those entries are summed), its boundary value and whether
// the cell is solvable.
int assemble_row(const pcfd_poisson_assembly& a, int64_t k, Entry* out, double& bc_out,
                 bool& solvable) {
  const int nx = a.nx, ny = a.ny;
  const int64_t i = k / ny, j = k - i * ny;
  auto fluid = [&](int64_t c) { return !a.solid || !a.solid[c]; };
  const bool f = fluid(k);
  double diag = 0.0, bc = 0.0;
  Entry e[5];
  int m = 0;
  for (int w = 0; w < 4; ++w) {  // left, right, bottom, top
    const bool xaxis = w < 2;
    const int shift = (w & 1) ? 1 : -1;
    const bool periodic = xaxis ? a.periodic_x != 0 : a.periodic_y != 0;
    int64_t ii = i, jj = j;
    bool inside = true;
    if (xaxis) {
      ii = i + shift;
      if (ii < 0 || ii >= nx) {
        inside = periodic;
        ii = (ii + nx) % nx;
      }
    } else {
      jj = j + shift;
      if (jj < 0 || jj >= ny) {
        inside = periodic;
        jj = (jj + ny) % ny;
      }
    }
    double coeff;
    if (xaxis)
      coeff = a.cx_face ? a.cx_face[(i + (shift > 0 ? 1 : 0)) * ny + j] : a.cx;
    else
      coeff = a.cy_face ? a.cy_face[i * (ny + 1) + j + (shift > 0 ? 1 : 0)] : a.cy;
    const int64_t nb = ii * ny + jj;
    const bool active = f && inside && fluid(nb);
    if (active) e[m++] = Entry{nb, coeff};
    diag -= coeff * (active ? 1.0 : 0.0);
    if (a.dirichlet[w]) {
      const double on_wall = (f && !inside) ? 1.0 : 0.0;
      diag -= 2.0 * coeff * on_wall;
      bc -= 2.0 * coeff * a.p_ref[w] * on_wall;
    }
  }
  solvable = f && !(diag == 0.0);
  e[m++] = Entry{k, solvable ? diag : 1.0};
  bc_out = solvable ? bc : 0.0;
  for (int q = 1; q < m; ++q) {  // insertion sort: stable, no allocation
    const Entry x = e[q];
    int p = q;
    for (; p > 0 && e[p - 1].col > x.col; --p) e[p] = e[p - 1];
    e[p] = x;
  }
  int r = 0;
  for (int q = 0; q < m; ++q) {
    if (r > 0 && out[r - 1].col == e[q].col)
      out[r - 1].val += e[q].val;
    else
      out[r++] = e[q];
  }
  return r;
}

template <class F>
void parallel_rows(int64_t n, F&& f) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int T = n < (1 << 16) ? 1 : (int)std::min(16u, hw);
  if (T == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + T - 1) / T;
  for (int t = 0; t < T; ++t) {
    const int64_t b0 = std::min<int64_t>(n, t * chunk), b1 = std::min<int64_t>(n, b0 + chunk);
    th.emplace_back([&f, b0, b1] { f(b0, b1); });
  }
  for (auto& x : th) x.join();
}

}  // namespace
}  // namespace pcfd

extern "C" {

int pcfd_poisson_assemble(const pcfd_poisson_assembly* a, int64_t* indptr, int32_t* indices,
                          double* data, double* bc, uint8_t* solvable) {
  using namespace pcfd;
  if (!a || !indptr || a->nx < 1 || a->ny < 1 || (indices && (!data || !bc || !solvable))) {
    set_last_error("pcfd_poisson_assemble: bad argument");
    return kErrBadArg;
  }
  const int64_t n = (int64_t)a->nx * a->ny;
  if (n >= ((int64_t)1 << 31)) {
    set_last_error("pcfd_poisson_assemble: %lld cells exceed 32-bit columns", (long long)n);
    return kErrUnsupported;
  }
  if (!indices) {  // row lengths
    indptr[0] = 0;
    parallel_rows(n, [&](int64_t b0, int64_t b1) {
      Entry row[5];
      double b;
      bool s;
      for (int64_t k = b0; k < b1; ++k) indptr[k + 1] = assemble_row(*a, k, row, b, s);
    });
    for (int64_t k = 0; k < n; ++k) indptr[k + 1] += indptr[k];
    return kOk;
  }
  parallel_rows(n, [&](int64_t b0, int64_t b1) {
    Entry row[5];
    for (int64_t k = b0; k < b1; ++k) {
      bool s;
      const int m = assemble_row(*a, k, row, bc[k], s);
      solvable[k] = s ? 1 : 0;
      const int64_t o = indptr[k];
      for (int q = 0; q < m; ++q) {
        indices[o + q] = (int32_t)row[q].col;
        data[o + q] = row[q].val;
      }
    }
  });
  return kOk;
}

}  // extern "C"
