#pragma once

#include <cstddef>

namespace wfe {

// Arakawa C-grid in one padded buffer shape shared by every 3-D field
// (src/core/grid.hpp:22-36 of the reference): i fastest, ng ghost cells on
// each side, and one extra z level so that w (faces k = 0..nz) fits the same
// shape as the cell-centred scalars. u(i) is the west face of cell i, v(j)
// the south face, w(k) the bottom face.
struct GDims {
  int nx = 0, ny = 0, nz = 0;  // interior cells
  int ng = 3;                  // ghost width (5th-order stencil)
  int NX = 0, NY = 0, NZ = 0;  // padded extents
  double dx = 0, dy = 0, dz = 0;

  size_t idx(int i, int j, int k) const {
    return ((size_t)(k + ng) * NY + (j + ng)) * NX + (i + ng);
  }
  size_t idx2(int i, int j) const { return (size_t)(j + ng) * NX + (i + ng); }
  size_t npts() const { return (size_t)NX * NY * NZ; }
  size_t n2() const { return (size_t)NX * NY; }
};

inline GDims make_grid(int nx, int ny, int nz, int ng, double dx, double dy, double dz) {
  GDims g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.ng = ng;
  g.NX = nx + 2 * ng;
  g.NY = ny + 2 * ng;
  g.NZ = nz + 1 + 2 * ng;
  g.dx = dx;
  g.dy = dy;
  g.dz = dz;
  return g;
}

// Column solvers keep at most this many levels (nz + 1 <= limit), as the
// reference does (src/core/grid.hpp:38-39).
inline constexpr int kMaxColumnLevels = 320;

}  // namespace wfe
