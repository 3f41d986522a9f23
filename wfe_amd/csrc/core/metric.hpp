#pragma once

#include <vector>

#include "core/config.hpp"
#include "core/grid.hpp"

namespace wfe {

// Gal-Chen terrain-following metric on the host, in FP64
// (src/core/metric.hpp:11-61, src/core/metric.cpp:12-168 of the reference):
//   zeta in [0, zt],   z(i,j,zeta) = h + zeta (zt - h) / zt
//   J = dz/dzeta = (zt - h) / zt   (constant within a column)
//   dz/dx|zeta = h_x (1 - zeta / zt)
// Vertical arrays have NZ entries indexed by k + ng; horizontal arrays are
// padded 2-D fields indexed by GDims::idx2.
//
// Keys: stretch = none | geometric (dz0, dz_ratio, dz_max);
//       terrain = none | schaer | gauss | file (mtn_h0, mtn_a, mtn_lambda,
//       mtn_xc, mtn_yc, mtn_ya); coriolis_f (constant f-plane).
// With terrain = file the interior h (ny*nx) comes from the input file,
// and f(lat), the map factor m likewise (edge-clamped into the ghosts).
struct Metric {
  double zt = 0;
  std::vector<double> zeta_w, zeta_c, dzeta_c, dzeta_w;          // [NZ]
  std::vector<double> h, hx, hy, hx_u, hy_v, jac, fcor, mapf;    // [NX*NY]

  // Throws std::runtime_error on an inconsistent configuration.
  void build(const GDims& g, const Config& cfg, const std::vector<double>* h_file = nullptr,
             const std::vector<double>* fcor_file = nullptr,
             const std::vector<double>* mapf_file = nullptr);

  double z_at(const GDims& g, int i, int j, double zeta) const {
    double hh = h[g.idx2(i, j)];
    return hh + zeta * (zt - hh) / zt;
  }
};

}  // namespace wfe
