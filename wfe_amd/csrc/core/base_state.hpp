#pragma once

#include <vector>

#include "core/config.hpp"
#include "core/grid.hpp"
#include "core/metric.hpp"

namespace wfe {

// Profile tables handed over by file mode (area-mean z, theta, qv, u).
struct ProfileTables {
  const std::vector<double>* z = nullptr;
  const std::vector<double>* th = nullptr;
  const std::vector<double>* qv = nullptr;
  const std::vector<double>* u = nullptr;
};

// Hydrostatic, horizontally homogeneous (in physical z) base state, built on
// the host in FP64 and evaluated at every padded grid point through z(i,j,k)
// (src/core/base_state.hpp:16-61, src/core/base_state.cpp:21-193 of the
// reference). profile = isentropic | constant_N | wk82 | file.
//
// Additive keys (absent => the reference's behaviour, bit for bit):
//   rh_max        caps the WK82 relative-humidity profile (default 1);
//   wind_v        constant base v (default 0);
//   wind_profile = quarter_circle: u_b = U_s(1 - cos a) + wind_offset,
//                  v_b = U_s sin a + wind_offset_v, a = (pi/2) min(z/wind_zs, 1).
struct BaseState {
  // 3-D fields, GDims::idx layout, npts entries each.
  std::vector<double> thb, pib, rhob, dthbdz, thbw, rhobw, thvb, thvbw, qvb;
  std::vector<double> ub, vb;  // 1-D base wind [NZ], index k + ng
  bool has_moisture = false;
  bool keep_qvb_ic = false;  // qv initial condition = qvb (wk82 only)

  // Throws std::runtime_error on an unknown profile or missing tables.
  void build(const GDims& g, const Config& cfg, const Metric& metric,
             const ProfileTables* tables = nullptr);
};

}  // namespace wfe
