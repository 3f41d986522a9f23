// Host side of the lateral boundary forcing (BdyManager,
// src/dynamics/boundary.cpp:9-81): Davies rim weights and the target states
// read from the boundary files. The per-context bracket and time fraction
// live with the case handle (wfe_case_update_bdy).
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "core/case.hpp"

namespace wfe {

namespace {

// Relaxation rate [1/s] of each interior column: a cos^2 ramp from 1/tau at
// the open edges over `width` cells, never below the interior nudging rate
// 1/nudge_tau (0 = none). Periodic directions have no rim.
std::vector<double> davies_weights(const GDims& g, const DynParams& dp, int width, double tau,
                                   double nudge_tau) {
  const double floor_rate = nudge_tau > 0 ? 1.0 / nudge_tau : 0.0;
  constexpr double kHalfPi = 1.5707963267948966;
  std::vector<double> w(g.n2(), 0.0);
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      int edge = 1 << 20;  // cells to the nearest open edge
      if (dp.bc_x_open) edge = std::min({edge, i, g.nx - 1 - i});
      if (dp.bc_y_open) edge = std::min({edge, j, g.ny - 1 - j});
      double rate = floor_rate;
      if (edge < width) {
        const double s = std::cos(kHalfPi * (double)edge / (double)width);
        rate = std::max(rate, s * s / tau);
      }
      w[g.idx2(i, j)] = rate;
    }
  return w;
}

}  // namespace

bool BoundaryHost::load_into(const GDims& g, const InputData& in,
                             const std::vector<double>& thb3, const std::vector<double>& pib3,
                             int idx, std::array<std::vector<double>, 5>& dst) const {
  std::vector<double> raw[5];
  if (!in.load_bdy(g, idx, raw)) return false;
  dst = perturbation_state(g, thb3, pib3, raw);
  return true;
}

void BoundaryHost::init(const GDims& g, const DynParams& dp, const InputData& in,
                        const std::vector<double>& thb3, const std::vector<double>& pib3,
                        int bdy_width, double bdy_tau, double nudge_tau) {
  active = true;
  wgt = davies_weights(g, dp, bdy_width, bdy_tau, nudge_tau);
  // the targets at the first two boundary times (the same file twice when
  // there is only one)
  if (!load_into(g, in, thb3, pib3, 0, lo))
    throw std::runtime_error("cannot read the first boundary file");
  if (!load_into(g, in, thb3, pib3, std::min(1, in.n_bdy - 1), hi))
    throw std::runtime_error("cannot read the second boundary file");
  t_lo = 0;
  t_hi = in.bdy_interval;
  idx_hi = 1;
}

}  // namespace wfe
