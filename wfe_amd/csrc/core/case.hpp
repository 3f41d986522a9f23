#pragma once

#include <array>
#include <string>
#include <vector>

#include "core/base_state.hpp"
#include "core/config.hpp"
#include "core/grid.hpp"
#include "core/metric.hpp"
#include "dynamics/params.hpp"
#include "io/input.hpp"

namespace wfe {

// Host side of BdyManager (src/dynamics/boundary.hpp:16-38,
// src/dynamics/boundary.cpp:9-81 of the reference): Davies rim weights
// (cos^2 over bdy_width cells, 1/bdy_tau at the edge, optional interior
// nudging 1/nudge_tau) and the relaxation targets for u, v, theta', pi', qv
// read from the boundary files as padded perturbation fields. The pair
// bracketing a context's time and the time fraction are advanced per context
// by wfe_case_update_bdy (boundary.cpp:83-97); lo/hi and t_lo/t_hi here hold
// the first pair, then the pair of the most recently advanced context.
struct BoundaryHost {
  bool active = false;
  std::vector<double> wgt;                    // [NX*NY]
  std::array<std::vector<double>, 5> lo, hi;  // padded 3-D targets
  double t_lo = 0, t_hi = 0;
  int idx_hi = 1;

  void init(const GDims& g, const DynParams& dp, const InputData& in,
            const std::vector<double>& thb3, const std::vector<double>& pib3, int bdy_width,
            double bdy_tau, double nudge_tau);
  // Targets of boundary file idx into dst; false on a file error.
  bool load_into(const GDims& g, const InputData& in, const std::vector<double>& thb3,
                 const std::vector<double>& pib3, int idx,
                 std::array<std::vector<double>, 5>& dst) const;
};

// A fully assembled case: everything src/main.cpp:248-383 builds before the
// time loop, kept on the host in FP64.
struct Case {
  Config cfg;
  GDims g;
  DynParams dp;
  Metric metric;
  BaseState base;
  InputData input;
  bool file_mode = false;
  double dt = 0.25, t_end = 1000, out_interval = 100, diag_interval = 10;
  std::string out_dir;
  std::string physics;  // "none" | "simple" (surface/PBL physics, file mode only)
  // Surface physics initial fields (SfcPBL::init, surface.cu:440-482):
  // padded 2-D arrays with the input's interior and zero ghosts; the hour
  // and day of year of the input's start time (src/main.cpp:364-377).
  struct Surface {
    bool on = false;
    bool nonlocal = false;
    double start_hour = 0;
    int doy = 1;
    std::vector<double> tsk, land, lat, lon, soilw;
  } sfc;
  // Initial state, padded 3-D arrays: u v w thp pip qv qc qr.
  std::array<std::vector<double>, 8> ic;
  BoundaryHost bdy;
  std::vector<std::string> warnings;  // e.g. the acoustic CFL advisory

  // Throws std::runtime_error on configuration or input errors.
  void build();
};

// Idealised initial state into c.ic (src/main.cpp:28-54, 146-195): the
// cos^2 theta' bubble, the base wind on its faces, WK82 vapour, the vapour
// blob.
void idealised_initial_state(Case& c);

// Full-field input (u, v, theta, pi, qv at the cell centres, nz*ny*nx, i
// fastest) as padded perturbation fields on the C-grid: u on the west faces
// 0..nx and v on the south faces 0..ny-1 from the neighbouring centres,
// theta' = theta - thb, pi' = pi - pib, qv (src/main.cpp:58-100; the same
// conversion as the boundary targets, src/dynamics/boundary.cpp:51-81).
std::array<std::vector<double>, 5> perturbation_state(const GDims& g,
                                                      const std::vector<double>& thb,
                                                      const std::vector<double>& pib,
                                                      const std::vector<double>* raw);

// Lambert conformal map factor m(lat) for standard parallels lat1, lat2
// (src/main.cpp:323-340); lat in degrees, interior ny*nx.
std::vector<double> lambert_map_factor(const std::vector<double>& lat_deg, double lat1_deg,
                                       double lat2_deg);

// Seeded theta' perturbation exactly as src/main.cpp:105-144 (mt19937 +
// normal_distribution<float>, six 3-D box-smoothing passes, RMS scaled to
// amp) added to thp (padded).
void add_ic_perturbation(const GDims& g, std::vector<double>& thp, double amp, unsigned seed);

}  // namespace wfe
