// Assembly of a case before the time loop, in the order of
// src/main.cpp:261-383: grid and run keys, dynamics parameters, file-mode
// input with its Lambert map factors, the metric, the base state, then the
// initial state and (file mode) the boundary forcing.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#include "core/case.hpp"

namespace wfe {

namespace {

DynParams read_params(const Config& cfg) {
  DynParams dp;
  dp.diff_K = cfg.get_real("diff_K", 0);
  dp.coriolis_f = cfg.get_real("coriolis_f", 0);
  dp.rayleigh_zd = cfg.get_real("rayleigh_zd", -1);
  dp.rayleigh_alpha = cfg.get_real("rayleigh_alpha", 0);
  dp.acoustic_ns = cfg.get_int("acoustic_ns", 6);
  dp.acoustic_beta = cfg.get_real("acoustic_beta", 0.2);
  dp.acoustic_smdiv = cfg.get_real("acoustic_smdiv", 0.1);
  dp.bc_x_open = cfg.get_str("bc_x", "periodic") == "open";
  dp.bc_y_open = cfg.get_str("bc_y", "periodic") == "open";
  dp.cstar = cfg.get_real("cstar", 30);
  dp.w_damping = cfg.get_str("w_damping", "off") == "on";
  dp.w_abort = cfg.get_real("w_abort", 150);
  dp.pd_moist = cfg.get_str("pd_moist", "off") == "on";
  dp.microphysics = cfg.get_str("microphysics", "on") == "on";
  return dp;
}

// Advisory when the horizontal acoustic Courant number of a substep with a
// 350 m/s sound speed exceeds 1/2 (src/main.cpp:300-311).
void acoustic_advisory(double dt, const DynParams& dp, double dx, double dy,
                       std::vector<std::string>& warnings) {
  const double cs = 350;
  const double dtau = dt / std::max(1, dp.acoustic_ns);
  const double lim = 0.5 * std::min(dx, dy) / cs;
  if (dtau <= lim) return;
  char buf[256];
  std::snprintf(buf, sizeof(buf),
                "WARNING: acoustic substep %.2fs > safe limit %.2fs "
                "(suggest dt<=%.1fs or acoustic_ns>=%d)",
                dtau, lim, lim * dp.acoustic_ns, (int)std::ceil(dt / lim));
  warnings.push_back(buf);
}

}  // namespace

void Case::build() {
  const int nx = cfg.get_int("nx", 100), ny = cfg.get_int("ny", 100), nz = cfg.get_int("nz", 50);
  const double dx = cfg.get_real("dx", 200), dy = cfg.get_real("dy", 200),
               dz = cfg.get_real("dz", 200);
  dt = cfg.get_real("dt", 0.25);
  t_end = cfg.get_real("t_end", 1000);
  out_interval = cfg.get_real("out_interval", 100);
  diag_interval = cfg.get_real("diag_interval", 10);
  out_dir = cfg.get_str("out_dir", "out/warm_bubble");
  if (nx < 4 || ny < 1 || nz < 4) throw std::runtime_error("grid too small (nx>=4, ny>=1, nz>=4)");
  g = make_grid(nx, ny, nz, 3, dx, dy, dz);
  dp = read_params(cfg);
  acoustic_advisory(dt, dp, dx, dy, warnings);

  file_mode = cfg.get_str("profile", "isentropic") == "file";
  std::vector<double> mapf;
  if (file_mode) {
    if (!input.load(g, cfg.get_str("input_dir", "")))
      throw std::runtime_error("cannot load file-mode input");
    dp.coriolis_use_ub = false;  // real data: Coriolis acts on the full wind
    mapf = lambert_map_factor(input.lat, cfg.get_real("proj_lat1", 35),
                              cfg.get_real("proj_lat2", 43));
  }
  metric.build(g, cfg, file_mode ? &input.h : nullptr, file_mode ? &input.fcor : nullptr,
               file_mode ? &mapf : nullptr);
  const ProfileTables tables{&input.prof_z, &input.prof_th, &input.prof_qv, &input.prof_u};
  base.build(g, cfg, metric, file_mode ? &tables : nullptr);
  dp.moisture = base.has_moisture || cfg.get_real("qvbub_dq", 0) != 0.0;
  if (!dp.moisture) dp.pd_moist = false;
  physics = cfg.get_str("physics", "none");

  for (auto& f : ic) f.assign(g.npts(), 0.0);
  if (!file_mode) {
    idealised_initial_state(*this);
    return;
  }
  // file mode: the input state as perturbations on the C-grid, the seeded
  // ensemble perturbation, and the boundary forcing
  const std::vector<double> raw[5] = {input.u, input.v, input.th, input.pi, input.qv};
  std::array<std::vector<double>, 5> p = perturbation_state(g, base.thb, base.pib, raw);
  ic[0].swap(p[0]);
  ic[1].swap(p[1]);
  ic[3].swap(p[2]);
  ic[4].swap(p[3]);
  ic[5].swap(p[4]);
  add_ic_perturbation(g, ic[3], cfg.get_real("ic_perturb", 0),
                      (unsigned)cfg.get_int("ens_seed", 0));
  bdy.init(g, dp, input, base.thb, base.pib, cfg.get_int("bdy_width", 8),
           cfg.get_real("bdy_tau", 600), cfg.get_real("nudge_tau", 0));
  if (physics == "simple") {
    // start = YYYYMMDDHH: day of year (no leap day) and UTC hour
    const std::string& st = input.start;
    if (st.size() < 10) throw std::runtime_error("physics=simple needs the input start time");
    const int mo = std::atoi(st.substr(4, 2).c_str()), day = std::atoi(st.substr(6, 2).c_str());
    if (mo < 1 || mo > 12) throw std::runtime_error("input start time has no valid month");
    static const int kCumDays[12] = {0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334};
    sfc.on = true;
    sfc.doy = kCumDays[mo - 1] + day;
    sfc.start_hour = std::atoi(st.substr(8, 2).c_str());
    sfc.nonlocal = cfg.get_str("pbl", "local") == "nonlocal";
    auto padded = [&](const std::vector<double>& src) {
      std::vector<double> h(g.n2(), 0.0);
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) h[g.idx2(i, j)] = src[(size_t)j * g.nx + i];
      return h;
    };
    sfc.tsk = padded(input.tsk);
    sfc.land = padded(input.land);
    sfc.lat = padded(input.lat);
    sfc.lon = padded(input.lon);
    sfc.soilw = padded(input.soilw);
  }
}

}  // namespace wfe
