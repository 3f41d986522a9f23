#include "core/case.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <stdexcept>

namespace wfe {

namespace {

const double kPi = 3.14159265358979323846;

// cos^2 theta' bubble measured in physical z (src/main.cpp:28-54).
void init_bubble(const Case& cs, std::vector<double>& thp) {
  const GDims& g = cs.g;
  const Config& cfg = cs.cfg;
  const double dth = cfg.get_real("bubble_dtheta", 0.0);
  if (dth == 0.0) return;
  const double xc = cfg.get_real("bubble_xc", g.nx * g.dx / 2);
  const double yc = cfg.get_real("bubble_yc", g.ny * g.dy / 2);
  const double zc = cfg.get_real("bubble_zc", 2000.0);
  const double xr = cfg.get_real("bubble_xr", 2000.0);
  const double yr = cfg.get_real("bubble_yr", 2000.0);
  const double zr = cfg.get_real("bubble_zr", 2000.0);
  std::fill(thp.begin(), thp.end(), 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double x = ((double)i + 0.5) * g.dx;
        const double y = ((double)j + 0.5) * g.dy;
        const double z = cs.metric.z_at(g, i, j, cs.metric.zeta_c[k + g.ng]);
        const double rx = (x - xc) / xr, ry = (y - yc) / yr, rz = (z - zc) / zr;
        const double r = std::sqrt(rx * rx + ry * ry + rz * rz);
        if (r < 1.0) {
          const double c = std::cos(kPi * r / 2);
          thp[g.idx(i, j, k)] = dth * c * c;
        }
      }
}

// Base wind as the initial wind (src/main.cpp:185-195); u faces 0..nx.
// The v faces 0..ny carry the additive v_b profile the same way.
void init_wind(const Case& cs, std::vector<double>& u, std::vector<double>& v) {
  const GDims& g = cs.g;
  std::fill(u.begin(), u.end(), 0.0);
  for (int k = 0; k < g.nz; ++k) {
    const double ub = cs.base.ub[k + g.ng];
    if (ub == 0.0) continue;
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i <= g.nx; ++i) u[g.idx(i, j, k)] = ub;
  }
  for (int k = 0; k < g.nz; ++k) {
    const double vb = cs.base.vb[k + g.ng];
    if (vb == 0.0) continue;
    for (int j = 0; j <= g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) v[g.idx(i, j, k)] = vb;
  }
}

// qv = qvbar (src/main.cpp:146-154).
void init_moisture(const Case& cs, std::vector<double>& qv) {
  const GDims& g = cs.g;
  if (!cs.base.has_moisture) return;
  std::fill(qv.begin(), qv.end(), 0.0);
  if (!cs.base.keep_qvb_ic) return;  // qvb_at() reads 0 outside wk82
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const size_t c = g.idx(i, j, k);
        qv[c] = cs.base.qvb[c];
      }
}

// Sharp cos^2 moisture blob (src/main.cpp:156-183); replaces qv.
void init_qv_blob(const Case& cs, std::vector<double>& qv) {
  const GDims& g = cs.g;
  const Config& cfg = cs.cfg;
  const double dq = cfg.get_real("qvbub_dq", 0.0);
  if (dq == 0.0) return;
  const double xc = cfg.get_real("qvbub_xc", g.nx * g.dx / 4);
  const double yc = cfg.get_real("qvbub_yc", g.ny * g.dy / 2);
  const double zc = cfg.get_real("qvbub_zc", 2000.0);
  const double xr = cfg.get_real("qvbub_xr", 2000.0);
  const double yr = cfg.get_real("qvbub_yr", 1e9);
  const double zr = cfg.get_real("qvbub_zr", 2000.0);
  std::fill(qv.begin(), qv.end(), 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double x = ((double)i + 0.5) * g.dx;
        const double y = ((double)j + 0.5) * g.dy;
        const double z = cs.metric.z_at(g, i, j, cs.metric.zeta_c[k + g.ng]);
        const double rx = (x - xc) / xr, ry = (y - yc) / yr, rz = (z - zc) / zr;
        const double r = std::sqrt(rx * rx + ry * ry + rz * rz);
        if (r < 1.0) {
          const double c = std::cos(kPi * r / 2);
          qv[g.idx(i, j, k)] = dq * c * c;
        }
      }
}

// Full input fields -> perturbation state on the C-grid (src/main.cpp:58-100).
void init_from_input(const Case& cs, std::array<std::vector<double>, 8>& ic) {
  const GDims& g = cs.g;
  const InputData& in = cs.input;
  auto src = [&](const std::vector<double>& a, int i, int j, int k) {
    return a[((size_t)k * g.ny + j) * g.nx + i];
  };
  for (auto& f : ic) std::fill(f.begin(), f.end(), 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i <= g.nx; ++i) {
        const int il = i > 0 ? i - 1 : 0, ir = i < g.nx ? i : g.nx - 1;
        ic[0][g.idx(i, j, k)] = 0.5 * (src(in.u, il, j, k) + src(in.u, ir, j, k));
      }
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const int jl = j > 0 ? j - 1 : 0;
        ic[1][g.idx(i, j, k)] = 0.5 * (src(in.v, i, jl, k) + src(in.v, i, j, k));
      }
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const size_t c = g.idx(i, j, k);
        ic[3][c] = src(in.th, i, j, k) - cs.base.thb[c];
        ic[4][c] = src(in.pi, i, j, k) - cs.base.pib[c];
        ic[5][c] = src(in.qv, i, j, k);
      }
}

}  // namespace

std::vector<double> lambert_map_factor(const std::vector<double>& lat_deg, double lat1_deg,
                                       double lat2_deg) {
  const double d2r = 0.01745329252;
  const double phi1 = lat1_deg * d2r, phi2 = lat2_deg * d2r;
  const double PI4 = 0.78539816339;
  const double n = (std::fabs(phi1 - phi2) < 1e-6)
                       ? std::sin(phi1)
                       : std::log(std::cos(phi1) / std::cos(phi2)) /
                             std::log(std::tan(PI4 + phi2 / 2) / std::tan(PI4 + phi1 / 2));
  const double A = std::cos(phi1) * std::pow(std::tan(PI4 + phi1 / 2), n);
  std::vector<double> m(lat_deg.size());
  for (size_t i = 0; i < m.size(); ++i) {
    const double phi = lat_deg[i] * d2r;
    m[i] = A / (std::pow(std::tan(PI4 + phi / 2), n) * std::cos(phi));
  }
  return m;
}

void add_ic_perturbation(const GDims& g, std::vector<double>& thp, double amp, unsigned seed) {
  if (amp <= 0.0) return;
  const size_t n = g.npts();
  std::vector<double> p(n, 0.0), tmp(n, 0.0);
  std::mt19937 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) p[g.idx(i, j, k)] = (double)nd(rng);
  for (int pass = 0; pass < 6; ++pass) {
    tmp = p;
    for (int k = 1; k < g.nz - 1; ++k)
      for (int j = 1; j < g.ny - 1; ++j)
        for (int i = 1; i < g.nx - 1; ++i) {
          const double s6 = tmp[g.idx(i - 1, j, k)] + tmp[g.idx(i + 1, j, k)] +
                            tmp[g.idx(i, j - 1, k)] + tmp[g.idx(i, j + 1, k)] +
                            tmp[g.idx(i, j, k - 1)] + tmp[g.idx(i, j, k + 1)];
          p[g.idx(i, j, k)] = 0.5 * tmp[g.idx(i, j, k)] + s6 / 12.0;
        }
  }
  double ss = 0;
  size_t cnt = 0;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double v = p[g.idx(i, j, k)];
        ss += v * v;
        ++cnt;
      }
  const double rms = std::sqrt(ss / cnt);
  if (rms < 1e-9) return;
  const double sc = amp / rms;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) thp[g.idx(i, j, k)] += sc * p[g.idx(i, j, k)];
}

// ---------------------------------------------------------------- boundary

void BoundaryHost::init(const GDims& g, const DynParams& dp, const InputData& in,
                        const std::vector<double>& thb3, const std::vector<double>& pib3,
                        int bdy_width, double bdy_tau, double nudge_tau) {
  active = true;
  const double inner = nudge_tau > 0 ? 1.0 / nudge_tau : 0.0;
  wgt.assign(g.n2(), 0.0);
  const double pi_half = 1.5707963267948966;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      double wij = inner;
      int dist = 1 << 20;
      if (dp.bc_x_open) dist = std::min({dist, i, g.nx - 1 - i});
      if (dp.bc_y_open) dist = std::min({dist, j, g.ny - 1 - j});
      if (dist < bdy_width) {
        const double s = std::cos(pi_half * (double)dist / (double)bdy_width);
        wij = std::max(wij, s * s / bdy_tau);
      }
      wgt[g.idx2(i, j)] = wij;
    }
  if (!load_into(g, in, thb3, pib3, 0, lo))
    throw std::runtime_error("cannot read the first boundary file");
  if (!load_into(g, in, thb3, pib3, std::min(1, in.n_bdy - 1), hi))
    throw std::runtime_error("cannot read the second boundary file");
  t_lo = 0;
  t_hi = in.bdy_interval;
  idx_hi = 1;
}

bool BoundaryHost::load_into(const GDims& g, const InputData& in,
                             const std::vector<double>& thb3, const std::vector<double>& pib3,
                             int idx, std::array<std::vector<double>, 5>& dst) {
  std::vector<double> f[5];
  if (!in.load_bdy(g, idx, f)) return false;
  auto src = [&](const std::vector<double>& a, int i, int j, int k) {
    return a[((size_t)k * g.ny + j) * g.nx + i];
  };
  for (int v = 0; v < 5; ++v) {
    dst[v].assign(g.npts(), 0.0);
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx + (v == 0 ? 1 : 0); ++i) {
          const size_t c = g.idx(i, j, k);
          double val;
          if (v == 0) {  // u: centres -> faces
            const int il = i > 0 ? i - 1 : 0, ir = i < g.nx ? i : g.nx - 1;
            val = 0.5 * (src(f[0], il, j, k) + src(f[0], ir, j, k));
          } else if (v == 1) {  // v: centres -> faces (j stays below ny)
            const int jl = j > 0 ? j - 1 : 0;
            val = 0.5 * (src(f[1], i, jl, k) + src(f[1], i, j, k));
          } else if (v == 2) {
            val = src(f[2], i, j, k) - thb3[c];
          } else if (v == 3) {
            val = src(f[3], i, j, k) - pib3[c];
          } else {
            val = src(f[4], i, j, k);
          }
          dst[v][c] = val;
        }
  }
  return true;
}

bool BoundaryHost::update(const GDims& g, const InputData& in, const std::vector<double>& thb3,
                          const std::vector<double>& pib3, double t) {
  while (t >= t_hi && idx_hi + 1 < in.n_bdy) {
    std::swap(lo, hi);
    swapped = !swapped;
    ++idx_hi;
    if (!load_into(g, in, thb3, pib3, idx_hi, hi)) return false;
    dirty_hi = true;
    t_lo = t_hi;
    t_hi += in.bdy_interval;
  }
  return true;
}

double BoundaryHost::tfrac(double t) const {
  if (t_hi <= t_lo) return 0;
  const double f = (t - t_lo) / (t_hi - t_lo);
  return f < 0 ? 0 : (f > 1 ? 1.0 : f);
}

// ------------------------------------------------------------------ build

void Case::build() {
  const int nx = cfg.get_int("nx", 100);
  const int ny = cfg.get_int("ny", 100);
  const int nz = cfg.get_int("nz", 50);
  const double dx = cfg.get_real("dx", 200);
  const double dy = cfg.get_real("dy", 200);
  const double dz = cfg.get_real("dz", 200);
  dt = cfg.get_real("dt", 0.25);
  t_end = cfg.get_real("t_end", 1000);
  out_interval = cfg.get_real("out_interval", 100);
  diag_interval = cfg.get_real("diag_interval", 10);
  out_dir = cfg.get_str("out_dir", "out/warm_bubble");
  if (nx < 4 || ny < 1 || nz < 4) throw std::runtime_error("grid too small (nx>=4, ny>=1, nz>=4)");
  g = make_grid(nx, ny, nz, 3, dx, dy, dz);

  dp = DynParams{};
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

  // Acoustic CFL advisory (src/main.cpp:300-311).
  {
    const double cs = 350;
    const double dtau = dt / std::max(1, dp.acoustic_ns);
    const double lim = 0.5 * std::min(dx, dy) / cs;
    if (dtau > lim) {
      char buf[256];
      std::snprintf(buf, sizeof(buf),
                    "WARNING: acoustic substep %.2fs > safe limit %.2fs "
                    "(suggest dt<=%.1fs or acoustic_ns>=%d)",
                    dtau, lim, lim * dp.acoustic_ns, (int)std::ceil(dt / lim));
      warnings.push_back(buf);
    }
  }

  file_mode = cfg.get_str("profile", "isentropic") == "file";
  if (file_mode) {
    if (!input.load(g, cfg.get_str("input_dir", "")))
      throw std::runtime_error("cannot load file-mode input");
    dp.coriolis_use_ub = false;
  }
  std::vector<double> mapf;
  if (file_mode)
    mapf = lambert_map_factor(input.lat, cfg.get_real("proj_lat1", 35),
                              cfg.get_real("proj_lat2", 43));
  metric.build(g, cfg, file_mode ? &input.h : nullptr, file_mode ? &input.fcor : nullptr,
               file_mode ? &mapf : nullptr);

  ProfileTables tables{&input.prof_z, &input.prof_th, &input.prof_qv, &input.prof_u};
  base.build(g, cfg, metric, file_mode ? &tables : nullptr);
  const bool qv_blob = cfg.get_real("qvbub_dq", 0) != 0.0;
  dp.moisture = base.has_moisture || qv_blob;
  if (!dp.moisture) dp.pd_moist = false;
  physics = cfg.get_str("physics", "none");

  for (auto& f : ic) f.assign(g.npts(), 0.0);
  if (file_mode) {
    init_from_input(*this, ic);
    add_ic_perturbation(g, ic[3], cfg.get_real("ic_perturb", 0),
                        (unsigned)cfg.get_int("ens_seed", 0));
    bdy.init(g, dp, input, base.thb, base.pib, cfg.get_int("bdy_width", 8),
             cfg.get_real("bdy_tau", 600), cfg.get_real("nudge_tau", 0));
  } else {
    init_bubble(*this, ic[3]);
    init_wind(*this, ic[0], ic[1]);
    init_moisture(*this, ic[5]);
    init_qv_blob(*this, ic[5]);
  }
}

}  // namespace wfe
