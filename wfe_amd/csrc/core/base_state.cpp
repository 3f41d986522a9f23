#include "core/base_state.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "core/constants.hpp"

namespace wfe {

namespace {
// Tetens saturation mixing ratio without the denominator floor, as used by the
// base-state sounding (src/core/base_state.cpp:14-18).
double qsat_sounding(double p, double T) {
  const double es = 610.78 * std::exp(17.269 * (T - 273.16) / (T - 35.86));
  return 0.622 * es / (p - es);
}
}  // namespace

// Follows src/core/base_state.cpp:21-193: theta(z) per profile, a 1 m table
// of Exner pi integrated hydrostatically from pi(0) = 1 with the virtual
// potential temperature, five moist iterations for WK82, then evaluation of
// every field at the padded grid points.
void BaseState::build(const GDims& g, const Config& cfg, const Metric& metric,
                      const ProfileTables* tables) {
  const std::string prof = cfg.get_str("profile", "isentropic");
  const double theta0 = cfg.get_real("theta0", 300);
  const double bvN = cfg.get_real("bv_N", 0.01);
  const bool constN = prof == "constant_N";
  const bool wk82 = prof == "wk82";
  const bool from_file = prof == "file";
  if (from_file && (!tables || !tables->z))
    throw std::runtime_error("profile=file but no profile tables were provided");
  has_moisture = wk82 || from_file;
  if (!constN && !wk82 && !from_file && prof != "isentropic")
    throw std::runtime_error("unknown profile: " + prof);
  const double rh_max = cfg.has("rh_max") ? cfg.get_real("rh_max", 1.0) : 1.0;

  auto table_interp = [&](const std::vector<double>& tab, double z) -> double {
    const std::vector<double>& zt = *tables->z;
    const double dzt = zt[1] - zt[0];
    const double s = (z - zt[0]) / dzt;
    int it = (int)s;
    if (it < 0) it = 0;
    if (it > (int)zt.size() - 2) it = (int)zt.size() - 2;
    const double f = s - it;
    return tab[it] * (1 - f) + tab[it + 1] * f;
  };

  // WK82 sounding constants.
  const double z_tr = 12000.0, th_tr = 343.0, T_tr = 213.0, qv_max = 0.014;
  auto theta_of_z = [&](double z) -> double {
    if (from_file) return table_interp(*tables->th, z);
    if (wk82) {
      if (z <= z_tr) return 300.0 + (th_tr - 300.0) * std::pow(z / z_tr, 1.25);
      return th_tr * std::exp(phys::grav * (z - z_tr) / (phys::cp * T_tr));
    }
    if (constN) return theta0 * std::exp(bvN * bvN * z / phys::grav);
    return theta0;
  };
  auto relhum_of_z = [&](double z) -> double {
    const double rh = (z <= z_tr) ? 1.0 - 0.75 * std::pow(z / z_tr, 1.25) : 0.25;
    return std::min(rh, rh_max);
  };

  // 1 m table spanning the padded column with margin.
  const double zmin = -(double)(g.ng + 1) * g.dz * 4 - 100.0;
  const double zmax = metric.zt + (double)(g.ng + 1) * g.dz * 4 + 100.0;
  const double dzt = 1.0;
  const int ntab = (int)((zmax - zmin) / dzt) + 2;
  const int i0 = (int)(-zmin / dzt);
  std::vector<double> pitab(ntab), qvtab(ntab, 0.0), thvtab(ntab);
  for (int it = 0; it < ntab; ++it) thvtab[it] = theta_of_z(zmin + it * dzt);

  auto integrate_pi = [&]() {
    pitab[i0] = 1.0;
    for (int it = i0 + 1; it < ntab; ++it) {
      const double thv = 0.5 * (thvtab[it - 1] + thvtab[it]);
      pitab[it] = pitab[it - 1] - phys::grav * dzt / (phys::cp * thv);
    }
    for (int it = i0 - 1; it >= 0; --it) {
      const double thv = 0.5 * (thvtab[it] + thvtab[it + 1]);
      pitab[it] = pitab[it + 1] + phys::grav * dzt / (phys::cp * thv);
    }
  };
  integrate_pi();

  if (wk82) {
    for (int iter = 0; iter < 5; ++iter) {
      for (int it = 0; it < ntab; ++it) {
        const double z = zmin + it * dzt;
        const double th = theta_of_z(z);
        const double pi = pitab[it];
        const double T = th * pi;
        const double p = phys::p00 * std::pow(pi, phys::cp / phys::Rd);
        double qv = std::min(relhum_of_z(z) * qsat_sounding(p, T), qv_max);
        // Below ground the table keeps the previous iterate (starts at 0),
        // exactly as the reference's index expression does.
        if (z < 0) qv = qvtab[it > 0 ? it : 0];
        qvtab[it] = qv;
        thvtab[it] = th * (1.0 + phys::eps61 * qv);
      }
      integrate_pi();
    }
  } else if (from_file) {
    for (int it = 0; it < ntab; ++it) {
      const double z = zmin + it * dzt;
      qvtab[it] = std::max(0.0, table_interp(*tables->qv, z));
      thvtab[it] = theta_of_z(z) * (1.0 + phys::eps61 * qvtab[it]);
    }
    integrate_pi();
  }

  auto interp = [&](const std::vector<double>& tab, double z) -> double {
    const double s = (z - zmin) / dzt;
    int it = (int)s;
    if (it < 0) it = 0;
    if (it > ntab - 2) it = ntab - 2;
    const double f = s - it;
    return tab[it] * (1 - f) + tab[it + 1] * f;
  };
  auto rho_of = [&](double pi, double thv) -> double {
    return phys::p00 * std::pow(pi, phys::cv / phys::Rd) / (phys::Rd * thv);
  };

  const size_t n = g.npts();
  thb.assign(n, 0.0);
  pib.assign(n, 0.0);
  rhob.assign(n, 0.0);
  dthbdz.assign(n, 0.0);
  thbw.assign(n, 0.0);
  rhobw.assign(n, 0.0);
  thvb.assign(n, 0.0);
  thvbw.assign(n, 0.0);
  qvb.assign(n, 0.0);
  for (int k = -g.ng; k < g.nz + 1 + g.ng; ++k)
    for (int j = -g.ng; j < g.ny + g.ng; ++j)
      for (int i = -g.ng; i < g.nx + g.ng; ++i) {
        const size_t c = g.idx(i, j, k);
        const double zc = metric.z_at(g, i, j, metric.zeta_c[k + g.ng]);
        const double zw = metric.z_at(g, i, j, metric.zeta_w[k + g.ng]);
        const double thc = theta_of_z(zc), thw = theta_of_z(zw);
        const double qvc = has_moisture ? interp(qvtab, zc) : 0.0;
        const double qvw = has_moisture ? interp(qvtab, zw) : 0.0;
        const double thvc = thc * (1.0 + phys::eps61 * qvc);
        const double thvw = thw * (1.0 + phys::eps61 * qvw);
        const double pic = interp(pitab, zc), piw = interp(pitab, zw);
        thb[c] = thc;
        pib[c] = pic;
        rhob[c] = rho_of(pic, thvc);
        dthbdz[c] = (theta_of_z(zc + 1.0) - theta_of_z(zc - 1.0)) / 2.0;
        thbw[c] = thw;
        rhobw[c] = rho_of(piw, thvw);
        thvb[c] = thvc;
        thvbw[c] = thvw;
        qvb[c] = qvc;
      }
  keep_qvb_ic = wk82;

  // Base wind profiles.
  const std::string wprof = cfg.get_str("wind_profile", "const");
  const double windu = cfg.get_real("wind_u", 0);
  const double windv = cfg.has("wind_v") ? cfg.get_real("wind_v", 0) : 0.0;
  const double us = cfg.get_real("wind_us", 25);
  const double zs = cfg.get_real("wind_zs", 3000);
  const double uoff = cfg.get_real("wind_offset", 0);
  ub.assign((size_t)g.NZ, windu);
  vb.assign((size_t)g.NZ, windv);
  if (wprof == "tanh") {
    for (int kk = 0; kk < g.NZ; ++kk) {
      double z = metric.zeta_c[kk];  // flat-bottomed storm cases: zeta ~ z
      if (z < 0) z = 0;
      ub[kk] = us * std::tanh(z / zs) + uoff;
    }
  } else if (wprof == "quarter_circle") {
    const double voff = cfg.get_real("wind_offset_v", 0);
    const double halfpi = 1.5707963267948966;
    for (int kk = 0; kk < g.NZ; ++kk) {
      double z = metric.zeta_c[kk];
      if (z < 0) z = 0;
      const double a = halfpi * std::min(z / zs, 1.0);
      ub[kk] = us * (1.0 - std::cos(a)) + uoff;
      vb[kk] = us * std::sin(a) + voff;
    }
  } else if (from_file && tables->u) {
    for (int kk = 0; kk < g.NZ; ++kk) {
      double z = metric.zeta_c[kk];
      if (z < 0) z = 0;
      ub[kk] = table_interp(*tables->u, z);
    }
  }
}

}  // namespace wfe
