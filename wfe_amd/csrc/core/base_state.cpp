#include "core/base_state.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "core/constants.hpp"

namespace wfe {

namespace {

// Values at heights z0 + n dz, read by linear interpolation; heights below
// or above the table extrapolate the end intervals (the index is truncated
// toward zero and then clamped, as the reference's table lookups do).
struct UniformTable {
  double z0 = 0, dz = 1;
  std::vector<double> v;
  double at(double z) const {
    const double s = (z - z0) / dz;
    const int last = (int)v.size() - 2;
    int n = (int)s;
    if (n < 0) n = 0;
    if (n > last) n = last;
    const double f = s - n;
    return v[n] * (1 - f) + v[n + 1] * f;
  }
};

// Horizontally uniform environment of a case as functions of height: the
// potential temperature theta(z) of the chosen profile and, for the WK82
// sounding, its relative humidity (src/core/base_state.cpp:52-67).
class Sounding {
 public:
  enum Kind { kIsentropic, kConstantN, kWK82, kFile };

  Sounding(const Config& cfg, const ProfileTables* tables) {
    const std::string prof = cfg.get_str("profile", "isentropic");
    theta0_ = cfg.get_real("theta0", 300);
    N_ = cfg.get_real("bv_N", 0.01);
    if (prof == "isentropic") kind_ = kIsentropic;
    else if (prof == "constant_N") kind_ = kConstantN;
    else if (prof == "wk82") kind_ = kWK82;
    else if (prof == "file") kind_ = kFile;
    else throw std::runtime_error("unknown profile: " + prof);
    if (kind_ == kFile) {
      if (!tables || !tables->z) throw std::runtime_error("profile=file but no profile tables were provided");
      const std::vector<double>& z = *tables->z;
      auto table = [&](const std::vector<double>* v) {
        UniformTable t;
        t.z0 = z[0];
        t.dz = z[1] - z[0];
        if (v) t.v = *v;
        return t;
      };
      theta_tab_ = table(tables->th);
      qv_tab_ = table(tables->qv);
      if (tables->u) u_tab_ = table(tables->u);
      has_u_tab_ = tables->u != nullptr;
    }
    rh_cap_ = cfg.has("rh_max") ? cfg.get_real("rh_max", 1.0) : 1.0;
  }

  Kind kind() const { return kind_; }
  bool moist() const { return kind_ == kWK82 || kind_ == kFile; }

  double theta(double z) const {
    switch (kind_) {
      case kFile:
        return theta_tab_.at(z);
      case kWK82:
        // theta rises as z^1.25 to the tropopause, isothermal above
        if (z <= kZtr) return 300.0 + (kThtr - 300.0) * std::pow(z / kZtr, 1.25);
        return kThtr * std::exp(phys::grav * (z - kZtr) / (phys::cp * kTtr));
      case kConstantN:
        return theta0_ * std::exp(N_ * N_ * z / phys::grav);
      default:
        return theta0_;
    }
  }
  // WK82 relative humidity, capped at rh_max (1 when absent).
  double relative_humidity(double z) const {
    const double rh = z <= kZtr ? 1.0 - 0.75 * std::pow(z / kZtr, 1.25) : 0.25;
    return std::min(rh, rh_cap_);
  }
  double max_vapour() const { return 0.014; }
  const UniformTable& vapour_table() const { return qv_tab_; }
  bool has_wind_table() const { return has_u_tab_; }
  const UniformTable& wind_table() const { return u_tab_; }

 private:
  static constexpr double kZtr = 12000.0, kThtr = 343.0, kTtr = 213.0;
  Kind kind_ = kIsentropic;
  double theta0_ = 300, N_ = 0.01, rh_cap_ = 1.0;
  UniformTable theta_tab_, qv_tab_, u_tab_;
  bool has_u_tab_ = false;
};

// Tetens saturation mixing ratio, without the reference's denominator
// floor of the microphysics (src/core/base_state.cpp:15-18).
double saturation_mixing_ratio(double p, double T) {
  const double es = 610.78 * std::exp(17.269 * (T - 273.16) / (T - 35.86));
  return 0.622 * es / (p - es);
}

// Hydrostatic column of the sounding on a 1 m grid reaching beyond the
// padded column (base_state.cpp:69-115): the Exner function from pi = 1 at
// z = 0, integrated up and down with the mean 1/theta_v of each metre; for
// WK82 five passes of vapour from the relative humidity at the current
// pressure (no vapour below ground), for file profiles the tabulated vapour.
struct HydrostaticColumn {
  UniformTable exner, vapour;

  HydrostaticColumn(const Sounding& snd, const GDims& g, double ztop) {
    constexpr double kStep = 1.0;
    const double margin = (double)(g.ng + 1) * g.dz * 4;
    const double zlo = -margin - 100.0;
    const double zhi = ztop + margin + 100.0;
    const int n = (int)((zhi - zlo) / kStep) + 2;
    surface_ = (int)(-zlo / kStep);
    exner.z0 = vapour.z0 = zlo;
    exner.dz = vapour.dz = kStep;
    exner.v.assign(n, 0.0);
    vapour.v.assign(n, 0.0);
    std::vector<double> thv(n);
    for (int m = 0; m < n; ++m) thv[m] = snd.theta(zlo + m * kStep);
    integrate(thv);
    if (snd.kind() == Sounding::kWK82) {
      for (int pass = 0; pass < 5; ++pass) {
        for (int m = 0; m < n; ++m) {
          const double z = zlo + m * kStep;
          const double th = snd.theta(z);
          double q = 0.0;
          if (z >= 0) {
            const double pi = exner.v[m];
            const double p = phys::p00 * std::pow(pi, phys::cp / phys::Rd);
            q = std::min(snd.relative_humidity(z) * saturation_mixing_ratio(p, th * pi),
                         snd.max_vapour());
          }
          vapour.v[m] = q;
          thv[m] = th * (1.0 + phys::eps61 * q);
        }
        integrate(thv);
      }
    } else if (snd.kind() == Sounding::kFile) {
      for (int m = 0; m < n; ++m) {
        const double z = zlo + m * kStep;
        vapour.v[m] = std::max(0.0, snd.vapour_table().at(z));
        thv[m] = snd.theta(z) * (1.0 + phys::eps61 * vapour.v[m]);
      }
      integrate(thv);
    }
  }

 private:
  void integrate(const std::vector<double>& thv) {
    std::vector<double>& pi = exner.v;
    const int n = (int)pi.size();
    const double g_dz = phys::grav * exner.dz;
    pi[surface_] = 1.0;
    for (int m = surface_ + 1; m < n; ++m)
      pi[m] = pi[m - 1] - g_dz / (phys::cp * (0.5 * (thv[m - 1] + thv[m])));
    for (int m = surface_ - 1; m >= 0; --m)
      pi[m] = pi[m + 1] + g_dz / (phys::cp * (0.5 * (thv[m] + thv[m + 1])));
  }
  int surface_ = 0;
};

double density(double pi, double thv) {
  return phys::p00 * std::pow(pi, phys::cv / phys::Rd) / (phys::Rd * thv);
}

// Base wind profiles u_b, v_b [NZ] at the zeta of the cell centres
// (base_state.cpp:170-189): constant wind_u (and wind_v), a tanh shear
// profile, the quarter-circle hodograph, or the file profile table.
void base_wind(const GDims& g, const Config& cfg, const Metric& metric, const Sounding& snd,
               std::vector<double>& ub, std::vector<double>& vb) {
  const std::string shape = cfg.get_str("wind_profile", "const");
  const double u0 = cfg.get_real("wind_u", 0);
  const double v0 = cfg.has("wind_v") ? cfg.get_real("wind_v", 0) : 0.0;
  const double us = cfg.get_real("wind_us", 25);
  const double zs = cfg.get_real("wind_zs", 3000);
  const double uoff = cfg.get_real("wind_offset", 0);
  ub.assign((size_t)g.NZ, u0);
  vb.assign((size_t)g.NZ, v0);
  auto height = [&](int kk) { return std::max(metric.zeta_c[kk], 0.0); };
  if (shape == "tanh") {
    for (int kk = 0; kk < g.NZ; ++kk) ub[kk] = us * std::tanh(height(kk) / zs) + uoff;
  } else if (shape == "quarter_circle") {
    const double voff = cfg.get_real("wind_offset_v", 0);
    for (int kk = 0; kk < g.NZ; ++kk) {
      const double turn = 1.5707963267948966 * std::min(height(kk) / zs, 1.0);
      ub[kk] = us * (1.0 - std::cos(turn)) + uoff;
      vb[kk] = us * std::sin(turn) + voff;
    }
  } else if (snd.kind() == Sounding::kFile && snd.has_wind_table()) {
    for (int kk = 0; kk < g.NZ; ++kk) ub[kk] = snd.wind_table().at(height(kk));
  }
}

}  // namespace

void BaseState::build(const GDims& g, const Config& cfg, const Metric& metric,
                      const ProfileTables* tables) {
  const Sounding snd(cfg, tables);
  has_moisture = snd.moist();
  keep_qvb_ic = snd.kind() == Sounding::kWK82;
  const HydrostaticColumn col(snd, g, metric.zt);

  const size_t n = g.npts();
  for (std::vector<double>* a : {&thb, &pib, &rhob, &dthbdz, &thbw, &rhobw, &thvb, &thvbw, &qvb})
    a->assign(n, 0.0);
  // every padded point, at the physical heights of its centre and w face
  for (int k = -g.ng; k < g.nz + 1 + g.ng; ++k) {
    const double zeta_c = metric.zeta_c[k + g.ng], zeta_w = metric.zeta_w[k + g.ng];
    for (int j = -g.ng; j < g.ny + g.ng; ++j)
      for (int i = -g.ng; i < g.nx + g.ng; ++i) {
        const size_t c = g.idx(i, j, k);
        const double zc = metric.z_at(g, i, j, zeta_c);
        const double zw = metric.z_at(g, i, j, zeta_w);
        const double th_c = snd.theta(zc), th_w = snd.theta(zw);
        const double q_c = has_moisture ? col.vapour.at(zc) : 0.0;
        const double q_w = has_moisture ? col.vapour.at(zw) : 0.0;
        const double thv_c = th_c * (1.0 + phys::eps61 * q_c);
        const double thv_w = th_w * (1.0 + phys::eps61 * q_w);
        const double pi_c = col.exner.at(zc);
        thb[c] = th_c;
        pib[c] = pi_c;
        rhob[c] = density(pi_c, thv_c);
        dthbdz[c] = (snd.theta(zc + 1.0) - snd.theta(zc - 1.0)) / 2.0;
        thbw[c] = th_w;
        rhobw[c] = density(col.exner.at(zw), thv_w);
        thvb[c] = thv_c;
        thvbw[c] = thv_w;
        qvb[c] = q_c;
      }
  }
  base_wind(g, cfg, metric, snd, ub, vb);
}

}  // namespace wfe
