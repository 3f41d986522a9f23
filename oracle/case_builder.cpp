// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.
// case_builder.cpp -- CPU FP64 restatement of the reference's host-side case
// assembly: Config (src/core/config.cpp:17-60), file-mode input
// (src/io/input.cpp:19-73), Metric::build (src/core/metric.cpp:12-168),
// BaseState::build (src/core/base_state.cpp:21-193), the initial conditions
// and Lambert map factor of src/main.cpp:28-195,261-383, and BdyManager
// (src/dynamics/boundary.cpp:9-97).
//
// TEST INFRASTRUCTURE ONLY (see oracle.h): it builds the oracle's own base
// state, metric, initial state and boundary targets from a case file, so the
// product's host builders and the whole product pipeline can be checked
// against an independent restatement. It follows the reference statement by
// statement, in FP64 (the reference's `real` is double in its FP64 build).
//
// Extensions shared with the product (each off by default, reproducing the
// reference when absent): rh_max (cap on the WK82 relative humidity),
// wind_v (constant base v), wind_profile = quarter_circle (u_b, v_b).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "oracle.h"

namespace orcb {

namespace {

constexpr double kGrav = 9.81, kCp = 1004.5, kRd = 287.04, kCv = kCp - kRd, kP00 = 1.0e5;
constexpr double kEps61 = 0.61;

// ---------------------------------------------------------------- config
// src/core/config.cpp:9-49
struct Cfg {
  std::map<std::string, std::string> kv;
  static std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
  }
  bool load(const std::string& path) {
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
      size_t c = line.find_first_of("#;");
      if (c != std::string::npos) line = line.substr(0, c);
      size_t eq = line.find('=');
      if (eq == std::string::npos) continue;
      std::string key = trim(line.substr(0, eq));
      std::string val = trim(line.substr(eq + 1));
      if (!key.empty()) kv[key] = val;
    }
    return true;
  }
  bool has(const std::string& k) const { return kv.count(k) > 0; }
  double get_real(const std::string& k, double def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::atof(it->second.c_str());
  }
  int get_int(const std::string& k, int def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::atoi(it->second.c_str());
  }
  std::string get_str(const std::string& k, const std::string& def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : it->second;
  }
};

// ----------------------------------------------------------------- grid
struct Dims {
  int nx, ny, nz, ng, NX, NY, NZ;
  double dx, dy, dz;
  size_t idx(int i, int j, int k) const {
    return ((size_t)(k + ng) * NY + (j + ng)) * NX + (i + ng);
  }
  size_t idx2(int i, int j) const { return (size_t)(j + ng) * NX + (i + ng); }
  size_t npts() const { return (size_t)NX * NY * NZ; }
  size_t n2() const { return (size_t)NX * NY; }
};

// ---------------------------------------------------------------- input
// src/io/input.cpp:10-73
struct Input {
  std::string dir, start;
  std::vector<double> prof_z, prof_th, prof_qv, prof_u;
  std::vector<double> h, fcor, tsk, land, lat, lon, soilw;
  std::vector<double> u, v, th, pi, qv;
  double bdy_interval = 10800;
  int n_bdy = 0;

  static bool read_f32(FILE* f, std::vector<double>& dst, size_t n) {
    dst.resize(n);
    std::vector<float> tmp(n);
    if (n && std::fread(tmp.data(), sizeof(float), n, f) != n) return false;
    for (size_t i = 0; i < n; ++i) dst[i] = (double)tmp[i];
    return true;
  }
  bool load(const Dims& g, const std::string& d) {
    dir = d;
    Cfg meta;
    if (!meta.load(d + "/wfe_input.ini")) return false;
    if (meta.get_int("nx", -1) != g.nx || meta.get_int("ny", -1) != g.ny ||
        meta.get_int("nz", -1) != g.nz)
      return false;
    if (meta.get_int("version", 1) < 3) return false;
    int npf = meta.get_int("np_prof", 0);
    bdy_interval = meta.get_real("bdy_interval", 10800);
    n_bdy = meta.get_int("n_bdy", 0);
    start = meta.get_str("start", "");
    FILE* f = std::fopen((d + "/wfe_init.bin").c_str(), "rb");
    if (!f) return false;
    size_t n2 = (size_t)g.nx * g.ny, n3 = n2 * g.nz;
    bool ok = read_f32(f, prof_z, npf) && read_f32(f, prof_th, npf) &&
              read_f32(f, prof_qv, npf) && read_f32(f, prof_u, npf) && read_f32(f, h, n2) &&
              read_f32(f, fcor, n2) && read_f32(f, tsk, n2) && read_f32(f, land, n2) &&
              read_f32(f, lat, n2) && read_f32(f, lon, n2) && read_f32(f, soilw, n2) &&
              read_f32(f, u, n3) && read_f32(f, v, n3) && read_f32(f, th, n3) &&
              read_f32(f, pi, n3) && read_f32(f, qv, n3);
    std::fclose(f);
    return ok;
  }
  bool load_bdy(const Dims& g, int idx, std::vector<double> fields[5]) const {
    char path[1024];
    std::snprintf(path, sizeof(path), "%s/wfe_bdy_%03d.bin", dir.c_str(),
                  (int)(idx * bdy_interval / 3600));
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    size_t n3 = (size_t)g.nx * g.ny * g.nz;
    bool ok = true;
    for (int v = 0; v < 5 && ok; ++v) ok = read_f32(f, fields[v], n3);
    std::fclose(f);
    return ok;
  }
};

// --------------------------------------------------------------- metric
// src/core/metric.cpp:12-168
struct Met {
  double zt = 0;
  std::vector<double> zeta_w, zeta_c, dzeta_c, dzeta_w;
  std::vector<double> h, hx, hy, hx_u, hy_v, jac, fcor, mapf;
  double z_at(const Dims& g, int i, int j, double zeta) const {
    double hh = h[g.idx2(i, j)];
    return hh + zeta * (zt - hh) / zt;
  }
};

bool build_metric(const Dims& g, const Cfg& cfg, const std::vector<double>* h_file,
                  const std::vector<double>* fcor_file, const std::vector<double>* mapf_file,
                  Met& m) {
  const size_t n1 = (size_t)g.NZ, n2 = g.n2();
  m.zeta_w.assign(n1, 0);
  m.zeta_c.assign(n1, 0);
  m.dzeta_c.assign(n1, 0);
  m.dzeta_w.assign(n1, 0);
  std::string stretch = cfg.get_str("stretch", "none");
  std::vector<double> zw(g.nz + 1, 0.0);
  if (stretch == "geometric") {  // metric.cpp:26-34
    double dz0 = cfg.get_real("dz0", g.dz);
    double ratio = cfg.get_real("dz_ratio", 1.05);
    double dzmax = cfg.get_real("dz_max", g.dz * 4);
    double d = dz0;
    for (int k = 1; k <= g.nz; ++k) {
      zw[k] = zw[k - 1] + d;
      d = std::min(d * ratio, dzmax);
    }
  } else {
    for (int k = 0; k <= g.nz; ++k) zw[k] = (double)k * g.dz;
  }
  m.zt = zw[g.nz];
  auto zw_at = [&](int k) -> double {  // metric.cpp:41-45
    if (k < 0) return zw[0] + (double)k * (zw[1] - zw[0]);
    if (k > g.nz) return zw[g.nz] + (double)(k - g.nz) * (zw[g.nz] - zw[g.nz - 1]);
    return zw[k];
  };
  for (int kk = 0; kk < (int)n1; ++kk) {
    int k = kk - g.ng;
    m.zeta_w[kk] = zw_at(k);
    m.zeta_c[kk] = 0.5 * (zw_at(k) + zw_at(k + 1));
  }
  for (int kk = 0; kk < (int)n1 - 1; ++kk) m.dzeta_c[kk] = m.zeta_w[kk + 1] - m.zeta_w[kk];
  m.dzeta_c[n1 - 1] = m.dzeta_c[n1 - 2];
  for (int kk = 1; kk < (int)n1; ++kk) m.dzeta_w[kk] = m.zeta_c[kk] - m.zeta_c[kk - 1];
  m.dzeta_w[0] = m.dzeta_w[1];

  m.h.assign(n2, 0);
  m.jac.assign(n2, 0);
  m.hx.assign(n2, 0);
  m.hy.assign(n2, 0);
  m.hx_u.assign(n2, 0);
  m.hy_v.assign(n2, 0);
  std::string terrain = cfg.get_str("terrain", "none");
  double h0 = cfg.get_real("mtn_h0", 250);
  double xa = cfg.get_real("mtn_a", 5000);
  double lam = cfg.get_real("mtn_lambda", 4000);
  double xc = cfg.get_real("mtn_xc", g.nx * g.dx / 2);
  double yc = cfg.get_real("mtn_yc", g.ny * g.dy / 2);
  double ya = cfg.get_real("mtn_ya", xa);
  const double pi = 3.14159265358979323846;
  auto h_of = [&](double x, double y) -> double {  // metric.cpp:72-84
    if (terrain == "schaer") {
      double xr = x - xc;
      double env = std::exp(-(xr / xa) * (xr / xa));
      double c = std::cos(pi * xr / lam);
      return h0 * env * c * c;
    }
    if (terrain == "gauss") {
      double xr = (x - xc) / xa, yr = (y - yc) / ya;
      return h0 * std::exp(-xr * xr - yr * yr);
    }
    return 0.0;
  };
  auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
  if (terrain == "file") {  // metric.cpp:91-109
    if (!h_file) return false;
    for (int j = -g.ng; j < g.ny + g.ng; ++j)
      for (int i = -g.ng; i < g.nx + g.ng; ++i) {
        int ic = clampi(i, 0, g.nx - 1), jc = clampi(j, 0, g.ny - 1);
        m.h[g.idx2(i, j)] = (*h_file)[(size_t)jc * g.nx + ic];
      }
    for (int j = -g.ng; j < g.ny + g.ng; ++j)
      for (int i = -g.ng; i < g.nx + g.ng; ++i) {
        size_t c = g.idx2(i, j);
        m.jac[c] = (m.zt - m.h[c]) / m.zt;
        int im = i > -g.ng ? i - 1 : i, ip = i < g.nx + g.ng - 1 ? i + 1 : i;
        int jm = j > -g.ng ? j - 1 : j, jp = j < g.ny + g.ng - 1 ? j + 1 : j;
        m.hx[c] = (m.h[g.idx2(ip, j)] - m.h[g.idx2(im, j)]) / ((double)(ip - im) * g.dx);
        m.hy[c] = (m.h[g.idx2(i, jp)] - m.h[g.idx2(i, jm)]) / ((double)(jp - jm) * g.dy);
        m.hx_u[c] = (i == im) ? 0.0 : (m.h[c] - m.h[g.idx2(im, j)]) / g.dx;
        m.hy_v[c] = (j == jm) ? 0.0 : (m.h[c] - m.h[g.idx2(i, jm)]) / g.dy;
      }
  } else {  // metric.cpp:111-122
    for (int j = -g.ng; j < g.ny + g.ng; ++j)
      for (int i = -g.ng; i < g.nx + g.ng; ++i) {
        double x = ((double)i + 0.5) * g.dx;
        double y = ((double)j + 0.5) * g.dy;
        size_t c = g.idx2(i, j);
        m.h[c] = h_of(x, y);
        m.jac[c] = (m.zt - m.h[c]) / m.zt;
        m.hx[c] = (h_of(x + g.dx, y) - h_of(x - g.dx, y)) / (2.0 * g.dx);
        m.hy[c] = (h_of(x, y + g.dy) - h_of(x, y - g.dy)) / (2.0 * g.dy);
        m.hx_u[c] = (h_of(x, y) - h_of(x - g.dx, y)) / g.dx;
        m.hy_v[c] = (h_of(x, y) - h_of(x, y - g.dy)) / g.dy;
      }
  }
  m.fcor.assign(n2, cfg.get_real("coriolis_f", 0));  // metric.cpp:126-134
  if (fcor_file)
    for (int j = -g.ng; j < g.ny + g.ng; ++j)
      for (int i = -g.ng; i < g.nx + g.ng; ++i) {
        int ic = clampi(i, 0, g.nx - 1), jc = clampi(j, 0, g.ny - 1);
        m.fcor[g.idx2(i, j)] = (*fcor_file)[(size_t)jc * g.nx + ic];
      }
  m.mapf.assign(n2, 1.0);  // metric.cpp:137-145
  if (mapf_file)
    for (int j = -g.ng; j < g.ny + g.ng; ++j)
      for (int i = -g.ng; i < g.nx + g.ng; ++i) {
        int ic = clampi(i, 0, g.nx - 1), jc = clampi(j, 0, g.ny - 1);
        m.mapf[g.idx2(i, j)] = (*mapf_file)[(size_t)jc * g.nx + ic];
      }
  return true;
}

// ----------------------------------------------------------- base state
// src/core/base_state.cpp:15-18
double qsat_of(double p, double T) {
  double es = 610.78 * std::exp(17.269 * (T - 273.16) / (T - 35.86));
  return 0.622 * es / (p - es);
}

struct Base {
  std::vector<double> thb, pib, rhob, dthbdz, thbw, rhobw, thvb, thvbw, qvb, ub, vb;
  bool has_moisture = false, wk82 = false;
};

// src/core/base_state.cpp:21-193
bool build_base(const Dims& g, const Cfg& cfg, const Met& metric, const Input* tables,
                Base& b) {
  std::string prof = cfg.get_str("profile", "isentropic");
  double theta0 = cfg.get_real("theta0", 300);
  double bvN = cfg.get_real("bv_N", 0.01);
  bool constN = prof == "constant_N", wk82 = prof == "wk82", from_file = prof == "file";
  if (from_file && (!tables || tables->prof_z.size() < 2)) return false;
  b.has_moisture = wk82 || from_file;
  b.wk82 = wk82;
  if (!constN && !wk82 && !from_file && prof != "isentropic") return false;
  double rh_max = cfg.has("rh_max") ? cfg.get_real("rh_max", 1.0) : 1.0;

  auto table_interp = [&](const std::vector<double>& tab, double z) -> double {  // :41-50
    const std::vector<double>& zt = tables->prof_z;
    double dzt = zt[1] - zt[0];
    double s = (z - zt[0]) / dzt;
    int it = (int)s;
    if (it < 0) it = 0;
    if (it > (int)zt.size() - 2) it = (int)zt.size() - 2;
    double f = s - it;
    return tab[it] * (1 - f) + tab[it + 1] * f;
  };
  const double z_tr = 12000.0, th_tr = 343.0, T_tr = 213.0, qv_max = 0.014;  // :53
  auto theta_of_z = [&](double z) -> double {  // :55-63
    if (from_file) return table_interp(tables->prof_th, z);
    if (wk82) {
      if (z <= z_tr) return 300.0 + (th_tr - 300.0) * std::pow(z / z_tr, 1.25);
      return th_tr * std::exp(kGrav * (z - z_tr) / (kCp * T_tr));
    }
    if (constN) return theta0 * std::exp(bvN * bvN * z / kGrav);
    return theta0;
  };
  auto relhum_of_z = [&](double z) -> double {  // :64-67 (+ rh_max cap)
    double rh = (z <= z_tr) ? 1.0 - 0.75 * std::pow(z / z_tr, 1.25) : 0.25;
    return std::min(rh, rh_max);
  };

  double zmin = -(double)(g.ng + 1) * g.dz * 4 - 100.0;  // :70-75
  double zmax = metric.zt + (double)(g.ng + 1) * g.dz * 4 + 100.0;
  const double dzt = 1.0;
  int ntab = (int)((zmax - zmin) / dzt) + 2;
  int i0 = (int)(-zmin / dzt);
  std::vector<double> pitab(ntab), qvtab(ntab, 0.0), thvtab(ntab);
  for (int it = 0; it < ntab; ++it) thvtab[it] = theta_of_z(zmin + it * dzt);
  auto integrate_pi = [&]() {  // :80-90
    pitab[i0] = 1.0;
    for (int it = i0 + 1; it < ntab; ++it) {
      double thv = 0.5 * (thvtab[it - 1] + thvtab[it]);
      pitab[it] = pitab[it - 1] - kGrav * dzt / (kCp * thv);
    }
    for (int it = i0 - 1; it >= 0; --it) {
      double thv = 0.5 * (thvtab[it] + thvtab[it + 1]);
      pitab[it] = pitab[it + 1] + kGrav * dzt / (kCp * thv);
    }
  };
  integrate_pi();
  if (wk82) {  // :93-107
    for (int iter = 0; iter < 5; ++iter) {
      for (int it = 0; it < ntab; ++it) {
        double z = zmin + it * dzt;
        double th = theta_of_z(z);
        double pi = pitab[it];
        double T = th * pi;
        double p = kP00 * std::pow(pi, kCp / kRd);
        double qv = std::min(relhum_of_z(z) * qsat_of(p, T), qv_max);
        if (z < 0) qv = qvtab[it > 0 ? it : 0];
        qvtab[it] = qv;
        thvtab[it] = th * (1.0 + kEps61 * qv);
      }
      integrate_pi();
    }
  } else if (from_file) {  // :108-115
    for (int it = 0; it < ntab; ++it) {
      double z = zmin + it * dzt;
      qvtab[it] = std::max(0.0, table_interp(tables->prof_qv, z));
      thvtab[it] = theta_of_z(z) * (1.0 + kEps61 * qvtab[it]);
    }
    integrate_pi();
  }
  auto interp = [&](const std::vector<double>& tab, double z) -> double {  // :117-124
    double s = (z - zmin) / dzt;
    int it = (int)s;
    if (it < 0) it = 0;
    if (it > ntab - 2) it = ntab - 2;
    double f = s - it;
    return tab[it] * (1 - f) + tab[it + 1] * f;
  };
  auto rho_of = [&](double pi, double thv) -> double {  // :125-128
    return kP00 * std::pow(pi, kCv / kRd) / (kRd * thv);
  };
  size_t n = g.npts();
  for (auto* v : {&b.thb, &b.pib, &b.rhob, &b.dthbdz, &b.thbw, &b.rhobw, &b.thvb, &b.thvbw,
                  &b.qvb})
    v->assign(n, 0.0);
  for (int k = -g.ng; k < g.nz + 1 + g.ng; ++k)  // :134-155
    for (int j = -g.ng; j < g.ny + g.ng; ++j)
      for (int i = -g.ng; i < g.nx + g.ng; ++i) {
        size_t c = g.idx(i, j, k);
        double zc = metric.z_at(g, i, j, metric.zeta_c[k + g.ng]);
        double zw = metric.z_at(g, i, j, metric.zeta_w[k + g.ng]);
        double thc = theta_of_z(zc), thw = theta_of_z(zw);
        double qvc = b.has_moisture ? interp(qvtab, zc) : 0.0;
        double qvw = b.has_moisture ? interp(qvtab, zw) : 0.0;
        double thvc = thc * (1.0 + kEps61 * qvc);
        double thvw = thw * (1.0 + kEps61 * qvw);
        double pic = interp(pitab, zc), piw = interp(pitab, zw);
        b.thb[c] = thc;
        b.pib[c] = pic;
        b.rhob[c] = rho_of(pic, thvc);
        b.dthbdz[c] = (theta_of_z(zc + 1.0) - theta_of_z(zc - 1.0)) / 2.0;
        b.thbw[c] = thw;
        b.rhobw[c] = rho_of(piw, thvw);
        b.thvb[c] = thvc;
        b.thvbw[c] = thvw;
        b.qvb[c] = qvc;
      }
  // base wind (:171-189, + wind_v and quarter_circle)
  std::string wprof = cfg.get_str("wind_profile", "const");
  double windu = cfg.get_real("wind_u", 0);
  double windv = cfg.has("wind_v") ? cfg.get_real("wind_v", 0) : 0.0;
  double us = cfg.get_real("wind_us", 25);
  double zs = cfg.get_real("wind_zs", 3000);
  double uoff = cfg.get_real("wind_offset", 0);
  b.ub.assign((size_t)g.NZ, windu);
  b.vb.assign((size_t)g.NZ, windv);
  if (wprof == "tanh") {
    for (int kk = 0; kk < g.NZ; ++kk) {
      double z = metric.zeta_c[kk];
      if (z < 0) z = 0;
      b.ub[kk] = us * std::tanh(z / zs) + uoff;
    }
  } else if (wprof == "quarter_circle") {
    double voff = cfg.get_real("wind_offset_v", 0);
    for (int kk = 0; kk < g.NZ; ++kk) {
      double z = metric.zeta_c[kk];
      if (z < 0) z = 0;
      double a = 1.5707963267948966 * std::min(z / zs, 1.0);
      b.ub[kk] = us * (1.0 - std::cos(a)) + uoff;
      b.vb[kk] = us * std::sin(a) + voff;
    }
  } else if (from_file && !tables->prof_u.empty()) {
    for (int kk = 0; kk < g.NZ; ++kk) {
      double z = metric.zeta_c[kk];
      if (z < 0) z = 0;
      b.ub[kk] = table_interp(tables->prof_u, z);
    }
  }
  return true;
}

// ------------------------------------------------------------ boundary
// src/dynamics/boundary.cpp:9-97
struct Bdy {
  bool active = false;
  std::vector<double> wgt;
  std::array<std::vector<double>, 5> lo, hi;
  double t_lo = 0, t_hi = 0;
  int idx_hi = 1;
  bool changed = false;

  bool load_into(const Dims& g, const Input& in, const Base& b, int idx,
                 std::array<std::vector<double>, 5>& dst) {  // :51-81
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
            size_t c = g.idx(i, j, k);
            double val;
            if (v == 0) {
              int il = i > 0 ? i - 1 : 0, ir = i < g.nx ? i : g.nx - 1;
              val = 0.5 * (src(f[0], il, j, k) + src(f[0], ir, j, k));
            } else if (v == 1) {
              int jl = j > 0 ? j - 1 : 0;
              val = 0.5 * (src(f[1], i, jl, k) + src(f[1], i, j, k));
            } else if (v == 2) {
              val = src(f[2], i, j, k) - b.thb[c];
            } else if (v == 3) {
              val = src(f[3], i, j, k) - b.pib[c];
            } else {
              val = src(f[4], i, j, k);
            }
            dst[v][c] = val;
          }
    }
    return true;
  }
  bool init(const Dims& g, bool bcx, bool bcy, const Input& in, const Base& b, int bdy_width,
            double bdy_tau, double nudge_tau) {  // :9-49
    active = true;
    double inner = nudge_tau > 0 ? 1.0 / nudge_tau : 0.0;
    wgt.assign(g.n2(), 0.0);
    const double pi_half = 1.5707963267948966;
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        double wij = inner;
        int dist = 1 << 20;
        if (bcx) dist = std::min({dist, i, g.nx - 1 - i});
        if (bcy) dist = std::min({dist, j, g.ny - 1 - j});
        if (dist < bdy_width) {
          double s = std::cos(pi_half * (double)dist / (double)bdy_width);
          wij = std::max(wij, s * s / bdy_tau);
        }
        wgt[(size_t)(j + g.ng) * g.NX + (i + g.ng)] = wij;
      }
    if (!load_into(g, in, b, 0, lo)) return false;
    if (!load_into(g, in, b, std::min(1, in.n_bdy - 1), hi)) return false;
    t_lo = 0;
    t_hi = in.bdy_interval;
    idx_hi = 1;
    return true;
  }
  bool update(const Dims& g, const Input& in, const Base& b, double t) {  // :83-91
    while (t >= t_hi && idx_hi + 1 < in.n_bdy) {
      std::swap(lo, hi);
      ++idx_hi;
      if (!load_into(g, in, b, idx_hi, hi)) return false;
      t_lo = t_hi;
      t_hi += in.bdy_interval;
      changed = true;
    }
    return true;
  }
  double tfrac(double t) const {  // :93-97
    if (t_hi <= t_lo) return 0;
    double f = (t - t_lo) / (t_hi - t_lo);
    return f < 0 ? 0 : (f > 1 ? 1.0 : f);
  }
};

// ------------------------------------------------------------------ case
struct Case {
  Cfg cfg;
  Dims g{};
  Met met;
  Base base;
  Input input;
  Bdy bdy;
  bool file_mode = false;
  std::map<std::string, double> scalars, params;
  std::array<std::vector<double>, 8> ic;  // u v w thp pip qv qc qr
  // surface physics initial fields (src/main.cpp:364-377, surface.cu:465-472)
  std::vector<double> sfc_tsk, sfc_land, sfc_lat, sfc_lon, sfc_soilw;
};

// src/main.cpp:28-54
void init_bubble(Case& cs) {
  const Dims& g = cs.g;
  double dth = cs.cfg.get_real("bubble_dtheta", 0.0);
  if (dth == 0.0) return;
  double xc = cs.cfg.get_real("bubble_xc", g.nx * g.dx / 2);
  double yc = cs.cfg.get_real("bubble_yc", g.ny * g.dy / 2);
  double zc = cs.cfg.get_real("bubble_zc", 2000.0);
  double xr = cs.cfg.get_real("bubble_xr", 2000.0);
  double yr = cs.cfg.get_real("bubble_yr", 2000.0);
  double zr = cs.cfg.get_real("bubble_zr", 2000.0);
  const double pi = 3.14159265358979323846;
  std::vector<double>& h = cs.ic[3];
  std::fill(h.begin(), h.end(), 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        double x = ((double)i + 0.5) * g.dx;
        double y = ((double)j + 0.5) * g.dy;
        double z = cs.met.z_at(g, i, j, cs.met.zeta_c[k + g.ng]);
        double rx = (x - xc) / xr, ry = (y - yc) / yr, rz = (z - zc) / zr;
        double r = std::sqrt(rx * rx + ry * ry + rz * rz);
        if (r < 1.0) {
          double c = std::cos(pi * r / 2);
          h[g.idx(i, j, k)] = dth * c * c;
        }
      }
}

// src/main.cpp:58-100
void init_from_input(Case& cs) {
  const Dims& g = cs.g;
  const Input& in = cs.input;
  auto src = [&](const std::vector<double>& a, int i, int j, int k) {
    return a[((size_t)k * g.ny + j) * g.nx + i];
  };
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i <= g.nx; ++i) {
        int il = i > 0 ? i - 1 : 0, ir = i < g.nx ? i : g.nx - 1;
        cs.ic[0][g.idx(i, j, k)] = 0.5 * (src(in.u, il, j, k) + src(in.u, ir, j, k));
      }
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        int jl = j > 0 ? j - 1 : 0;
        cs.ic[1][g.idx(i, j, k)] = 0.5 * (src(in.v, i, jl, k) + src(in.v, i, j, k));
      }
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        size_t c = g.idx(i, j, k);
        cs.ic[3][c] = src(in.th, i, j, k) - cs.base.thb[c];
        cs.ic[4][c] = src(in.pi, i, j, k) - cs.base.pib[c];
        cs.ic[5][c] = src(in.qv, i, j, k);
      }
}

// src/main.cpp:105-144
void add_ic_perturbation(Case& cs, double amp, unsigned seed) {
  const Dims& g = cs.g;
  if (amp <= 0.0) return;
  size_t n = g.npts();
  std::vector<double> p(n, 0), tmp(n, 0);
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
          double s6 = tmp[g.idx(i - 1, j, k)] + tmp[g.idx(i + 1, j, k)] +
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
        double v = p[g.idx(i, j, k)];
        ss += v * v;
        ++cnt;
      }
  double rms = std::sqrt(ss / cnt);
  if (rms < 1e-9) return;
  double sc = amp / rms;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) cs.ic[3][g.idx(i, j, k)] += sc * p[g.idx(i, j, k)];
}

// src/main.cpp:147-154
void init_moisture(Case& cs) {
  const Dims& g = cs.g;
  if (!cs.base.has_moisture) return;
  std::fill(cs.ic[5].begin(), cs.ic[5].end(), 0.0);
  if (!cs.base.wk82) return;  // BaseState::qvb_at reads 0 unless wk82 (base_state.cpp:168)
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) cs.ic[5][g.idx(i, j, k)] = cs.base.qvb[g.idx(i, j, k)];
}

// src/main.cpp:158-183
void init_qv_blob(Case& cs) {
  const Dims& g = cs.g;
  double dq = cs.cfg.get_real("qvbub_dq", 0.0);
  if (dq == 0.0) return;
  double xc = cs.cfg.get_real("qvbub_xc", g.nx * g.dx / 4);
  double yc = cs.cfg.get_real("qvbub_yc", g.ny * g.dy / 2);
  double zc = cs.cfg.get_real("qvbub_zc", 2000.0);
  double xr = cs.cfg.get_real("qvbub_xr", 2000.0);
  double yr = cs.cfg.get_real("qvbub_yr", 1e9);
  double zr = cs.cfg.get_real("qvbub_zr", 2000.0);
  const double pi = 3.14159265358979323846;
  std::vector<double>& h = cs.ic[5];
  std::fill(h.begin(), h.end(), 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        double x = ((double)i + 0.5) * g.dx;
        double y = ((double)j + 0.5) * g.dy;
        double z = cs.met.z_at(g, i, j, cs.met.zeta_c[k + g.ng]);
        double rx = (x - xc) / xr, ry = (y - yc) / yr, rz = (z - zc) / zr;
        double r = std::sqrt(rx * rx + ry * ry + rz * rz);
        if (r < 1.0) {
          double c = std::cos(pi * r / 2);
          h[g.idx(i, j, k)] = dq * c * c;
        }
      }
}

// src/main.cpp:186-195 (u); the v faces carry v_b the same way.
void init_wind(Case& cs) {
  const Dims& g = cs.g;
  for (int k = 0; k < g.nz; ++k) {
    double ub = cs.base.ub[k + g.ng];
    if (ub == 0.0) continue;
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i <= g.nx; ++i) cs.ic[0][g.idx(i, j, k)] = ub;
  }
  for (int k = 0; k < g.nz; ++k) {
    double vb = cs.base.vb[k + g.ng];
    if (vb == 0.0) continue;
    for (int j = 0; j <= g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) cs.ic[1][g.idx(i, j, k)] = vb;
  }
}

// src/main.cpp:261-383
bool build_case(Case& cs) {
  Cfg& cfg = cs.cfg;
  Dims& g = cs.g;
  g.nx = cfg.get_int("nx", 100);
  g.ny = cfg.get_int("ny", 100);
  g.nz = cfg.get_int("nz", 50);
  g.dx = cfg.get_real("dx", 200);
  g.dy = cfg.get_real("dy", 200);
  g.dz = cfg.get_real("dz", 200);
  g.ng = 3;
  if (g.nx < 4 || g.ny < 1 || g.nz < 4) return false;
  g.NX = g.nx + 2 * g.ng;
  g.NY = g.ny + 2 * g.ng;
  g.NZ = g.nz + 1 + 2 * g.ng;
  cs.scalars["dt"] = cfg.get_real("dt", 0.25);
  cs.scalars["t_end"] = cfg.get_real("t_end", 1000);
  auto& P = cs.params;  // main.cpp:284-298
  P["diff_K"] = cfg.get_real("diff_K", 0);
  P["coriolis_f"] = cfg.get_real("coriolis_f", 0);
  P["coriolis_use_ub"] = 1;
  P["rayleigh_zd"] = cfg.get_real("rayleigh_zd", -1);
  P["rayleigh_alpha"] = cfg.get_real("rayleigh_alpha", 0);
  P["acoustic_ns"] = cfg.get_int("acoustic_ns", 6);
  P["acoustic_beta"] = cfg.get_real("acoustic_beta", 0.2);
  P["acoustic_smdiv"] = cfg.get_real("acoustic_smdiv", 0.1);
  P["bc_x_open"] = cfg.get_str("bc_x", "periodic") == "open";
  P["bc_y_open"] = cfg.get_str("bc_y", "periodic") == "open";
  P["cstar"] = cfg.get_real("cstar", 30);
  P["w_damping"] = cfg.get_str("w_damping", "off") == "on";
  P["pd_moist"] = cfg.get_str("pd_moist", "off") == "on";
  P["microphysics"] = cfg.get_str("microphysics", "on") == "on";

  cs.file_mode = cfg.get_str("profile", "isentropic") == "file";  // :313-321
  if (cs.file_mode) {
    if (!cs.input.load(g, cfg.get_str("input_dir", ""))) return false;
    P["coriolis_use_ub"] = 0;
  }
  std::vector<double> mapf;
  if (cs.file_mode) {  // :325-340
    double phi1 = cfg.get_real("proj_lat1", 35) * 0.01745329252;
    double phi2 = cfg.get_real("proj_lat2", 43) * 0.01745329252;
    const double PI4 = 0.78539816339;
    double n = (std::fabs(phi1 - phi2) < 1e-6)
                   ? std::sin(phi1)
                   : std::log(std::cos(phi1) / std::cos(phi2)) /
                         std::log(std::tan(PI4 + phi2 / 2) / std::tan(PI4 + phi1 / 2));
    double A = std::cos(phi1) * std::pow(std::tan(PI4 + phi1 / 2), n);
    mapf.resize((size_t)g.nx * g.ny);
    for (size_t idx = 0; idx < mapf.size(); ++idx) {
      double phi = cs.input.lat[idx] * 0.01745329252;
      mapf[idx] = A / (std::pow(std::tan(PI4 + phi / 2), n) * std::cos(phi));
    }
  }
  if (!build_metric(g, cfg, cs.file_mode ? &cs.input.h : nullptr,
                    cs.file_mode ? &cs.input.fcor : nullptr, cs.file_mode ? &mapf : nullptr,
                    cs.met))
    return false;
  if (!build_base(g, cfg, cs.met, cs.file_mode ? &cs.input : nullptr, cs.base)) return false;
  bool qv_blob = cfg.get_real("qvbub_dq", 0) != 0.0;  // :349-351
  P["moisture"] = cs.base.has_moisture || qv_blob;
  if (!P["moisture"]) P["pd_moist"] = 0;
  cs.scalars["zt"] = cs.met.zt;

  for (auto& f : cs.ic) f.assign(g.npts(), 0.0);
  if (cs.file_mode) {  // :357-363
    init_from_input(cs);
    add_ic_perturbation(cs, cfg.get_real("ic_perturb", 0), (unsigned)cfg.get_int("ens_seed", 0));
    if (!cs.bdy.init(g, P["bc_x_open"] != 0, P["bc_y_open"] != 0, cs.input, cs.base,
                     cfg.get_int("bdy_width", 8), cfg.get_real("bdy_tau", 600),
                     cfg.get_real("nudge_tau", 0)))
      return false;
    // physics = simple (main.cpp:364-377): day of year and UTC hour of the
    // input's start time, local or nonlocal boundary layer
    cs.scalars["sfc_on"] = 0;
    if (cfg.get_str("physics", "none") == "simple") {
      const std::string& st = cs.input.start;
      if (st.size() < 10) return false;
      int mo = std::atoi(st.substr(4, 2).c_str());
      int dy = std::atoi(st.substr(6, 2).c_str());
      int hh = std::atoi(st.substr(8, 2).c_str());
      static const int cum[12] = {0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334};
      if (mo < 1 || mo > 12) return false;
      cs.scalars["sfc_on"] = 1;
      cs.scalars["sfc_doy"] = cum[mo - 1] + dy;
      cs.scalars["sfc_start_hour"] = hh;
      cs.scalars["sfc_nonlocal"] = cfg.get_str("pbl", "local") == "nonlocal";
      auto up2 = [&](const std::vector<double>& src) {  // surface.cu:465-471
        std::vector<double> h(g.n2(), 0.0);
        for (int j = 0; j < g.ny; ++j)
          for (int i = 0; i < g.nx; ++i)
            h[(size_t)(j + g.ng) * g.NX + (i + g.ng)] = src[(size_t)j * g.nx + i];
        return h;
      };
      cs.sfc_tsk = up2(cs.input.tsk);
      cs.sfc_land = up2(cs.input.land);
      cs.sfc_lat = up2(cs.input.lat);
      cs.sfc_lon = up2(cs.input.lon);
      cs.sfc_soilw = up2(cs.input.soilw);
    }
  } else {  // :379-382
    cs.scalars["sfc_on"] = 0;
    init_bubble(cs);
    init_wind(cs);
    init_moisture(cs);
    init_qv_blob(cs);
  }
  return true;
}

const std::vector<double>* array_of(const Case& cs, const std::string& n) {
  const Base& b = cs.base;
  const Met& m = cs.met;
  const std::pair<const char*, const std::vector<double>*> t[] = {
      {"thb", &b.thb},       {"pib", &b.pib},       {"rhob", &b.rhob},
      {"dthbdz", &b.dthbdz}, {"thbw", &b.thbw},     {"rhobw", &b.rhobw},
      {"thvb", &b.thvb},     {"thvbw", &b.thvbw},   {"qvb", &b.qvb},
      {"ub", &b.ub},         {"vb", &b.vb},         {"zeta_w", &m.zeta_w},
      {"zeta_c", &m.zeta_c}, {"dzeta_c", &m.dzeta_c}, {"dzeta_w", &m.dzeta_w},
      {"h", &m.h},           {"hx", &m.hx},         {"hy", &m.hy},
      {"hx_u", &m.hx_u},     {"hy_v", &m.hy_v},     {"jac", &m.jac},
      {"fcor", &m.fcor},     {"mapf", &m.mapf},     {"u", &cs.ic[0]},
      {"v", &cs.ic[1]},      {"w", &cs.ic[2]},      {"thp", &cs.ic[3]},
      {"pip", &cs.ic[4]},    {"qv", &cs.ic[5]},     {"qc", &cs.ic[6]},
      {"qr", &cs.ic[7]},     {"bdy_wgt", &cs.bdy.wgt}, {"sfc_tsk", &cs.sfc_tsk},
      {"sfc_land", &cs.sfc_land}, {"sfc_lat", &cs.sfc_lat}, {"sfc_lon", &cs.sfc_lon},
      {"sfc_soilw", &cs.sfc_soilw}};
  for (const auto& e : t)
    if (n == e.first) return e.second;
  static const char* vn[5] = {"u", "v", "thp", "pip", "qv"};
  for (int v = 0; v < 5; ++v) {
    if (n == std::string("bdy_lo_") + vn[v]) return &cs.bdy.lo[v];
    if (n == std::string("bdy_hi_") + vn[v]) return &cs.bdy.hi[v];
  }
  return nullptr;
}

}  // namespace
}  // namespace orcb

struct orc_case {
  orcb::Case c;
};

extern "C" {

orc_case* orc_case_load(const char* ini, int n_over, const char* const* over) {
  auto* p = new orc_case();
  orcb::Case& cs = p->c;
  if (!ini || !cs.cfg.load(ini)) {
    delete p;
    return nullptr;
  }
  for (int a = 0; a < n_over; ++a) {  // main.cpp:255-259
    std::string s = over[a];
    size_t eq = s.find('=');
    if (eq != std::string::npos) cs.cfg.kv[s.substr(0, eq)] = s.substr(eq + 1);
  }
  if (!orcb::build_case(cs)) {
    delete p;
    return nullptr;
  }
  return p;
}

void orc_case_destroy(orc_case* c) { delete c; }

int orc_case_dims(const orc_case* c, int* dims3, double* d3) {
  if (!c) return -1;
  dims3[0] = c->c.g.nx;
  dims3[1] = c->c.g.ny;
  dims3[2] = c->c.g.nz;
  d3[0] = c->c.g.dx;
  d3[1] = c->c.g.dy;
  d3[2] = c->c.g.dz;
  return 0;
}

int orc_case_value(const orc_case* c, const char* name, double* out) {
  if (!c || !name || !out) return -1;
  auto it = c->c.scalars.find(name);
  if (it != c->c.scalars.end()) {
    *out = it->second;
    return 0;
  }
  it = c->c.params.find(name);
  if (it != c->c.params.end()) {
    *out = it->second;
    return 0;
  }
  if (std::string(name) == "file_mode") {
    *out = c->c.file_mode;
    return 0;
  }
  return -1;
}

long orc_case_array_size(const orc_case* c, const char* name) {
  if (!c || !name) return 0;
  const std::vector<double>* v = orcb::array_of(c->c, name);
  return v ? (long)v->size() : 0;
}

int orc_case_array(const orc_case* c, const char* name, double* out, long n) {
  if (!c || !name || !out) return -1;
  const std::vector<double>* v = orcb::array_of(c->c, name);
  if (!v || (long)v->size() != n) return -1;
  std::memcpy(out, v->data(), sizeof(double) * n);
  return 0;
}

int orc_case_update_bdy(orc_case* c, orc_model* m, double t) {
  if (!c || !m) return -1;
  orcb::Case& cs = c->c;
  if (!cs.file_mode) return 0;
  if (!cs.bdy.update(cs.g, cs.input, cs.base, t)) return -2;
  if (cs.bdy.changed) {
    for (int v = 0; v < 5; ++v) {
      char lo[16], hi[16];
      std::snprintf(lo, sizeof lo, "bdy_lo_%d", v);
      std::snprintf(hi, sizeof hi, "bdy_hi_%d", v);
      if (orc_set_array(m, lo, cs.bdy.lo[v].data(), (long)cs.bdy.lo[v].size()) != 0) return -3;
      if (orc_set_array(m, hi, cs.bdy.hi[v].data(), (long)cs.bdy.hi[v].size()) != 0) return -3;
    }
    cs.bdy.changed = false;
  }
  return orc_set_bdy(m, 1, cs.bdy.tfrac(t));
}

}  // extern "C"
