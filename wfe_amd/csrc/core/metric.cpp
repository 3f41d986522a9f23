#include "core/metric.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace wfe {

namespace {
int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Edge-clamped copy of an interior (ny*nx) field into a padded 2-D array.
void pad_clamped(const GDims& g, const std::vector<double>& src, std::vector<double>& dst) {
  for (int j = -g.ng; j < g.ny + g.ng; ++j)
    for (int i = -g.ng; i < g.nx + g.ng; ++i) {
      int ic = clampi(i, 0, g.nx - 1), jc = clampi(j, 0, g.ny - 1);
      dst[g.idx2(i, j)] = src[(size_t)jc * g.nx + ic];
    }
}
}  // namespace

// Follows src/core/metric.cpp:12-168 (vertical levels :19-56, terrain :58-123,
// Coriolis :125-134, map factor :136-145).
void Metric::build(const GDims& g, const Config& cfg, const std::vector<double>* h_file,
                   const std::vector<double>* fcor_file, const std::vector<double>* mapf_file) {
  const size_t n1 = (size_t)g.NZ, n2 = g.n2();

  // Vertical zeta levels: uniform, or geometric growth dz0 -> dz_max.
  zeta_w.assign(n1, 0.0);
  zeta_c.assign(n1, 0.0);
  dzeta_c.assign(n1, 0.0);
  dzeta_w.assign(n1, 0.0);
  std::vector<double> zw(g.nz + 1, 0.0);
  const std::string stretch = cfg.get_str("stretch", "none");
  if (stretch == "geometric") {
    double d = cfg.get_real("dz0", g.dz);
    const double ratio = cfg.get_real("dz_ratio", 1.05);
    const double dzmax = cfg.get_real("dz_max", g.dz * 4);
    for (int k = 1; k <= g.nz; ++k) {
      zw[k] = zw[k - 1] + d;
      d = std::min(d * ratio, dzmax);
    }
  } else {
    for (int k = 0; k <= g.nz; ++k) zw[k] = (double)k * g.dz;
  }
  zt = zw[g.nz];
  // Ghost levels extend the boundary spacing linearly.
  auto zw_ext = [&](int k) -> double {
    if (k < 0) return zw[0] + (double)k * (zw[1] - zw[0]);
    if (k > g.nz) return zw[g.nz] + (double)(k - g.nz) * (zw[g.nz] - zw[g.nz - 1]);
    return zw[k];
  };
  for (int kk = 0; kk < (int)n1; ++kk) {
    const int k = kk - g.ng;
    zeta_w[kk] = zw_ext(k);
    zeta_c[kk] = 0.5 * (zw_ext(k) + zw_ext(k + 1));
  }
  for (int kk = 0; kk + 1 < (int)n1; ++kk) dzeta_c[kk] = zeta_w[kk + 1] - zeta_w[kk];
  dzeta_c[n1 - 1] = dzeta_c[n1 - 2];
  for (int kk = 1; kk < (int)n1; ++kk) dzeta_w[kk] = zeta_c[kk] - zeta_c[kk - 1];
  dzeta_w[0] = dzeta_w[1];

  // Terrain on the padded 2-D grid.
  h.assign(n2, 0.0);
  jac.assign(n2, 0.0);
  hx.assign(n2, 0.0);
  hy.assign(n2, 0.0);
  hx_u.assign(n2, 0.0);
  hy_v.assign(n2, 0.0);
  const std::string terrain = cfg.get_str("terrain", "none");
  const double h0 = cfg.get_real("mtn_h0", 250);
  const double xa = cfg.get_real("mtn_a", 5000);
  const double lam = cfg.get_real("mtn_lambda", 4000);
  const double xc = cfg.get_real("mtn_xc", g.nx * g.dx / 2);
  const double yc = cfg.get_real("mtn_yc", g.ny * g.dy / 2);
  const double ya = cfg.get_real("mtn_ya", xa);
  const double pi = 3.14159265358979323846;
  auto h_of = [&](double x, double y) -> double {
    if (terrain == "schaer") {
      const double xr = x - xc;
      const double env = std::exp(-(xr / xa) * (xr / xa));
      const double c = std::cos(pi * xr / lam);
      return h0 * env * c * c;
    }
    if (terrain == "gauss") {
      const double xr = (x - xc) / xa, yr = (y - yc) / ya;
      return h0 * std::exp(-xr * xr - yr * yr);
    }
    return 0.0;
  };

  if (terrain == "file") {
    if (!h_file) throw std::runtime_error("terrain=file but no terrain data was provided");
    pad_clamped(g, *h_file, h);
    // Derivatives by differences, one-sided at the padded edges.
    for (int j = -g.ng; j < g.ny + g.ng; ++j)
      for (int i = -g.ng; i < g.nx + g.ng; ++i) {
        const size_t c = g.idx2(i, j);
        jac[c] = (zt - h[c]) / zt;
        const int im = i > -g.ng ? i - 1 : i, ip = i < g.nx + g.ng - 1 ? i + 1 : i;
        const int jm = j > -g.ng ? j - 1 : j, jp = j < g.ny + g.ng - 1 ? j + 1 : j;
        hx[c] = (h[g.idx2(ip, j)] - h[g.idx2(im, j)]) / ((double)(ip - im) * g.dx);
        hy[c] = (h[g.idx2(i, jp)] - h[g.idx2(i, jm)]) / ((double)(jp - jm) * g.dy);
        hx_u[c] = (i == im) ? 0.0 : (h[c] - h[g.idx2(im, j)]) / g.dx;
        hy_v[c] = (j == jm) ? 0.0 : (h[c] - h[g.idx2(i, jm)]) / g.dy;
      }
  } else {
    for (int j = -g.ng; j < g.ny + g.ng; ++j)
      for (int i = -g.ng; i < g.nx + g.ng; ++i) {
        const double x = ((double)i + 0.5) * g.dx;
        const double y = ((double)j + 0.5) * g.dy;
        const size_t c = g.idx2(i, j);
        h[c] = h_of(x, y);
        jac[c] = (zt - h[c]) / zt;
        hx[c] = (h_of(x + g.dx, y) - h_of(x - g.dx, y)) / (2.0 * g.dx);
        hy[c] = (h_of(x, y + g.dy) - h_of(x, y - g.dy)) / (2.0 * g.dy);
        hx_u[c] = (h_of(x, y) - h_of(x - g.dx, y)) / g.dx;  // u point (i - 1/2)
        hy_v[c] = (h_of(x, y) - h_of(x, y - g.dy)) / g.dy;  // v point (j - 1/2)
      }
  }

  fcor.assign(n2, cfg.get_real("coriolis_f", 0));
  if (fcor_file) pad_clamped(g, *fcor_file, fcor);
  mapf.assign(n2, 1.0);
  if (mapf_file) pad_clamped(g, *mapf_file, mapf);
}

}  // namespace wfe
