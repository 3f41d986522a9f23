#include "core/metric.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <stdexcept>
#include <string>

namespace wfe {

namespace {

// Heights of the interior w faces 0..nz of the vertical coordinate: uniform
// spacing dz, or a geometric progression of layer depths from dz0 growing by
// dz_ratio per layer up to dz_max (metric.cpp:19-38 of the reference).
std::vector<double> interior_faces(const GDims& g, const Config& cfg) {
  std::vector<double> z(g.nz + 1, 0.0);
  if (cfg.get_str("stretch", "none") == "geometric") {
    double layer = cfg.get_real("dz0", g.dz);
    const double growth = cfg.get_real("dz_ratio", 1.05);
    const double cap = cfg.get_real("dz_max", g.dz * 4);
    for (int k = 1; k <= g.nz; ++k) {
      z[k] = z[k - 1] + layer;
      layer = std::min(layer * growth, cap);
    }
  } else {
    for (int k = 0; k <= g.nz; ++k) z[k] = (double)k * g.dz;
  }
  return z;
}

// Face k of the padded column (k may lie in the ghost layers): beyond the
// interior, faces continue with the depth of the outermost interior layer.
double padded_face(const std::vector<double>& z, int k) {
  const int top = (int)z.size() - 1;
  if (k < 0) return z[0] + (double)k * (z[1] - z[0]);
  if (k > top) return z[top] + (double)(k - top) * (z[top] - z[top - 1]);
  return z[k];
}

using Surface = std::function<double(double, double)>;

// Analytic terrain heights h(x, y) of the idealised cases (metric.cpp:63-84):
// Schaer's cos^2-modulated Gaussian ridge in x, or an elliptic Gaussian hill;
// flat otherwise. Every mountain key is read, as the reference reads them.
Surface analytic_surface(const GDims& g, const Config& cfg, const std::string& kind) {
  const double h0 = cfg.get_real("mtn_h0", 250);
  const double ax = cfg.get_real("mtn_a", 5000);
  const double wavelength = cfg.get_real("mtn_lambda", 4000);
  const double xc = cfg.get_real("mtn_xc", g.nx * g.dx / 2);
  const double yc = cfg.get_real("mtn_yc", g.ny * g.dy / 2);
  const double ay = cfg.get_real("mtn_ya", ax);
  constexpr double kPi = 3.14159265358979323846;
  if (kind == "schaer")
    return [=](double x, double) {
      const double dxr = x - xc;
      const double envelope = std::exp(-(dxr / ax) * (dxr / ax));
      const double wave = std::cos(kPi * dxr / wavelength);
      return h0 * envelope * wave * wave;
    };
  if (kind == "gauss")
    return [=](double x, double y) {
      const double sx = (x - xc) / ax, sy = (y - yc) / ay;
      return h0 * std::exp(-sx * sx - sy * sy);
    };
  return [](double, double) { return 0.0; };
}

// Interior (ny * nx) field spread over the padded plane, each ghost cell
// taking the value of the nearest interior cell.
void spread_to_padded(const GDims& g, const std::vector<double>& interior,
                      std::vector<double>& padded) {
  for (int j = -g.ng; j < g.ny + g.ng; ++j) {
    const int js = std::clamp(j, 0, g.ny - 1);
    for (int i = -g.ng; i < g.nx + g.ng; ++i)
      padded[g.idx2(i, j)] = interior[(size_t)js * g.nx + std::clamp(i, 0, g.nx - 1)];
  }
}

}  // namespace

void Metric::build(const GDims& g, const Config& cfg, const std::vector<double>* h_file,
                   const std::vector<double>* fcor_file, const std::vector<double>* mapf_file) {
  // --- vertical coordinate
  const std::vector<double> faces = interior_faces(g, cfg);
  zt = faces[g.nz];
  const size_t n1 = (size_t)g.NZ;
  zeta_w.assign(n1, 0.0);
  zeta_c.assign(n1, 0.0);
  dzeta_c.assign(n1, 0.0);
  dzeta_w.assign(n1, 0.0);
  for (int kk = 0; kk < g.NZ; ++kk) {
    const double lo = padded_face(faces, kk - g.ng), hi = padded_face(faces, kk - g.ng + 1);
    zeta_w[kk] = lo;
    zeta_c[kk] = 0.5 * (lo + hi);
  }
  // layer depths at centres (face to face above) and at faces (centre to
  // centre below); the outermost entries repeat their neighbour
  for (int kk = 0; kk + 1 < g.NZ; ++kk) dzeta_c[kk] = zeta_w[kk + 1] - zeta_w[kk];
  dzeta_c[n1 - 1] = dzeta_c[n1 - 2];
  for (int kk = 1; kk < g.NZ; ++kk) dzeta_w[kk] = zeta_c[kk] - zeta_c[kk - 1];
  dzeta_w[0] = dzeta_w[1];

  // --- terrain and its slopes on the padded plane
  const size_t n2 = g.n2();
  for (std::vector<double>* a : {&h, &jac, &hx, &hy, &hx_u, &hy_v}) a->assign(n2, 0.0);
  const std::string kind = cfg.get_str("terrain", "none");
  const Surface surface = analytic_surface(g, cfg, kind);
  if (kind == "file") {
    if (!h_file) throw std::runtime_error("terrain=file but no terrain data was provided");
    spread_to_padded(g, *h_file, h);
    // slopes from neighbouring cells: centred inside the padded plane,
    // one-sided on its outer rows and columns (no face slope there)
    const int i_lo = -g.ng, i_hi = g.nx + g.ng - 1, j_lo = -g.ng, j_hi = g.ny + g.ng - 1;
    for (int j = j_lo; j <= j_hi; ++j)
      for (int i = i_lo; i <= i_hi; ++i) {
        const size_t c = g.idx2(i, j);
        const int w = std::max(i - 1, i_lo), e = std::min(i + 1, i_hi);
        const int s = std::max(j - 1, j_lo), n = std::min(j + 1, j_hi);
        jac[c] = (zt - h[c]) / zt;
        hx[c] = (h[g.idx2(e, j)] - h[g.idx2(w, j)]) / ((double)(e - w) * g.dx);
        hy[c] = (h[g.idx2(i, n)] - h[g.idx2(i, s)]) / ((double)(n - s) * g.dy);
        hx_u[c] = w == i ? 0.0 : (h[c] - h[g.idx2(w, j)]) / g.dx;
        hy_v[c] = s == j ? 0.0 : (h[c] - h[g.idx2(i, s)]) / g.dy;
      }
  } else {
    // analytic slopes by differences of the surface one grid length apart:
    // centred at cell centres, backward at the u (i - 1/2) and v (j - 1/2)
    // faces
    for (int j = -g.ng; j < g.ny + g.ng; ++j) {
      const double y = ((double)j + 0.5) * g.dy;
      for (int i = -g.ng; i < g.nx + g.ng; ++i) {
        const double x = ((double)i + 0.5) * g.dx;
        const size_t c = g.idx2(i, j);
        const double here = surface(x, y);
        h[c] = here;
        jac[c] = (zt - here) / zt;
        hx[c] = (surface(x + g.dx, y) - surface(x - g.dx, y)) / (2.0 * g.dx);
        hy[c] = (surface(x, y + g.dy) - surface(x, y - g.dy)) / (2.0 * g.dy);
        hx_u[c] = (here - surface(x - g.dx, y)) / g.dx;
        hy_v[c] = (here - surface(x, y - g.dy)) / g.dy;
      }
    }
  }

  // --- Coriolis parameter and map factor: constant, or from the input
  fcor.assign(n2, cfg.get_real("coriolis_f", 0));
  if (fcor_file) spread_to_padded(g, *fcor_file, fcor);
  mapf.assign(n2, 1.0);
  if (mapf_file) spread_to_padded(g, *mapf_file, mapf);
}

}  // namespace wfe
