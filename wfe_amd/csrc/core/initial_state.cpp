// Initial states of the cases (the idealised fields of src/main.cpp:28-54,
// 146-195 and the file-mode conversion of src/main.cpp:58-100), the seeded
// ensemble perturbation (src/main.cpp:105-144) and the Lambert map factor
// (src/main.cpp:323-340).
#include <algorithm>
#include <cmath>
#include <random>

#include "core/case.hpp"

namespace wfe {

namespace {

// A cos^2 blob amp * cos^2(pi r / 2), r < 1 being the ellipsoidal distance
// from (xc, yc, zc) in units of the radii, sampled at the cell centres in
// physical height. Keys are <prefix>xc ... <prefix>zr with the given
// defaults; the interior of `out` is cleared first.
struct BlobDefaults {
  double xc, yc, zc, xr, yr, zr;
};

void cos2_blob(const Case& cs, const char* prefix, double amp, const BlobDefaults& d,
               std::vector<double>& out) {
  const GDims& g = cs.g;
  const std::string p = prefix;
  const double xc = cs.cfg.get_real(p + "xc", d.xc), yc = cs.cfg.get_real(p + "yc", d.yc);
  const double zc = cs.cfg.get_real(p + "zc", d.zc), xr = cs.cfg.get_real(p + "xr", d.xr);
  const double yr = cs.cfg.get_real(p + "yr", d.yr), zr = cs.cfg.get_real(p + "zr", d.zr);
  constexpr double kPi = 3.14159265358979323846;
  std::fill(out.begin(), out.end(), 0.0);
  for (int k = 0; k < g.nz; ++k) {
    const double zeta = cs.metric.zeta_c[k + g.ng];
    for (int j = 0; j < g.ny; ++j) {
      const double ry = (((double)j + 0.5) * g.dy - yc) / yr;
      for (int i = 0; i < g.nx; ++i) {
        const double rx = (((double)i + 0.5) * g.dx - xc) / xr;
        const double rz = (cs.metric.z_at(g, i, j, zeta) - zc) / zr;
        const double r = std::sqrt(rx * rx + ry * ry + rz * rz);
        if (!(r < 1.0)) continue;
        const double c = std::cos(kPi * r / 2);
        out[g.idx(i, j, k)] = amp * c * c;
      }
    }
  }
}

// The base wind on its faces at the interior levels: u on faces 0..nx, v
// on faces 0..ny; levels with a zero base wind keep +0.
void base_wind_state(const Case& cs, std::vector<double>& u, std::vector<double>& v) {
  const GDims& g = cs.g;
  std::fill(u.begin(), u.end(), 0.0);
  for (int k = 0; k < g.nz; ++k) {
    const double ub = cs.base.ub[k + g.ng], vb = cs.base.vb[k + g.ng];
    for (int j = 0; j <= g.ny; ++j)
      for (int i = 0; i <= g.nx; ++i) {
        if (ub != 0.0 && j < g.ny) u[g.idx(i, j, k)] = ub;
        if (vb != 0.0 && i < g.nx) v[g.idx(i, j, k)] = vb;
      }
  }
}

}  // namespace

void idealised_initial_state(Case& cs) {
  const GDims& g = cs.g;
  const Config& cfg = cs.cfg;
  // warm/cold theta' bubble
  const double dth = cfg.get_real("bubble_dtheta", 0.0);
  if (dth != 0.0)
    cos2_blob(cs, "bubble_", dth,
              {g.nx * g.dx / 2, g.ny * g.dy / 2, 2000.0, 2000.0, 2000.0, 2000.0}, cs.ic[3]);
  base_wind_state(cs, cs.ic[0], cs.ic[1]);
  // vapour: the base-state profile of the WK82 sounding (zero for other
  // moist profiles) ...
  if (cs.base.has_moisture) {
    std::vector<double>& qv = cs.ic[5];
    std::fill(qv.begin(), qv.end(), 0.0);
    if (cs.base.keep_qvb_ic)
      for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j)
          for (int i = 0; i < g.nx; ++i) qv[g.idx(i, j, k)] = cs.base.qvb[g.idx(i, j, k)];
  }
  // ... replaced by a sharp vapour blob when one is configured
  const double dq = cfg.get_real("qvbub_dq", 0.0);
  if (dq != 0.0)
    cos2_blob(cs, "qvbub_", dq, {g.nx * g.dx / 4, g.ny * g.dy / 2, 2000.0, 2000.0, 1e9, 2000.0},
              cs.ic[5]);
}

std::array<std::vector<double>, 5> perturbation_state(const GDims& g,
                                                      const std::vector<double>& thb,
                                                      const std::vector<double>& pib,
                                                      const std::vector<double>* raw) {
  // raw: u, v, theta, pi, qv at the cell centres, nz * ny * nx, i fastest
  auto at = [&](int f, int i, int j, int k) { return raw[f][((size_t)k * g.ny + j) * g.nx + i]; };
  std::array<std::vector<double>, 5> out;
  for (auto& a : out) a.assign(g.npts(), 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j) {
      // u on the west faces 0..nx from the neighbouring centres (the edge
      // faces repeat the edge centre), v on the south faces 0..ny-1
      for (int i = 0; i <= g.nx; ++i)
        out[0][g.idx(i, j, k)] =
            0.5 * (at(0, std::max(i - 1, 0), j, k) + at(0, std::min(i, g.nx - 1), j, k));
      for (int i = 0; i < g.nx; ++i) {
        const size_t c = g.idx(i, j, k);
        out[1][c] = 0.5 * (at(1, i, std::max(j - 1, 0), k) + at(1, i, j, k));
        out[2][c] = at(2, i, j, k) - thb[c];
        out[3][c] = at(3, i, j, k) - pib[c];
        out[4][c] = at(4, i, j, k);
      }
    }
  return out;
}

std::vector<double> lambert_map_factor(const std::vector<double>& lat_deg, double lat1_deg,
                                       double lat2_deg) {
  // constants at the precision the reference writes them
  constexpr double kDeg = 0.01745329252, kQuarterPi = 0.78539816339;
  const double phi1 = lat1_deg * kDeg, phi2 = lat2_deg * kDeg;
  auto t = [&](double phi) { return std::tan(kQuarterPi + phi / 2); };
  const double cone = std::fabs(phi1 - phi2) < 1e-6
                          ? std::sin(phi1)
                          : std::log(std::cos(phi1) / std::cos(phi2)) / std::log(t(phi2) / t(phi1));
  const double scale = std::cos(phi1) * std::pow(t(phi1), cone);
  std::vector<double> m(lat_deg.size());
  std::transform(lat_deg.begin(), lat_deg.end(), m.begin(), [&](double lat) {
    const double phi = lat * kDeg;
    return scale / (std::pow(t(phi), cone) * std::cos(phi));
  });
  return m;
}

void add_ic_perturbation(const GDims& g, std::vector<double>& thp, double amp, unsigned seed) {
  if (amp <= 0.0) return;
  // white noise from the seeded stream, interior cells in k, j, i order
  std::vector<double> noise(g.npts(), 0.0);
  std::mt19937 stream(seed);
  std::normal_distribution<float> normal(0.f, 1.f);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) noise[g.idx(i, j, k)] = (double)normal(stream);
  // six Jacobi passes of the 7-point filter: half the cell plus a twelfth
  // of each of its six neighbours, on cells one away from every face
  const long sx = 1, sy = g.NX, sz = (long)g.NX * g.NY;
  for (int pass = 0; pass < 6; ++pass) {
    const std::vector<double> prev = noise;
    for (int k = 1; k < g.nz - 1; ++k)
      for (int j = 1; j < g.ny - 1; ++j)
        for (int i = 1; i < g.nx - 1; ++i) {
          const size_t c = g.idx(i, j, k);
          const double nb = prev[c - sx] + prev[c + sx] + prev[c - sy] + prev[c + sy] +
                            prev[c - sz] + prev[c + sz];
          noise[c] = 0.5 * prev[c] + nb / 12.0;
        }
  }
  // scale to the requested interior RMS
  double sum_sq = 0;
  size_t count = 0;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i, ++count) {
        const double v = noise[g.idx(i, j, k)];
        sum_sq += v * v;
      }
  const double rms = std::sqrt(sum_sq / count);
  if (rms < 1e-9) return;
  const double gain = amp / rms;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) thp[g.idx(i, j, k)] += gain * noise[g.idx(i, j, k)];
}

}  // namespace wfe
