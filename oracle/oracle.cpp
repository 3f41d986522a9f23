// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.
// CPU FP64 restatement of the WFE large-step dynamics hot path.
// TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP path and
// the CPU baseline. Every function cites the reference file:line it follows.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

using Vec = std::vector<double>;

// Physical constants (src/core/constants.hpp:7-17).
constexpr double kGrav = 9.81, kCp = 1004.5, kRd = 287.04, kCv = kCp - kRd, kP00 = 1.0e5;
constexpr double kLv = 2.5e6, kLf = 3.34e5, kLs = kLv + kLf, kRv = 461.5, kEps61 = 0.61,
                 kTfrz = 273.15;

struct Grid {  // src/core/grid.hpp:22-36
  int nx, ny, nz, ng, NX, NY, NZ;
  double dx, dy, dz;
  long id(int i, int j, int k) const { return ((long)(k + ng) * NY + (j + ng)) * NX + (i + ng); }
  long id2(int i, int j) const { return (long)(j + ng) * NX + (i + ng); }
  long n3() const { return (long)NX * NY * NZ; }
  long n2() const { return (long)NX * NY; }
};

struct Params {  // src/dynamics/params.hpp:8-25
  double diff_K = 0;
  int coriolis_use_ub = 1;
  double rayleigh_zd = -1, rayleigh_alpha = 0;
  int acoustic_ns = 6;
  double acoustic_beta = 0.2, acoustic_smdiv = 0.1;
  int bc_x_open = 0, bc_y_open = 0;
  double cstar = 30;
  int moisture = 0, microphysics = 1, pd_moist = 0, w_damping = 0;
};

struct State {  // src/dynamics/state.hpp:13-46
  Vec u, v, w, thp, pip, qv, qc, qr;
  void alloc(long n) {
    for (Vec* f : fields()) f->assign(n, 0.0);
  }
  std::vector<Vec*> fields() { return {&u, &v, &w, &thp, &pip, &qv, &qc, &qr}; }
  void copy_from(const State& o) {
    u = o.u; v = o.v; w = o.w; thp = o.thp; pip = o.pip; qv = o.qv; qc = o.qc; qr = o.qr;
  }
  void swap(State& o) {
    u.swap(o.u); v.swap(o.v); w.swap(o.w); thp.swap(o.thp);
    pip.swap(o.pip); qv.swap(o.qv); qc.swap(o.qc); qr.swap(o.qr);
  }
};

// 5th-order upwind interface value (kernels.cu:51-57).
inline double iface5(double qm2, double qm1, double q0, double q1, double q2, double q3,
                     double vel) {
  const double c = 1.0 / 60.0;
  const double cen = c * (37.0 * (q0 + q1) - 8.0 * (qm1 + q2) + (qm2 + q3));
  const double upw = c * ((q3 - qm2) - 5.0 * (q2 - qm1) + 10.0 * (q1 - q0));
  return (vel >= 0.0) ? (cen - upw) : (cen + upw);
}

// Tetens saturation mixing ratios with the denominator floor (thermo.hpp:11-26).
inline double qsat_tetens(double p, double T) {
  const double es = 610.78 * std::exp(17.269 * (T - 273.16) / (T - 35.86));
  double den = p - es;
  if (den < 1.0) den = 1.0;
  return 0.622 * es / den;
}
inline double qsat_ice(double p, double T) {
  const double es = 610.78 * std::exp(21.875 * (T - 273.16) / (T - 7.66));
  double den = p - es;
  if (den < 1.0) den = 1.0;
  return 0.622 * es / den;
}

}  // namespace

struct orc_model {
  Grid g{};
  Params p{};
  double zt = 0;
  Vec thb, pib, rhob, dthbdz, thbw, rhobw, thvb, thvbw, qvb, ub, vb;
  Vec zeta_w, zeta_c, dzeta_c, dzeta_w, h, hx, hy, hxu, hyv, jac, fcor, mapf;
  State sn, sstage, swork, tend;
  Vec div, piprev, mfx, mfy, mfz, rain;
  Vec rtjx, rtjy, kt3, aw;
  Vec pfx, pfy, pfz, pratio;
  int bdy_active = 0;
  double tfrac = 0;
  Vec bdy_lo[5], bdy_hi[5], bdy_wgt;
  // surface/PBL physics (SfcPBL, src/physics/surface.hpp:19-43)
  int phys_on = 0, phys_nonlocal = 0, phys_doy = 1;
  double phys_start_hour = 0;
  Vec tsk, tdeep, land, lat, lon, soilw, cdv, pblh, t2m, u10, km, soilt;

  // ------------------------------------------------------------- helpers
  double ray_alpha(double zeta) const {  // kernels.cu:42-48
    if (p.rayleigh_zd < 0.0) return 0.0;
    if (zeta <= p.rayleigh_zd || zt <= p.rayleigh_zd) return 0.0;
    const double halfpi = 1.5707963267948966;
    const double s = std::sin(halfpi * (zeta - p.rayleigh_zd) / (zt - p.rayleigh_zd));
    return p.rayleigh_alpha * s * s;
  }

  // ------------------------------------------------- ghost fills (kernels.cu:695-801)
  void bc_z_zerograd(Vec& f) {  // kernels.cu:695-707
    const long stride = (long)g.NY * g.NX;
#pragma omp parallel for
    for (long col = 0; col < (long)g.NX * g.NY; ++col) {
      const double bot = f[col + stride * g.ng];
      const double top = f[col + stride * (g.nz - 1 + g.ng)];
      for (int m = 1; m <= g.ng; ++m) f[col + stride * (g.ng - m)] = bot;
      for (int m = 1; m <= g.ng + 1; ++m) f[col + stride * (g.nz - 1 + g.ng + m)] = top;
    }
  }
  void bc_z_w(Vec& f) {  // kernels.cu:710-723
    const long stride = (long)g.NY * g.NX;
#pragma omp parallel for
    for (long col = 0; col < (long)g.NX * g.NY; ++col) {
      auto at = [&](int k) -> double& { return f[col + stride * (k + g.ng)]; };
      at(g.nz) = 0;
      const double w0 = at(0);
      for (int m = 1; m <= g.ng; ++m) {
        at(-m) = 2 * w0 - at(m);
        at(g.nz + m) = -at(g.nz - m);
      }
    }
  }
  // Lateral x ghosts: periodic (kernels.cu:725-733) or zero-gradient with the
  // last valid index ilast (kernels.cu:748-758, kept inside the row).
  void bc_lateral_x(Vec& f, int ilast) {
#pragma omp parallel for
    for (int kr = 0; kr < g.NZ; ++kr)
      for (int jr = 0; jr < g.NY; ++jr) {
        const long row = ((long)kr * g.NY + jr) * g.NX;
        for (int m = 0; m < g.ng; ++m) {
          if (p.bc_x_open) {
            const double left = f[row + g.ng];
            const double right = f[row + g.ng + ilast];
            f[row + g.ng - 1 - m] = left;
            const long ir = g.ng + ilast + 1 + m;
            if (ir < g.NX) f[row + ir] = right;
          } else {
            f[row + g.ng + g.nx + m] = f[row + g.ng + m];
            f[row + g.ng - 1 - m] = f[row + g.ng + g.nx - 1 - m];
          }
        }
      }
  }
  void bc_lateral_y(Vec& f, int jlast) {  // kernels.cu:735-745, 760-770
#pragma omp parallel for
    for (int kr = 0; kr < g.NZ; ++kr) {
      const long plane = (long)kr * g.NY * g.NX;
      for (int m = 0; m < g.ng; ++m)
        for (int ir = 0; ir < g.NX; ++ir) {
          if (p.bc_y_open) {
            const double lo = f[plane + (long)g.ng * g.NX + ir];
            const double hi = f[plane + (long)(g.ng + jlast) * g.NX + ir];
            f[plane + (long)(g.ng - 1 - m) * g.NX + ir] = lo;
            const long jr = g.ng + jlast + 1 + m;
            if (jr < g.NY) f[plane + jr * g.NX + ir] = hi;
          } else {
            f[plane + (long)(g.ng + g.ny + m) * g.NX + ir] =
                f[plane + (long)(g.ng + m) * g.NX + ir];
            f[plane + (long)(g.ng - 1 - m) * g.NX + ir] =
                f[plane + (long)(g.ng + g.ny - 1 - m) * g.NX + ir];
          }
        }
    }
  }
  void apply_bcs(State& s) {  // kernels.cu:805-835
    bc_z_zerograd(s.u);
    bc_z_zerograd(s.v);
    bc_z_zerograd(s.thp);
    bc_z_zerograd(s.pip);
    bc_z_w(s.w);
    bc_lateral_x(s.u, g.nx);
    bc_lateral_x(s.v, g.nx - 1);
    bc_lateral_x(s.w, g.nx - 1);
    bc_lateral_x(s.thp, g.nx - 1);
    bc_lateral_x(s.pip, g.nx - 1);
    bc_lateral_y(s.u, g.ny - 1);
    bc_lateral_y(s.v, g.ny);
    bc_lateral_y(s.w, g.ny - 1);
    bc_lateral_y(s.thp, g.ny - 1);
    bc_lateral_y(s.pip, g.ny - 1);
    if (p.moisture) {
      for (Vec* f : {&s.qv, &s.qc, &s.qr}) {
        bc_z_zerograd(*f);
        bc_lateral_x(*f, g.nx - 1);
        bc_lateral_y(*f, g.ny - 1);
      }
    }
  }

  // --------------------------------------------- mass fluxes (kernels.cu:68-110)
  void mass_fluxes(const State& s) {
#pragma omp parallel for
    for (int k = 0; k < g.nz; ++k)
      for (int j = -2; j <= g.ny + 2; ++j)
        for (int i = -2; i <= g.nx + 2; ++i) {  // k_massflux_xy :68-81
          const long c = g.id(i, j, k);
          const double rjx = 0.5 * (rhob[g.id(i - 1, j, k)] * jac[g.id2(i - 1, j)] +
                                    rhob[c] * jac[g.id2(i, j)]);
          mfx[c] = rjx * s.u[c];
          const double rjy = 0.5 * (rhob[g.id(i, j - 1, k)] * jac[g.id2(i, j - 1)] +
                                    rhob[c] * jac[g.id2(i, j)]);
          mfy[c] = rjy * s.v[c];
        }
#pragma omp parallel for
    for (int k = 1; k < g.nz; ++k)
      for (int j = -2; j <= g.ny + 1; ++j)
        for (int i = -2; i <= g.nx + 1; ++i) {  // k_massflux_z :83-97
          const int kk = k + g.ng;
          const long c = g.id(i, j, k);
          const double fac = 1.0 - zeta_w[kk] / zt;
          const double uw = 0.25 * (s.u[g.id(i, j, k - 1)] + s.u[g.id(i + 1, j, k - 1)] +
                                    s.u[g.id(i, j, k)] + s.u[g.id(i + 1, j, k)]);
          const double vw = 0.25 * (s.v[g.id(i, j, k - 1)] + s.v[g.id(i, j + 1, k - 1)] +
                                    s.v[g.id(i, j, k)] + s.v[g.id(i, j + 1, k)]);
          const double cross = (uw * hx[g.id2(i, j)] + vw * hy[g.id2(i, j)]) * fac;
          mfz[c] = rhobw[c] * (s.w[c] - cross);
        }
  }
  void divergence() {  // k_divergence :99-110
#pragma omp parallel for
    for (int k = 0; k < g.nz; ++k)
      for (int j = -1; j <= g.ny; ++j)
        for (int i = -1; i <= g.nx; ++i) {
          const int kk = k + g.ng;
          div[g.id(i, j, k)] = (mfx[g.id(i + 1, j, k)] - mfx[g.id(i, j, k)]) / g.dx +
                               (mfy[g.id(i, j + 1, k)] - mfy[g.id(i, j, k)]) / g.dy +
                               (mfz[g.id(i, j, k + 1)] - mfz[g.id(i, j, k)]) / dzeta_c[kk];
        }
  }

  // ------------------------------------------------ scalar advection (:114-179)
  void tend_scalar(const Vec& q, Vec& out, bool rayleigh) {
#pragma omp parallel for
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const int kk = k + g.ng;
          auto Q = [&](int ii, int jj, int kz) { return q[g.id(ii, jj, kz)]; };
          const double fxL = mfx[g.id(i, j, k)], fxR = mfx[g.id(i + 1, j, k)];
          const double fyL = mfy[g.id(i, j, k)], fyR = mfy[g.id(i, j + 1, k)];
          const double fzB = mfz[g.id(i, j, k)], fzT = mfz[g.id(i, j, k + 1)];
          const double FxL = fxL * iface5(Q(i - 3, j, k), Q(i - 2, j, k), Q(i - 1, j, k),
                                          Q(i, j, k), Q(i + 1, j, k), Q(i + 2, j, k), fxL);
          const double FxR = fxR * iface5(Q(i - 2, j, k), Q(i - 1, j, k), Q(i, j, k),
                                          Q(i + 1, j, k), Q(i + 2, j, k), Q(i + 3, j, k), fxR);
          const double FyL = fyL * iface5(Q(i, j - 3, k), Q(i, j - 2, k), Q(i, j - 1, k),
                                          Q(i, j, k), Q(i, j + 1, k), Q(i, j + 2, k), fyL);
          const double FyR = fyR * iface5(Q(i, j - 2, k), Q(i, j - 1, k), Q(i, j, k),
                                          Q(i, j + 1, k), Q(i, j + 2, k), Q(i, j + 3, k), fyR);
          const double FzB = fzB * iface5(Q(i, j, k - 3), Q(i, j, k - 2), Q(i, j, k - 1),
                                          Q(i, j, k), Q(i, j, k + 1), Q(i, j, k + 2), fzB);
          const double FzT = fzT * iface5(Q(i, j, k - 2), Q(i, j, k - 1), Q(i, j, k),
                                          Q(i, j, k + 1), Q(i, j, k + 2), Q(i, j, k + 3), fzT);
          const double fdiv =
              (FxR - FxL) / g.dx + (FyR - FyL) / g.dy + (FzT - FzB) / dzeta_c[kk];
          const double rt = rhob[g.id(i, j, k)] * jac[g.id2(i, j)];
          double t = (-fdiv + Q(i, j, k) * div[g.id(i, j, k)]) / rt;
          if (rayleigh) t -= ray_alpha(zeta_c[kk]) * Q(i, j, k);
          out[g.id(i, j, k)] = t;
        }
  }

  // --------------------------------------------------- momentum (:308-425)
  void tend_u(const State& s, Vec& out) {  // k_tend_u :308-346
#pragma omp parallel for
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const int kk = k + g.ng;
          auto U = [&](int ii, int jj, int kz) { return s.u[g.id(ii, jj, kz)]; };
          const double fcL = 0.5 * (mfx[g.id(i - 1, j, k)] + mfx[g.id(i, j, k)]);
          const double fcR = 0.5 * (mfx[g.id(i, j, k)] + mfx[g.id(i + 1, j, k)]);
          const double FxL = fcL * iface5(U(i - 3, j, k), U(i - 2, j, k), U(i - 1, j, k),
                                          U(i, j, k), U(i + 1, j, k), U(i + 2, j, k), fcL);
          const double FxR = fcR * iface5(U(i - 2, j, k), U(i - 1, j, k), U(i, j, k),
                                          U(i + 1, j, k), U(i + 2, j, k), U(i + 3, j, k), fcR);
          const double fyB = 0.5 * (mfy[g.id(i - 1, j, k)] + mfy[g.id(i, j, k)]);
          const double fyT = 0.5 * (mfy[g.id(i - 1, j + 1, k)] + mfy[g.id(i, j + 1, k)]);
          const double FyB = fyB * iface5(U(i, j - 3, k), U(i, j - 2, k), U(i, j - 1, k),
                                          U(i, j, k), U(i, j + 1, k), U(i, j + 2, k), fyB);
          const double FyT = fyT * iface5(U(i, j - 2, k), U(i, j - 1, k), U(i, j, k),
                                          U(i, j + 1, k), U(i, j + 2, k), U(i, j + 3, k), fyT);
          const double fzB = 0.5 * (mfz[g.id(i - 1, j, k)] + mfz[g.id(i, j, k)]);
          const double fzT = 0.5 * (mfz[g.id(i - 1, j, k + 1)] + mfz[g.id(i, j, k + 1)]);
          const double FzB = fzB * iface5(U(i, j, k - 3), U(i, j, k - 2), U(i, j, k - 1),
                                          U(i, j, k), U(i, j, k + 1), U(i, j, k + 2), fzB);
          const double FzT = fzT * iface5(U(i, j, k - 2), U(i, j, k - 1), U(i, j, k),
                                          U(i, j, k + 1), U(i, j, k + 2), U(i, j, k + 3), fzT);
          const double fdiv =
              (FxR - FxL) / g.dx + (FyT - FyB) / g.dy + (FzT - FzB) / dzeta_c[kk];
          const double dv = 0.5 * (div[g.id(i - 1, j, k)] + div[g.id(i, j, k)]);
          const double ru = 0.5 * (rhob[g.id(i - 1, j, k)] * jac[g.id2(i - 1, j)] +
                                   rhob[g.id(i, j, k)] * jac[g.id2(i, j)]);
          const double vavg = 0.25 * (s.v[g.id(i - 1, j, k)] + s.v[g.id(i - 1, j + 1, k)] +
                                      s.v[g.id(i, j, k)] + s.v[g.id(i, j + 1, k)]);
          const double fc = 0.5 * (fcor[g.id2(i - 1, j)] + fcor[g.id2(i, j)]);
          // vb extension: f (v - v_ref) and Rayleigh toward v_b mirror u_b.
          const double vref = p.coriolis_use_ub ? vb[kk] : 0.0;
          const double relax = -ray_alpha(zeta_c[kk]) * (U(i, j, k) - ub[kk]);
          out[g.id(i, j, k)] = (-fdiv + U(i, j, k) * dv) / ru + fc * (vavg - vref) + relax;
        }
  }
  void tend_v(const State& s, Vec& out) {  // k_tend_v :348-385
#pragma omp parallel for
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const int kk = k + g.ng;
          auto V = [&](int ii, int jj, int kz) { return s.v[g.id(ii, jj, kz)]; };
          const double fcB = 0.5 * (mfy[g.id(i, j - 1, k)] + mfy[g.id(i, j, k)]);
          const double fcT = 0.5 * (mfy[g.id(i, j, k)] + mfy[g.id(i, j + 1, k)]);
          const double FyB = fcB * iface5(V(i, j - 3, k), V(i, j - 2, k), V(i, j - 1, k),
                                          V(i, j, k), V(i, j + 1, k), V(i, j + 2, k), fcB);
          const double FyT = fcT * iface5(V(i, j - 2, k), V(i, j - 1, k), V(i, j, k),
                                          V(i, j + 1, k), V(i, j + 2, k), V(i, j + 3, k), fcT);
          const double fxL = 0.5 * (mfx[g.id(i, j - 1, k)] + mfx[g.id(i, j, k)]);
          const double fxR = 0.5 * (mfx[g.id(i + 1, j - 1, k)] + mfx[g.id(i + 1, j, k)]);
          const double FxL = fxL * iface5(V(i - 3, j, k), V(i - 2, j, k), V(i - 1, j, k),
                                          V(i, j, k), V(i + 1, j, k), V(i + 2, j, k), fxL);
          const double FxR = fxR * iface5(V(i - 2, j, k), V(i - 1, j, k), V(i, j, k),
                                          V(i + 1, j, k), V(i + 2, j, k), V(i + 3, j, k), fxR);
          const double fzB = 0.5 * (mfz[g.id(i, j - 1, k)] + mfz[g.id(i, j, k)]);
          const double fzT = 0.5 * (mfz[g.id(i, j - 1, k + 1)] + mfz[g.id(i, j, k + 1)]);
          const double FzB = fzB * iface5(V(i, j, k - 3), V(i, j, k - 2), V(i, j, k - 1),
                                          V(i, j, k), V(i, j, k + 1), V(i, j, k + 2), fzB);
          const double FzT = fzT * iface5(V(i, j, k - 2), V(i, j, k - 1), V(i, j, k),
                                          V(i, j, k + 1), V(i, j, k + 2), V(i, j, k + 3), fzT);
          const double fdiv =
              (FxR - FxL) / g.dx + (FyT - FyB) / g.dy + (FzT - FzB) / dzeta_c[kk];
          const double dv = 0.5 * (div[g.id(i, j - 1, k)] + div[g.id(i, j, k)]);
          const double rv = 0.5 * (rhob[g.id(i, j - 1, k)] * jac[g.id2(i, j - 1)] +
                                   rhob[g.id(i, j, k)] * jac[g.id2(i, j)]);
          const double uavg = 0.25 * (s.u[g.id(i, j - 1, k)] + s.u[g.id(i + 1, j - 1, k)] +
                                      s.u[g.id(i, j, k)] + s.u[g.id(i + 1, j, k)]);
          const double fc = 0.5 * (fcor[g.id2(i, j - 1)] + fcor[g.id2(i, j)]);
          const double uref = p.coriolis_use_ub ? ub[kk] : 0.0;
          const double relax = -ray_alpha(zeta_c[kk]) * (V(i, j, k) - vb[kk]);
          out[g.id(i, j, k)] = (-fdiv + V(i, j, k) * dv) / rv - fc * (uavg - uref) + relax;
        }
  }
  void tend_w(const State& s, double dt, Vec& out) {  // k_tend_w :387-425
#pragma omp parallel for
    for (int k = 1; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const int kk = k + g.ng;
          auto W = [&](int ii, int jj, int kz) { return s.w[g.id(ii, jj, kz)]; };
          const double fxL = 0.5 * (mfx[g.id(i, j, k - 1)] + mfx[g.id(i, j, k)]);
          const double fxR = 0.5 * (mfx[g.id(i + 1, j, k - 1)] + mfx[g.id(i + 1, j, k)]);
          const double FxL = fxL * iface5(W(i - 3, j, k), W(i - 2, j, k), W(i - 1, j, k),
                                          W(i, j, k), W(i + 1, j, k), W(i + 2, j, k), fxL);
          const double FxR = fxR * iface5(W(i - 2, j, k), W(i - 1, j, k), W(i, j, k),
                                          W(i + 1, j, k), W(i + 2, j, k), W(i + 3, j, k), fxR);
          const double fyB = 0.5 * (mfy[g.id(i, j, k - 1)] + mfy[g.id(i, j, k)]);
          const double fyT = 0.5 * (mfy[g.id(i, j + 1, k - 1)] + mfy[g.id(i, j + 1, k)]);
          const double FyB = fyB * iface5(W(i, j - 3, k), W(i, j - 2, k), W(i, j - 1, k),
                                          W(i, j, k), W(i, j + 1, k), W(i, j + 2, k), fyB);
          const double FyT = fyT * iface5(W(i, j - 2, k), W(i, j - 1, k), W(i, j, k),
                                          W(i, j + 1, k), W(i, j + 2, k), W(i, j + 3, k), fyT);
          const double fzB = 0.5 * (mfz[g.id(i, j, k - 1)] + mfz[g.id(i, j, k)]);
          const double fzT = 0.5 * (mfz[g.id(i, j, k)] + mfz[g.id(i, j, k + 1)]);
          const double FzB = fzB * iface5(W(i, j, k - 3), W(i, j, k - 2), W(i, j, k - 1),
                                          W(i, j, k), W(i, j, k + 1), W(i, j, k + 2), fzB);
          const double FzT = fzT * iface5(W(i, j, k - 2), W(i, j, k - 1), W(i, j, k),
                                          W(i, j, k + 1), W(i, j, k + 2), W(i, j, k + 3), fzT);
          const double fdiv =
              (FxR - FxL) / g.dx + (FyT - FyB) / g.dy + (FzT - FzB) / dzeta_w[kk];
          const double dv = 0.5 * (div[g.id(i, j, k - 1)] + div[g.id(i, j, k)]);
          const double rw = rhobw[g.id(i, j, k)] * jac[g.id2(i, j)];
          double relax = -ray_alpha(zeta_w[kk]) * W(i, j, k);
          const double wk = W(i, j, k);
          if (p.w_damping) {
            const double cr = std::fabs(wk) * dt / (dzeta_w[kk] * jac[g.id2(i, j)]);
            if (cr > 1.0) relax -= 0.3 * (cr - 1.0) * wk / dt;
          }
          out[g.id(i, j, k)] = (-fdiv + wk * dv) / rw + relax;
        }
  }

  // Constant-K Laplacian added to a tendency (k_diffuse :430-443).
  void diffuse(const Vec& f, Vec& out, int kmin, int knum, bool on_w) {
    const double K = p.diff_K;
#pragma omp parallel for
    for (int k = kmin; k < kmin + knum; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const int kk = k + g.ng;
          const double dzl = (on_w ? dzeta_w[kk] : dzeta_c[kk]) * jac[g.id2(i, j)];
          const double c = f[g.id(i, j, k)];
          const double lap =
              (f[g.id(i + 1, j, k)] - 2 * c + f[g.id(i - 1, j, k)]) / (g.dx * g.dx) +
              (f[g.id(i, j + 1, k)] - 2 * c + f[g.id(i, j - 1, k)]) / (g.dy * g.dy) +
              (f[g.id(i, j, k + 1)] - 2 * c + f[g.id(i, j, k - 1)]) / (dzl * dzl);
          out[g.id(i, j, k)] += K * lap;
        }
  }

  // Positive-definite moisture transport, one species (kernels.cu:187-264, :855-869).
  void advect_pd(double dt, const Vec& q, Vec& out) {
#pragma omp parallel for
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i <= g.nx; ++i) {  // k_qflux_x :187-196
          auto Q = [&](int ii) { return q[g.id(ii, j, k)]; };
          const double f = mfx[g.id(i, j, k)];
          pfx[g.id(i, j, k)] =
              f * iface5(Q(i - 3), Q(i - 2), Q(i - 1), Q(i), Q(i + 1), Q(i + 2), f);
        }
#pragma omp parallel for
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j <= g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {  // k_qflux_y :198-207
          auto Q = [&](int jj) { return q[g.id(i, jj, k)]; };
          const double f = mfy[g.id(i, j, k)];
          pfy[g.id(i, j, k)] =
              f * iface5(Q(j - 3), Q(j - 2), Q(j - 1), Q(j), Q(j + 1), Q(j + 2), f);
        }
#pragma omp parallel for
    for (int k = 0; k <= g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {  // k_qflux_z :209-218
          auto Q = [&](int kz) { return q[g.id(i, j, kz)]; };
          const double f = mfz[g.id(i, j, k)];
          pfz[g.id(i, j, k)] =
              f * iface5(Q(k - 3), Q(k - 2), Q(k - 1), Q(k), Q(k + 1), Q(k + 2), f);
        }
#pragma omp parallel for
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {  // k_pd_ratio :221-238
          const int kk = k + g.ng;
          const double dzc = dzeta_c[kk];
          const double out_ = std::fmax(0.0, pfx[g.id(i + 1, j, k)]) / g.dx +
                              std::fmax(0.0, -pfx[g.id(i, j, k)]) / g.dx +
                              std::fmax(0.0, pfy[g.id(i, j + 1, k)]) / g.dy +
                              std::fmax(0.0, -pfy[g.id(i, j, k)]) / g.dy +
                              std::fmax(0.0, pfz[g.id(i, j, k + 1)]) / dzc +
                              std::fmax(0.0, -pfz[g.id(i, j, k)]) / dzc;
          const double avail =
              rhob[g.id(i, j, k)] * jac[g.id2(i, j)] * std::fmax(0.0, q[g.id(i, j, k)]);
          const double r = (dt * out_ > 1e-20) ? avail / (dt * out_) : 1.0;
          pratio[g.id(i, j, k)] = std::fmin(1.0, r);
        }
    bc_lateral_x(pratio, g.nx - 1);
    bc_lateral_y(pratio, g.ny - 1);
#pragma omp parallel for
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {  // k_pd_apply :241-264
          const int kk = k + g.ng;
          auto sx = [&](int i0) {
            const double f = pfx[g.id(i0, j, k)];
            return f * pratio[g.id(f >= 0.0 ? i0 - 1 : i0, j, k)];
          };
          auto sy = [&](int j0) {
            const double f = pfy[g.id(i, j0, k)];
            return f * pratio[g.id(i, f >= 0.0 ? j0 - 1 : j0, k)];
          };
          auto sz = [&](int k0) {
            const double f = pfz[g.id(i, j, k0)];
            return f * pratio[g.id(i, j, f >= 0.0 ? k0 - 1 : k0)];
          };
          const double fdiv = (sx(i + 1) - sx(i)) / g.dx + (sy(j + 1) - sy(j)) / g.dy +
                              (sz(k + 1) - sz(k)) / dzeta_c[kk];
          const double rt = rhob[g.id(i, j, k)] * jac[g.id2(i, j)];
          out[g.id(i, j, k)] = (-fdiv + q[g.id(i, j, k)] * div[g.id(i, j, k)]) / rt;
        }
  }

  // compute_tendencies (kernels.cu:872-913).
  void compute_tendencies(const State& s, double dt) {
    tend_u(s, tend.u);
    tend_v(s, tend.v);
    tend_w(s, dt, tend.w);
    tend_scalar(s.thp, tend.thp, true);
    if (p.moisture) {
      if (p.pd_moist) {
        advect_pd(dt, s.qv, tend.qv);
        advect_pd(dt, s.qc, tend.qc);
        advect_pd(dt, s.qr, tend.qr);
      } else {
        tend_scalar(s.qv, tend.qv, false);
        tend_scalar(s.qc, tend.qc, false);
        tend_scalar(s.qr, tend.qr, false);
      }
    }
    std::fill(tend.pip.begin(), tend.pip.end(), 0.0);
    if (p.diff_K > 0.0) {
      diffuse(s.u, tend.u, 0, g.nz, false);
      diffuse(s.v, tend.v, 0, g.nz, false);
      diffuse(s.w, tend.w, 1, g.nz - 1, true);
      diffuse(s.thp, tend.thp, 0, g.nz, false);
    }
  }

  // Davies relaxation (k_bdy_relax :292-304, bdy_relax :981-995).
  void bdy_relax(const State& s) {
    auto relax = [&](int imax, int jmax, const Vec& f, const Vec& lo, const Vec& hi, Vec& t) {
#pragma omp parallel for
      for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < jmax; ++j)
          for (int i = 0; i < imax; ++i) {
            const int iw = i < g.nx ? i : g.nx - 1;
            const int jw = j < g.ny ? j : g.ny - 1;
            const double w = bdy_wgt[g.id2(iw, jw)];
            if (w <= 0.0) continue;
            const long c = g.id(i, j, k);
            t[c] += w * ((1.0 - tfrac) * lo[c] + tfrac * hi[c] - f[c]);
          }
    };
    relax(g.nx + 1, g.ny, s.u, bdy_lo[0], bdy_hi[0], tend.u);
    relax(g.nx, g.ny + 1, s.v, bdy_lo[1], bdy_hi[1], tend.v);
    relax(g.nx, g.ny, s.thp, bdy_lo[2], bdy_hi[2], tend.thp);
    relax(g.nx, g.ny, s.pip, bdy_lo[3], bdy_hi[3], tend.pip);
    relax(g.nx, g.ny, s.qv, bdy_lo[4], bdy_hi[4], tend.qv);
  }

  // Constant acoustic coefficients (k_acou_coef_c/f :452-474).
  void precompute_acoustic_coef() {
    const long n = g.n3();
    rtjx.assign(n, 0.0);
    rtjy.assign(n, 0.0);
    kt3.assign(n, 0.0);
    aw.assign(n, 0.0);
    Vec rtjc(n, 0.0);
    for (int kr = 0; kr < g.NZ; ++kr)
      for (int jr = 0; jr < g.NY; ++jr)
        for (int ir = 0; ir < g.NX; ++ir) {
          const long c = ((long)kr * g.NY + jr) * g.NX + ir;
          const double J = jac[(long)jr * g.NX + ir];
          rtjc[c] = rhob[c] * thvb[c] * J;
          kt3[c] = kRd * pib[c] / (kCv * rhob[c] * thvb[c] * J);
          aw[c] = rhobw[c] * thvbw[c];
        }
    for (int kr = 0; kr < g.NZ; ++kr)
      for (int jr = 0; jr < g.NY; ++jr)
        for (int ir = 0; ir < g.NX; ++ir) {
          const long c = ((long)kr * g.NY + jr) * g.NX + ir;
          rtjx[c] = (ir > 0) ? 0.5 * (rtjc[c - 1] + rtjc[c]) : rtjc[c];
          rtjy[c] = (jr > 0) ? 0.5 * (rtjc[c - g.NX] + rtjc[c]) : rtjc[c];
        }
  }

  // Horizontal forward step of u, v (k_acou_uv :480-527).
  void acou_uv(double dtau, State& s) {
#pragma omp parallel for
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const int kk = k + g.ng;
          auto ps = [&](int ii, int jj, int kz) {
            const long c = g.id(ii, jj, kz);
            return s.pip[c] + p.acoustic_smdiv * (s.pip[c] - piprev[c]);
          };
          auto thv = [&](int ii, int jj) {
            const long c = g.id(ii, jj, k);
            double t = thb[c] + s.thp[c];
            if (p.moisture) t *= 1.0 + kEps61 * s.qv[c];
            return t;
          };
          const double ps0 = ps(i, j, k);
          const double dzc2 = zeta_c[kk + 1] - zeta_c[kk - 1];
          if (!(p.bc_x_open && i == 0)) {
            const double thbu = 0.5 * (thv(i - 1, j) + thv(i, j));
            const double mu = 0.5 * (mapf[g.id2(i - 1, j)] + mapf[g.id2(i, j)]);
            double grad = (ps0 - ps(i - 1, j, k)) / g.dx;
            const double zxu = hxu[g.id2(i, j)] * (1.0 - zeta_c[kk] / zt);
            if (zxu != 0.0) {
              const double ju = 0.5 * (jac[g.id2(i - 1, j)] + jac[g.id2(i, j)]);
              const double dpdz = 0.5 *
                                  ((ps(i - 1, j, k + 1) - ps(i - 1, j, k - 1)) +
                                   (ps(i, j, k + 1) - ps(i, j, k - 1))) /
                                  dzc2;
              grad -= zxu / ju * dpdz;
            }
            s.u[g.id(i, j, k)] += dtau * (-kCp * thbu * mu * grad + tend.u[g.id(i, j, k)]);
          }
          if (!(p.bc_y_open && j == 0)) {
            const double thbv = 0.5 * (thv(i, j - 1) + thv(i, j));
            const double mv = 0.5 * (mapf[g.id2(i, j - 1)] + mapf[g.id2(i, j)]);
            double grad = (ps0 - ps(i, j - 1, k)) / g.dy;
            const double zyv = hyv[g.id2(i, j)] * (1.0 - zeta_c[kk] / zt);
            if (zyv != 0.0) {
              const double jv = 0.5 * (jac[g.id2(i, j - 1)] + jac[g.id2(i, j)]);
              const double dpdz = 0.5 *
                                  ((ps(i, j - 1, k + 1) - ps(i, j - 1, k - 1)) +
                                   (ps(i, j, k + 1) - ps(i, j, k - 1))) /
                                  dzc2;
              grad -= zyv / jv * dpdz;
            }
            s.v[g.id(i, j, k)] += dtau * (-kCp * thbv * mv * grad + tend.v[g.id(i, j, k)]);
          }
        }
  }

  // Klemp-Wilhelmson radiation at open boundaries (:649-688).
  void bc_rad_u_x(double dtau, Vec& u) {
#pragma omp parallel for
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j) {
        const int kk = k + g.ng;
        const double u0 = u[g.id(0, j, k)];
        if (u0 >= 0.0) {
          u[g.id(0, j, k)] = ub[kk];
        } else {
          const double cb = u0 - p.cstar;
          u[g.id(0, j, k)] = u0 - dtau * cb * (u[g.id(1, j, k)] - u0) / g.dx;
        }
        const double un = u[g.id(g.nx, j, k)];
        if (un <= 0.0) {
          u[g.id(g.nx, j, k)] = ub[kk];
        } else {
          const double cb = un + p.cstar;
          u[g.id(g.nx, j, k)] = un - dtau * cb * (un - u[g.id(g.nx - 1, j, k)]) / g.dx;
        }
      }
  }
  void bc_rad_v_y(double dtau, Vec& v) {  // inflow -> v_b (0 in the reference)
#pragma omp parallel for
    for (int k = 0; k < g.nz; ++k)
      for (int i = 0; i < g.nx; ++i) {
        const int kk = k + g.ng;
        const double v0 = v[g.id(i, 0, k)];
        if (v0 >= 0.0) {
          v[g.id(i, 0, k)] = vb[kk];
        } else {
          const double cb = v0 - p.cstar;
          v[g.id(i, 0, k)] = v0 - dtau * cb * (v[g.id(i, 1, k)] - v0) / g.dy;
        }
        const double vn = v[g.id(i, g.ny, k)];
        if (vn <= 0.0) {
          v[g.id(i, g.ny, k)] = vb[kk];
        } else {
          const double cb = vn + p.cstar;
          v[g.id(i, g.ny, k)] = vn - dtau * cb * (vn - v[g.id(i, g.ny - 1, k)]) / g.dy;
        }
      }
  }

  // Vertically implicit w-pi' column solve (k_acou_wpi :531-645).
  void acou_wpi(double dtau, State& s) {
    const double c1 = (1.0 + p.acoustic_beta) * 0.5;
    const double c2 = (1.0 - p.acoustic_beta) * 0.5;
#pragma omp parallel
    {
      const int L = g.nz + 1;
      Vec P(L), E(L), cs(L), ds(L), Wn(L), Av(L), Sv(L);
#pragma omp for collapse(2)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const double J = jac[g.id2(i, j)];
          const double hxc = hx[g.id2(i, j)];
          const double hyc = hy[g.id2(i, j)];
          Av[0] = 0;
          Av[g.nz] = 0;
          Sv[0] = 0;
          Sv[g.nz] = 0;
          for (int k = 1; k < g.nz; ++k) {
            const int kk = k + g.ng;
            const long c = g.id(i, j, k);
            Av[k] = aw[c];
            const double fac = 1.0 - zeta_w[kk] / zt;
            const double uw = 0.25 * (s.u[g.id(i, j, k - 1)] + s.u[g.id(i + 1, j, k - 1)] +
                                      s.u[g.id(i, j, k)] + s.u[g.id(i + 1, j, k)]);
            const double vw = 0.25 * (s.v[g.id(i, j, k - 1)] + s.v[g.id(i, j + 1, k - 1)] +
                                      s.v[g.id(i, j, k)] + s.v[g.id(i, j + 1, k)]);
            Sv[k] = Av[k] * (uw * hxc + vw * hyc) * fac;
          }
          for (int k = 0; k < g.nz; ++k) {
            const int kk = k + g.ng;
            const long c = g.id(i, j, k);
            const double Kt = kt3[c];
            const double mc = mapf[g.id2(i, j)];
            const double Dh = mc * ((rtjx[g.id(i + 1, j, k)] * s.u[g.id(i + 1, j, k)] -
                                     rtjx[c] * s.u[g.id(i, j, k)]) /
                                        g.dx +
                                    (rtjy[g.id(i, j + 1, k)] * s.v[g.id(i, j + 1, k)] -
                                     rtjy[c] * s.v[g.id(i, j, k)]) /
                                        g.dy);
            const double dzc = dzeta_c[kk];
            const double wk = s.w[g.id(i, j, k)];
            const double wkp = s.w[g.id(i, j, k + 1)];
            P[k] = s.pip[c] + dtau * tend.pip[c] -
                   dtau * Kt *
                       (Dh - (Sv[k + 1] - Sv[k]) / dzc +
                        c2 * (Av[k + 1] * wkp - Av[k] * wk) / dzc);
            E[k] = dtau * Kt * c1 / dzc;
          }
          for (int k = 1; k < g.nz; ++k) {
            const int kk = k + g.ng;
            const long c = g.id(i, j, k);
            const long cm = g.id(i, j, k - 1);
            double thvw = thbw[c] + 0.5 * (s.thp[cm] + s.thp[c]);
            if (p.moisture) thvw *= 1.0 + kEps61 * 0.5 * (s.qv[cm] + s.qv[c]);
            const double G = dtau * kCp * thvw * c1 / (J * dzeta_w[kk]);
            const double A = -G * E[k - 1] * Av[k - 1];
            const double B = 1.0 + G * Av[k] * (E[k] + E[k - 1]);
            const double C = -G * E[k] * Av[k + 1];
            double buoy = kGrav * 0.5 * (s.thp[cm] + s.thp[c]) / thbw[c];
            if (p.moisture) {
              const double dqv = 0.5 * (s.qv[cm] + s.qv[c]) - 0.5 * (qvb[cm] + qvb[c]);
              const double qcr = 0.5 * (s.qc[cm] + s.qc[c]) + 0.5 * (s.qr[cm] + s.qr[c]);
              buoy += kGrav * (kEps61 * dqv - qcr);
            }
            const double RHS = s.w[c] + dtau * (buoy + tend.w[c]) -
                               dtau * kCp * thvw * c2 *
                                   (s.pip[g.id(i, j, k)] - s.pip[g.id(i, j, k - 1)]) /
                                   (J * dzeta_w[kk]);
            const double D = RHS - G * (P[k] - P[k - 1]);
            if (k == 1) {
              cs[k] = C / B;
              ds[k] = D / B;
            } else {
              const double mm = B - A * cs[k - 1];
              cs[k] = C / mm;
              ds[k] = (D - A * ds[k - 1]) / mm;
            }
          }
          Wn[g.nz] = 0;
          double wabove = 0;
          for (int k = g.nz - 1; k >= 1; --k) {
            const double wk = ds[k] - cs[k] * wabove;
            Wn[k] = wk;
            wabove = wk;
          }
          {
            const double us = 0.5 * (s.u[g.id(i, j, 0)] + s.u[g.id(i + 1, j, 0)]);
            const double vs = 0.5 * (s.v[g.id(i, j, 0)] + s.v[g.id(i, j + 1, 0)]);
            Wn[0] = us * hxc + vs * hyc;
          }
          for (int k = 0; k < g.nz; ++k) {
            const long c = g.id(i, j, k);
            const double oldpi = s.pip[c];
            s.pip[c] = P[k] - E[k] * (Av[k + 1] * Wn[k + 1] - Av[k] * Wn[k]);
            piprev[c] = oldpi;
            const double wc = 0.5 * (Wn[k] + Wn[k + 1]);
            s.thp[c] += dtau * (tend.thp[c] - wc * dthbdz[c]);
          }
          for (int k = 0; k < g.nz; ++k) s.w[g.id(i, j, k)] = Wn[k];
        }
    }
  }

  // One acoustic substep (acoustic_substep :929-965).
  void acoustic_substep(double dtau, State& s) {
    acou_uv(dtau, s);
    if (p.bc_x_open) bc_rad_u_x(dtau, s.u);
    if (p.bc_y_open) bc_rad_v_y(dtau, s.v);
    bc_lateral_x(s.u, g.nx);
    bc_lateral_y(s.u, g.ny - 1);
    bc_lateral_x(s.v, g.nx - 1);
    bc_lateral_y(s.v, g.ny);
    acou_wpi(dtau, s);
    bc_lateral_x(s.pip, g.nx - 1);
    bc_lateral_y(s.pip, g.ny - 1);
    bc_z_zerograd(s.pip);
  }

  // Moisture stage update over all padded points (k_stage_q :268-273).
  void update_moisture_stage(const State& s0, double dt, State& out) {
    const long n = g.n3();
    for (int q = 0; q < 3; ++q) {
      const Vec& a = q == 0 ? s0.qv : q == 1 ? s0.qc : s0.qr;
      const Vec& t = q == 0 ? tend.qv : q == 1 ? tend.qc : tend.qr;
      Vec& o = q == 0 ? out.qv : q == 1 ? out.qc : out.qr;
#pragma omp parallel for
      for (long c = 0; c < n; ++c) o[c] = a[c] + dt * t[c];
    }
  }

  // Kessler warm rain with the mixed-phase blend (k_kessler kessler.cu:22-139).
  void kessler(double dt, State& s) {
    const double k1 = 0.001, qc0 = 0.001;
#pragma omp parallel
    {
      const int L = g.nz + 1;
      Vec qrl(L), rhol(L), dzl(L), vr(L);
#pragma omp for collapse(2)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const double J = jac[g.id2(i, j)];
          auto fallfac = [&](long c) {
            const double T = (thb[c] + s.thp[c]) * (pib[c] + s.pip[c]);
            return T < kTfrz ? 0.35 : 1.0;
          };
          const double rho_sfc = rhob[g.id(i, j, 0)];
          double vrmax = 0, dzmin = 1e30;
          for (int k = 0; k < g.nz; ++k) {
            const int kk = k + g.ng;
            const long c = g.id(i, j, k);
            qrl[k] = s.qr[c] > 0.0 ? s.qr[c] : 0.0;
            rhol[k] = rhob[c];
            dzl[k] = dzeta_c[kk] * J;
            const double rq = rhol[k] * qrl[k];
            vr[k] = rq > 1e-9 ? 36.34 * std::pow(rq, 0.1364) * std::sqrt(rho_sfc / rhol[k]) *
                                    fallfac(c)
                              : 0.0;
            if (vr[k] > vrmax) vrmax = vr[k];
            if (dzl[k] < dzmin) dzmin = dzl[k];
          }
          const int nsub = 1 + (int)(vrmax * dt / (0.8 * dzmin));
          const double dts = dt / (double)nsub;
          double acc = 0;
          for (int sc = 0; sc < nsub; ++sc) {
            const double fb = rhol[0] * qrl[0] * vr[0];
            acc += dts * fb;
            for (int k = 0; k < g.nz; ++k) {
              const double ftop = (k + 1 < g.nz) ? rhol[k + 1] * qrl[k + 1] * vr[k + 1] : 0.0;
              const double fbot = rhol[k] * qrl[k] * vr[k];
              qrl[k] += dts * (ftop - fbot) / (rhol[k] * dzl[k]);
            }
            if (sc + 1 < nsub) {
              for (int k = 0; k < g.nz; ++k) {
                const double rq = rhol[k] * (qrl[k] > 0.0 ? qrl[k] : 0.0);
                vr[k] = rq > 1e-9 ? 36.34 * std::pow(rq, 0.1364) *
                                        std::sqrt(rho_sfc / rhol[k]) * fallfac(g.id(i, j, k))
                                  : 0.0;
              }
            }
          }
          rain[g.id2(i, j)] += acc;
          for (int k = 0; k < g.nz; ++k) {
            const long c = g.id(i, j, k);
            double qvk = s.qv[c] > 0.0 ? s.qv[c] : 0.0;
            double qck = s.qc[c] > 0.0 ? s.qc[c] : 0.0;
            double qrk = qrl[k] > 0.0 ? qrl[k] : 0.0;
            const double pi = pib[c] + s.pip[c];
            double th = thb[c] + s.thp[c];
            double T = th * pi;
            const double pres = kP00 * std::pow(pi, kCp / kRd);
            double fl = (T - 233.15) / 40.0;
            fl = std::fmax(0.0, std::fmin(1.0, fl));
            const double qsw = qsat_tetens(pres, T);
            double qs = fl * qsw + (1.0 - fl) * qsat_ice(pres, T);
            const double Leff = fl * kLv + (1.0 - fl) * kLs;
            const double dqauto = k1 * (qck > qc0 ? qck - qc0 : 0.0);
            const double dqacc = 2.2 * qck * std::pow(qrk > 1e-9 ? qrk : 1e-9, 0.875);
            double dqr = dt * (dqauto + dqacc);
            if (dqr > qck) dqr = qck;
            qck -= dqr;
            qrk += dqr;
            if (qvk < qs && qrk > 0.0) {
              const double rq = rhol[k] * qrk;
              const double vent = 1.6 + 124.9 * std::pow(rq, 0.2046);
              const double ern = vent * std::pow(rq, 0.525) / (2.55e8 / (pres * qs) + 5.4e5) *
                                 (qs - qvk) / (rhol[k] * qs);
              double dqe = dt * ern;
              if (dqe > qrk) dqe = qrk;
              if (dqe > qs - qvk) dqe = qs - qvk;
              qrk -= dqe;
              qvk += dqe;
              s.thp[c] -= Leff * dqe / (kCp * pi);
              th = thb[c] + s.thp[c];
              T = th * pi;
              const double fl2 = std::fmax(0.0, std::fmin(1.0, (T - 233.15) / 40.0));
              qs = fl2 * qsat_tetens(pres, T) + (1.0 - fl2) * qsat_ice(pres, T);
            }
            const double gam = Leff * Leff * qs / (kCp * kRv * T * T);
            double dq = (qvk - qs) / (1.0 + gam);
            if (dq < 0.0) {
              double evap = -dq;
              if (evap > qck) evap = qck;
              dq = -evap;
            }
            qvk -= dq;
            qck += dq;
            s.thp[c] += Leff * dq / (kCp * pi);
            s.qv[c] = qvk > 0.0 ? qvk : 0.0;
            s.qc[c] = qck > 0.0 ? qck : 0.0;
            s.qr[c] = qrk > 0.0 ? qrk : 0.0;
          }
        }
    }
  }


  // ------------------------------------- surface/PBL physics (src/physics/surface.cu)
  static constexpr double kKappa = 0.4, kSigma = 5.67e-8, kS0 = 1361;  // surface.cu:24-27
  static constexpr double kKmMax = 1000, kKmMin = 0.1;
  static double psi_m(double zeta) {  // surface.cu:31-37
    if (zeta >= 0) return -5 * std::fmin(zeta, 2.0);
    const double x = std::pow(1 - 16 * zeta, 0.25);
    const double halfpi = 1.5707963267948966;
    return 2 * std::log((1 + x) * 0.5) + std::log((1 + x * x) * 0.5) - 2 * std::atan(x) +
           halfpi;
  }
  static double psi_h(double zeta) {  // surface.cu:38-42
    if (zeta >= 0) return -5 * std::fmin(zeta, 2.0);
    const double x = std::pow(1 - 16 * zeta, 0.25);
    return 2 * std::log((1 + x * x) * 0.5);
  }
  static double soil_C(double w) { return 1.2e6 + 3.2e6 * std::fmin(std::fmax(w, 0.0), 0.4); }
  static double soil_K(double w) { return 0.25 + 3.0 * std::fmin(std::fmax(w, 0.0), 0.4); }
  // Column-implicit vertical diffusion (surface.cu:61-84).
  static void diffuse_column(int nz, double dt, const double* rho, const double* rhow,
                             const double* dzc, const double* dzw, const double* Kw,
                             double src0, double* phi, const double* fcg) {
    std::vector<double> A(nz), B(nz), C(nz), D(nz);
    for (int k = 0; k < nz; ++k) {
      const double a = (k > 0) ? rhow[k] * Kw[k] / dzw[k] : 0.0;
      const double c = (k < nz - 1) ? rhow[k + 1] * Kw[k + 1] / dzw[k + 1] : 0.0;
      const double m = rho[k] * dzc[k] / dt;
      A[k] = -a;
      B[k] = m + a + c;
      C[k] = -c;
      D[k] = m * phi[k];
      if (fcg) D[k] += fcg[k] - fcg[k + 1];
    }
    D[0] += src0;
    for (int k = 1; k < nz; ++k) {
      const double w = A[k] / B[k - 1];
      B[k] -= w * C[k - 1];
      D[k] -= w * D[k - 1];
    }
    phi[nz - 1] = D[nz - 1] / B[nz - 1];
    for (int k = nz - 2; k >= 0; --k) phi[k] = (D[k] - C[k] * phi[k + 1]) / B[k];
  }

  // Louis (1979) bulk exchange coefficients for momentum and heat at
  // height z1 over roughness z0 (surface.cu:140-159).
  static void louis(double Rib, double z1, double z0, double* Cd, double* Ch) {
    double a2 = kKappa / std::log(z1 / z0);
    a2 *= a2;
    double Fm, Fh;
    if (Rib < 0) {
      const double sq = std::sqrt(-Rib * z1 / z0);
      Fm = 1 - 9.4 * Rib / (1 + 7.4 * 9.4 * a2 * sq);
      Fh = 1 - 9.4 * Rib / (1 + 5.3 * 9.4 * a2 * sq);
    } else {
      const double d = 1 + 4.7 * Rib;
      Fm = Fh = 1 / (d * d);
    }
    *Cd = a2 * Fm;
    *Ch = a2 * Fh;
  }
  // 10 m wind and 2 m potential temperature from the lowest level (z1, V1,
  // th1) and the surface (ths) through the Businger-Dyer profiles
  // (surface.cu:166-184); z0h = z0 / 10.
  static void sfc_diag(double V1, double Rib, double z1, double z0, double ths, double th1,
                       double* u10, double* th2) {
    const double z0h = z0 * 0.1;
    const double lnz1z0 = std::log(z1 / z0);
    double zeta1 = (Rib >= 0) ? Rib * lnz1z0 / (1 - 5 * std::fmin(Rib, 0.19)) : Rib * lnz1z0;
    zeta1 = std::fmax(-8.0, std::fmin(2.0, zeta1));
    const double zol = zeta1 / z1;
    const double den_m = lnz1z0 - psi_m(zeta1) + psi_m(z0 * zol);
    const double num_m = std::log(10 / z0) - psi_m(10 * zol) + psi_m(z0 * zol);
    const double r10 = std::fmax(0.0, std::fmin(1.2, num_m / std::fmax(den_m, 0.1)));
    *u10 = V1 * r10;
    const double den_h = std::log(z1 / z0h) - psi_h(zeta1) + psi_h(z0h * zol);
    const double num_h = std::log(2 / z0h) - psi_h(2 * zol) + psi_h(z0h * zol);
    const double r2 = std::fmax(0.0, std::fmin(1.2, num_h / std::fmax(den_h, 0.1)));
    *th2 = ths + (th1 - ths) * r2;
  }
  // Cosine of the solar zenith angle, clipped at 0 (surface.cu:188-196).
  static double solar_cosz(int doy, double utc, double lat, double lon) {
    const double decl = 0.40928 * std::sin(6.2832 * (284 + doy) / 365.0);
    const double latr = lat * 0.0174533;
    const double hsol = utc + lon / 15;
    const double H = 0.2618 * (hsol - 12);
    const double cosz = std::sin(latr) * std::sin(decl) + std::cos(latr) * std::cos(decl) * std::cos(H);
    return cosz < 0 ? 0 : cosz;
  }
  // Grey long-wave fluxes of a column of nz layers with temperatures T and
  // emissivities eps over a surface at Ts (emissivity 0.97): Fdn, Fup at the
  // nz + 1 interfaces (surface.cu:200-216).
  static void lw_fluxes(int nz, const double* T, const double* eps, double Ts, double* Fup,
                        double* Fdn) {
    Fdn[nz] = 0;
    for (int k = nz - 1; k >= 0; --k)
      Fdn[k] = Fdn[k + 1] * (1 - eps[k]) + eps[k] * kSigma * T[k] * T[k] * T[k] * T[k];
    Fup[0] = 0.97 * kSigma * Ts * Ts * Ts * Ts + 0.03 * Fdn[0];
    for (int k = 0; k < nz; ++k)
      Fup[k + 1] = Fup[k] * (1 - eps[k]) + eps[k] * kSigma * T[k] * T[k] * T[k] * T[k];
  }
  // Short-wave absorption of a column (vapour path du, liquid path dl per
  // layer) under the sun at cosz with cloud albedo from the liquid water
  // path lwp: swabs per layer, returns the flux reaching the surface
  // (surface.cu:218-232).
  static double sw_column(int nz, double cosz, double lwp, const double* du, const double* dl,
                          double* swabs) {
    const double mu = std::fmax(cosz, 0.05);
    const double cldalb = lwp / (lwp + 60);
    double Sdn = kS0 * cosz * 0.95 * (1 - cldalb);
    for (int k = nz - 1; k >= 0; --k) {
      const double tau = std::exp(-(0.008 * du[k] + 0.15 * dl[k]) / mu);
      swabs[k] = Sdn * (1 - tau);
      Sdn *= tau;
    }
    return Sdn;
  }
  // One implicit step of the four-layer soil (layer depths 0.1, 0.3, 0.6,
  // 1.0 m) with surface flux G into layer 0 and a deep temperature below
  // layer 3; water content w sets the heat capacity and conductivity
  // (surface.cu:247-289). Tso is updated in place and clipped to 200-350 K.
  static void soil_step(double dt, double w, double G, double tdeep, double* Tso) {
    static const double kDz[4] = {0.1, 0.3, 0.6, 1.0};
    const double Cso = soil_C(w), Kso = soil_K(w);
    double A[4], B[4], C[4], D[4];
    for (int L = 0; L < 4; ++L) {
      const double cap = Cso * kDz[L] / dt;
      const double ku = (L > 0) ? Kso / (0.5 * (kDz[L - 1] + kDz[L])) : 0.0;
      const double kl = (L < 3) ? Kso / (0.5 * (kDz[L] + kDz[L + 1])) : Kso / kDz[L];
      A[L] = -ku;
      C[L] = (L < 3) ? -kl : 0.0;
      B[L] = cap + ku + kl;
      D[L] = cap * Tso[L];
    }
    D[0] += G;
    D[3] += (Kso / kDz[3]) * tdeep;
    for (int L = 1; L < 4; ++L) {
      const double w2 = A[L] / B[L - 1];
      B[L] -= w2 * C[L - 1];
      D[L] -= w2 * D[L - 1];
    }
    Tso[3] = D[3] / B[3];
    for (int L = 2; L >= 0; --L) Tso[L] = (D[L] - C[L] * Tso[L + 1]) / B[L];
    for (int L = 0; L < 4; ++L) {
      if (Tso[L] < 200) Tso[L] = 200;
      if (Tso[L] > 350) Tso[L] = 350;
    }
  }

  // k_sfc_scalar (surface.cu:87-387) for column (i, j) of state s.
  void sfc_scalar_column(int i, int j, double dt, double utc, State& s) {
    const int nz = g.nz;
    const long c2 = g.id2(i, j);
    const double J = jac[c2];
    std::vector<double> rho(nz), rhow(nz), dzc(nz), dzw(nz), uc(nz), vc(nz), thv(nz),
        Kw(nz + 1), thcol(nz), qvcol(nz);
    double lwp = 0;
    for (int k = 0; k < nz; ++k) {  // :101-114
      const int kk = k + g.ng;
      const long c = g.id(i, j, k);
      rho[k] = rhob[c];
      rhow[k] = rhobw[c];
      dzc[k] = dzeta_c[kk] * J;
      dzw[k] = dzeta_w[kk] * J;
      uc[k] = 0.5 * (s.u[g.id(i, j, k)] + s.u[g.id(i + 1, j, k)]);
      vc[k] = 0.5 * (s.v[g.id(i, j, k)] + s.v[g.id(i, j + 1, k)]);
      thcol[k] = thb[c] + s.thp[c];
      qvcol[k] = s.qv[c];
      thv[k] = thcol[k] * (1 + kEps61 * qvcol[k]);
      lwp += rho[k] * (s.qc[c] > 0 ? s.qc[c] : 0.0) * dzc[k] * 1000;
    }
    // surface layer (Louis 1979), :117-159
    const bool onland = land[c2] > 0.5;
    const double z0 = onland ? 0.1 : 2e-4;
    const double alb = onland ? 0.2 : 0.08;
    double beta;
    if (onland) {
      double f = (soilw[c2] - 0.08) / (0.32 - 0.08);
      f = std::fmin(1.0, std::fmax(0.0, f));
      beta = std::fmax(0.05, f * f);
    } else {
      beta = 1;
    }
    const double z1 = 0.5 * dzc[0];
    double V1 = std::sqrt(uc[0] * uc[0] + vc[0] * vc[0]);
    if (V1 < 0.5) V1 = 0.5;
    const long c0 = g.id(i, j, 0);
    const double pi1 = pib[c0] + s.pip[c0];
    const double p1 = kP00 * std::pow(pi1, kCp / kRd);
    const double Ts = tsk[c2];
    const double pis = pi1 + kGrav * z1 / (kCp * thv[0]);
    const double ths = Ts / pis;
    const double qs_s = qsat_tetens(p1, Ts);
    const double thvs = ths * (1 + kEps61 * (beta * qs_s + (1 - beta) * qvcol[0]));
    const double Rib = kGrav * z1 * (thv[0] - thvs) / (thvs * V1 * V1);
    double Cd, Ch;
    louis(Rib, z1, z0, &Cd, &Ch);
    cdv[c2] = Cd * V1;
    // 2 m and 10 m diagnostics (:166-184)
    double th2;
    sfc_diag(V1, Rib, z1, z0, ths, thcol[0], &u10[c2], &th2);
    t2m[c2] = th2 * pi1;
    double shf = Ch * V1 * (ths - thcol[0]);
    double qflx = Ch * V1 * beta * (qs_s - qvcol[0]);
    if (!onland && qflx < 0) qflx = 0;
    // radiation (:188-244)
    const double cosz = solar_cosz(phys_doy, utc, lat[c2], lon[c2]);
    std::vector<double> Tlev(nz), epslw(nz), Fup(nz + 1), Fdn(nz + 1), radht(nz), swabs(nz),
        du(nz), dl(nz);
    for (int k = 0; k < nz; ++k) {
      const long c = g.id(i, j, k);
      const double pik = pib[c] + s.pip[c];
      Tlev[k] = thcol[k] * pik;
      du[k] = rho[k] * qvcol[k] * dzc[k];
      dl[k] = rho[k] * (s.qc[c] > 0 ? s.qc[c] : 0.0) * dzc[k] * 1000;
      epslw[k] = 1 - std::exp(-(0.22 * du[k] + 0.30 * dl[k]));
    }
    lw_fluxes(nz, Tlev.data(), epslw.data(), Ts, Fup.data(), Fdn.data());
    const double lwd = Fdn[0];
    const double Sdn = sw_column(nz, cosz, lwp, du.data(), dl.data(), swabs.data());
    const double swn = Sdn * (1 - alb);
    for (int k = 0; k < nz; ++k) {
      const long c = g.id(i, j, k);
      const double pik = pib[c] + s.pip[c];
      const double lwconv = (Fdn[k + 1] - Fdn[k]) + (Fup[k] - Fup[k + 1]);
      const double net = lwconv + swabs[k];
      radht[k] = net / (rho[k] * kCp * dzc[k]) / pik;
    }
    // multi-layer soil (:247-289)
    const double lwu = 0.97 * kSigma * Ts * Ts * Ts * Ts;
    const double G = swn + 0.97 * lwd - lwu - rho[0] * kCp * shf * pi1 - rho[0] * kLv * qflx;
    if (onland) {
      const long n2 = g.n2();
      double Tso[4];
      for (int L = 0; L < 4; ++L) Tso[L] = soilt[L * n2 + c2];
      soil_step(dt, soilw[c2], G, tdeep[c2], Tso);
      for (int L = 0; L < 4; ++L) soilt[L * n2 + c2] = Tso[L];
      tsk[c2] = Tso[0];
    }
    // diffusivity (:292-305)
    auto local_K = [&](int k) -> double {
      const double dz = dzw[k];
      const double du = (uc[k] - uc[k - 1]) / dz;
      const double dv = (vc[k] - vc[k - 1]) / dz;
      const double S2 = du * du + dv * dv + 1e-8;
      const double thvw = 0.5 * (thv[k - 1] + thv[k]);
      const double N2 = kGrav / thvw * (thv[k] - thv[k - 1]) / dz;
      const double Ri = N2 / S2;
      const double zw = zeta_w[k + g.ng] * J;
      const double l = kKappa * zw / (1 + kKappa * zw / 150);
      if (Ri < 0) return l * l * std::sqrt(S2) * std::sqrt(1 - 16 * Ri);
      const double d = 1 + 5 * Ri;
      return l * l * std::sqrt(S2) / (d * d);
    };
    std::vector<double> fcg_th(nz + 1, 0.0), fcg_qv(nz + 1, 0.0);
    Kw[0] = 0;
    if (phys_nonlocal) {  // :311-370
      const double ust = std::sqrt(Cd) * V1;
      const double thv1 = thv[0];
      const double wthv0 = shf * (1 + kEps61 * qvcol[0]) + kEps61 * thcol[0] * qflx;
      const double Ric = 0.25;
      double hpbl = zeta_c[(nz - 1) + g.ng] * J;
      double ws = ust > 0.1 ? ust : 0.1;
      for (int it = 0; it < 3; ++it) {
        const double exc = (wthv0 > 0) ? 6.8 * wthv0 / ws : 0.0;
        const double thv_s = thv1 + exc;
        double prevRb = 0;
        hpbl = zeta_c[(nz - 1) + g.ng] * J;
        for (int k = 1; k < nz; ++k) {
          const double zc = zeta_c[k + g.ng] * J;
          const double shear = uc[k] * uc[k] + vc[k] * vc[k] + 0.1;
          const double Rb = kGrav * (thv[k] - thv_s) * zc / (thv1 * shear);
          if (Rb >= Ric) {
            const double zcm = zeta_c[(k - 1) + g.ng] * J;
            double f = (Ric - prevRb) / (Rb - prevRb + 1e-9);
            f = std::fmin(1.0, std::fmax(0.0, f));
            hpbl = zcm + f * (zc - zcm);
            break;
          }
          prevRb = Rb;
        }
        const double zc0 = zeta_c[g.ng] * J;
        if (hpbl < zc0) hpbl = zc0;
        const double wstar = (wthv0 > 0) ? std::cbrt(kGrav / thv1 * wthv0 * hpbl) : 0.0;
        ws = std::cbrt(ust * ust * ust + 0.6 * wstar * wstar * wstar);
        if (ws < 0.1) ws = 0.1;
      }
      pblh[c2] = hpbl;
      const double gth = (wthv0 > 0) ? 6.8 * shf / (ws * hpbl) : 0.0;
      const double gqv = (wthv0 > 0) ? 6.8 * qflx / (ws * hpbl) : 0.0;
      for (int k = 1; k < nz; ++k) {
        const double zw = zeta_w[k + g.ng] * J;
        double Kv;
        if (wthv0 > 0 && zw < hpbl) {
          const double zf = zw / hpbl;
          Kv = kKappa * ws * zw * (1 - zf) * (1 - zf);
          fcg_th[k] = rhow[k] * Kv * gth;
          fcg_qv[k] = rhow[k] * Kv * gqv;
        } else {
          Kv = local_K(k);
          fcg_th[k] = fcg_qv[k] = 0;
        }
        if (Kv < kKmMin) Kv = kKmMin;
        if (Kv > kKmMax) Kv = kKmMax;
        Kw[k] = Kv;
        km[g.id(i, j, k)] = Kv;
      }
    } else {  // :371-382
      pblh[c2] = 0;
      for (int k = 1; k < nz; ++k) {
        double Kv = local_K(k);
        if (Kv < kKmMin) Kv = kKmMin;
        if (Kv > kKmMax) Kv = kKmMax;
        Kw[k] = Kv;
        km[g.id(i, j, k)] = Kv;
      }
    }
    // implicit scalar mixing (:385-386) and radiative heating (:389-397)
    diffuse_column(nz, dt, rho.data(), rhow.data(), dzc.data(), dzw.data(), Kw.data(),
                   rho[0] * shf, thcol.data(), fcg_th.data());
    diffuse_column(nz, dt, rho.data(), rhow.data(), dzc.data(), dzw.data(), Kw.data(),
                   rho[0] * qflx, qvcol.data(), fcg_qv.data());
    for (int k = 0; k < nz; ++k) {
      const long c = g.id(i, j, k);
      double dth = dt * radht[k];
      if (dth > 0.02) dth = 0.02;
      if (dth < -0.02) dth = -0.02;
      s.thp[c] = thcol[k] - thb[c] + dth;
      const double q = qvcol[k];
      s.qv[c] = q > 0 ? q : 0.0;
    }
  }
  // k_sfc_u / k_sfc_v (surface.cu:390-437): face column (fi, fj) averaged
  // from the cells (fi - di, fj - dj) and (fi, fj); surface drag + mixing.
  void sfc_face_column(Vec& f, int fi, int fj, int di, int dj, double dt) {
    const int nz = g.nz;
    const long a2 = g.id2(fi - di, fj - dj), b2 = g.id2(fi, fj);
    const double J = 0.5 * (jac[a2] + jac[b2]);
    std::vector<double> rho(nz), rhow(nz), dzc(nz), dzw(nz), Kw(nz), phi(nz);
    for (int k = 0; k < nz; ++k) {
      const int kk = k + g.ng;
      const long a = g.id(fi - di, fj - dj, k), c = g.id(fi, fj, k);
      rho[k] = 0.5 * (rhob[a] + rhob[c]);
      rhow[k] = 0.5 * (rhobw[a] + rhobw[c]);
      dzc[k] = dzeta_c[kk] * J;
      dzw[k] = dzeta_w[kk] * J;
      Kw[k] = (k > 0) ? 0.5 * (km[a] + km[c]) : 0.0;
      phi[k] = f[c];
    }
    const double cdva = 0.5 * (cdv[a2] + cdv[b2]);
    const double src0 = -rho[0] * cdva * phi[0];
    diffuse_column(nz, dt, rho.data(), rhow.data(), dzc.data(), dzw.data(), Kw.data(), src0,
                   phi.data(), nullptr);
    for (int k = 0; k < nz; ++k) f[g.id(fi, fj, k)] = phi[k];
  }
  // SfcPBL::step (surface.cu:499-514): t is the time at the end of the step.
  void surface_step(double dt, double t, State& s) {
    double utc = phys_start_hour + t / 3600;
    utc = utc - 24 * std::floor(utc / 24);
#pragma omp parallel for collapse(2)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) sfc_scalar_column(i, j, dt, utc, s);
#pragma omp parallel for collapse(2)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 1; i < g.nx; ++i) sfc_face_column(s.u, i, j, 1, 0, dt);
#pragma omp parallel for collapse(2)
    for (int j = 1; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) sfc_face_column(s.v, i, j, 0, 1, dt);
  }

  // Integrator::step (integrator.cpp:71-111).
  void step(double dt, double t = 0) {
    const double stage_dt[3] = {dt / 3, dt / 2, dt};
    const int ns = std::max(1, p.acoustic_ns);
    const int ns_stage[3] = {1, std::max(1, ns / 2), ns};
    for (int m = 0; m < 3; ++m) {
      State& est = (m == 0) ? sn : sstage;
      apply_bcs(est);
      mass_fluxes(est);
      divergence();
      compute_tendencies(est, dt);
      if (bdy_active) bdy_relax(est);
      swork.copy_from(sn);
      piprev = sn.pip;
      const double dtau = stage_dt[m] / ns_stage[m];
      for (int t = 0; t < ns_stage[m]; ++t) acoustic_substep(dtau, swork);
      if (p.moisture) update_moisture_stage(sn, stage_dt[m], swork);
      sstage.swap(swork);
    }
    sn.swap(sstage);
    if (p.moisture && p.microphysics) kessler(dt, sn);
    if (phys_on) surface_step(dt, t + dt, sn);  // integrator.cpp:109
  }

  Vec* array(const std::string& n) {
    static const char* tn[8] = {"tu", "tv", "tw", "tthp", "tpip", "tqv", "tqc", "tqr"};
    static const char* sn_[8] = {"u", "v", "w", "thp", "pip", "qv", "qc", "qr"};
    std::vector<Vec*> sf = sn.fields(), tf = tend.fields();
    for (int q = 0; q < 8; ++q) {
      if (n == sn_[q]) return sf[q];
      if (n == tn[q]) return tf[q];
    }
    struct NV { const char* n; Vec* v; };
    const NV t[] = {{"thb", &thb}, {"pib", &pib}, {"rhob", &rhob}, {"dthbdz", &dthbdz},
                    {"thbw", &thbw}, {"rhobw", &rhobw}, {"thvb", &thvb}, {"thvbw", &thvbw},
                    {"qvb", &qvb}, {"ub", &ub}, {"vb", &vb}, {"zeta_w", &zeta_w},
                    {"zeta_c", &zeta_c}, {"dzeta_c", &dzeta_c}, {"dzeta_w", &dzeta_w},
                    {"h", &h}, {"hx", &hx}, {"hy", &hy}, {"hx_u", &hxu}, {"hy_v", &hyv},
                    {"jac", &jac}, {"fcor", &fcor}, {"mapf", &mapf}, {"rain", &rain},
                    {"bdy_wgt", &bdy_wgt}, {"mfx", &mfx}, {"mfy", &mfy}, {"mfz", &mfz},
                    {"div", &div}, {"rtjx", &rtjx}, {"rtjy", &rtjy}, {"kt3", &kt3}, {"aw", &aw},
                    {"piprev", &piprev}, {"tsk", &tsk}, {"land", &land}, {"lat", &lat},
                    {"lon", &lon}, {"soilw", &soilw}, {"cdv", &cdv}, {"pblh", &pblh},
                    {"t2m", &t2m}, {"u10", &u10}, {"km", &km}, {"soilt", &soilt},
                    {"tdeep", &tdeep}};
    for (const NV& e : t)
      if (n == e.n) return e.v;
    for (int v = 0; v < 5; ++v) {
      if (n == "bdy_lo_" + std::to_string(v)) return &bdy_lo[v];
      if (n == "bdy_hi_" + std::to_string(v)) return &bdy_hi[v];
    }
    return nullptr;
  }
  long expected_size(const std::string& n) const {
    if (n == "ub" || n == "vb" || n == "zeta_w" || n == "zeta_c" || n == "dzeta_c" ||
        n == "dzeta_w")
      return g.NZ;
    if (n == "h" || n == "hx" || n == "hy" || n == "hx_u" || n == "hy_v" || n == "jac" ||
        n == "fcor" || n == "mapf" || n == "rain" || n == "bdy_wgt" || n == "tsk" ||
        n == "land" || n == "lat" || n == "lon" || n == "soilw" || n == "cdv" || n == "pblh" ||
        n == "t2m" || n == "u10" || n == "tdeep")
      return g.n2();
    if (n == "soilt") return 4 * g.n2();
    return g.n3();
  }
};

extern "C" {

orc_model* orc_create(int nx, int ny, int nz, int ng, double dx, double dy, double dz,
                      double zt) {
  if (nx < 4 || ny < 1 || nz < 4 || ng != 3) return nullptr;
  orc_model* m = new orc_model();
  m->g = Grid{nx, ny, nz, ng, nx + 2 * ng, ny + 2 * ng, nz + 1 + 2 * ng, dx, dy, dz};
  m->zt = zt;
  const long n3 = m->g.n3(), n2 = m->g.n2(), n1 = m->g.NZ;
  for (Vec* v : {&m->thb, &m->pib, &m->rhob, &m->dthbdz, &m->thbw, &m->rhobw, &m->thvb,
                 &m->thvbw, &m->qvb, &m->div, &m->piprev, &m->mfx, &m->mfy, &m->mfz, &m->pfx,
                 &m->pfy, &m->pfz, &m->pratio})
    v->assign(n3, 0.0);
  for (Vec* v : {&m->ub, &m->vb, &m->zeta_w, &m->zeta_c, &m->dzeta_c, &m->dzeta_w})
    v->assign(n1, 0.0);
  for (Vec* v : {&m->h, &m->hx, &m->hy, &m->hxu, &m->hyv, &m->jac, &m->fcor, &m->mapf,
                 &m->rain, &m->bdy_wgt, &m->tsk, &m->tdeep, &m->land, &m->lat, &m->lon,
                 &m->soilw, &m->cdv, &m->pblh, &m->t2m, &m->u10})
    v->assign(n2, 0.0);
  m->km.assign(n3, 0.0);
  m->soilt.assign(4 * n2, 0.0);
  m->sn.alloc(n3);
  m->sstage.alloc(n3);
  m->swork.alloc(n3);
  m->tend.alloc(n3);
  for (int v = 0; v < 5; ++v) {
    m->bdy_lo[v].assign(n3, 0.0);
    m->bdy_hi[v].assign(n3, 0.0);
  }
  return m;
}

void orc_destroy(orc_model* m) { delete m; }

int orc_set_param(orc_model* m, const char* name, double v) {
  Params& p = m->p;
  const std::string n(name);
  if (n == "diff_K") p.diff_K = v;
  else if (n == "coriolis_use_ub") p.coriolis_use_ub = (int)v;
  else if (n == "rayleigh_zd") p.rayleigh_zd = v;
  else if (n == "rayleigh_alpha") p.rayleigh_alpha = v;
  else if (n == "acoustic_ns") p.acoustic_ns = (int)v;
  else if (n == "acoustic_beta") p.acoustic_beta = v;
  else if (n == "acoustic_smdiv") p.acoustic_smdiv = v;
  else if (n == "bc_x_open") p.bc_x_open = (int)v;
  else if (n == "bc_y_open") p.bc_y_open = (int)v;
  else if (n == "cstar") p.cstar = v;
  else if (n == "moisture") p.moisture = (int)v;
  else if (n == "microphysics") p.microphysics = (int)v;
  else if (n == "pd_moist") p.pd_moist = (int)v;
  else if (n == "w_damping") p.w_damping = (int)v;
  else return -1;
  return 0;
}

int orc_set_array(orc_model* m, const char* name, const double* data, long n) {
  Vec* v = m->array(name);
  if (!v || n != m->expected_size(name)) return -1;
  std::memcpy(v->data(), data, sizeof(double) * n);
  return 0;
}

int orc_get_array(const orc_model* m, const char* name, double* data, long n) {
  Vec* v = const_cast<orc_model*>(m)->array(name);
  if (!v || n != m->expected_size(name) || (long)v->size() != n) return -1;
  std::memcpy(data, v->data(), sizeof(double) * n);
  return 0;
}

int orc_init(orc_model* m) {
  m->precompute_acoustic_coef();
  return 0;
}

int orc_set_bdy(orc_model* m, int active, double tfrac) {
  m->bdy_active = active;
  m->tfrac = tfrac;
  return 0;
}

int orc_set_physics(orc_model* m, int nonlocal, double start_hour_utc, int doy) {
  // SfcPBL::init (surface.cu:440-482): deep soil and the four soil layers
  // start at the surface temperature
  m->phys_on = 1;
  m->phys_nonlocal = nonlocal;
  m->phys_start_hour = start_hour_utc;
  m->phys_doy = doy;
  m->tdeep = m->tsk;
  const long n2 = m->g.n2();
  for (int L = 0; L < 4; ++L)
    std::copy(m->tsk.begin(), m->tsk.end(), m->soilt.begin() + L * n2);
  return 0;
}

int orc_step(orc_model* m, double dt, double t) {
  m->step(dt, t);
  return 0;
}

int orc_apply_bcs(orc_model* m) {
  m->apply_bcs(m->sn);
  return 0;
}

int orc_tendencies(orc_model* m, double dt) {
  m->mass_fluxes(m->sn);
  m->divergence();
  m->compute_tendencies(m->sn, dt);
  if (m->bdy_active) m->bdy_relax(m->sn);
  return 0;
}

int orc_acoustic(orc_model* m, double dtau, int nsub) {
  m->swork.copy_from(m->sn);
  m->piprev = m->sn.pip;
  for (int t = 0; t < nsub; ++t) m->acoustic_substep(dtau, m->swork);
  m->sn.swap(m->swork);
  return 0;
}

int orc_kessler(orc_model* m, double dt) {
  m->kessler(dt, m->sn);
  return 0;
}

double orc_absmax(const orc_model* m, const char* name) {
  Vec* v = const_cast<orc_model*>(m)->array(name);
  if (!v) return -1;
  const Grid& g = m->g;
  const int kmax = std::string(name) == "w" ? g.nz : g.nz - 1;
  double mx = 0;
  for (int k = 0; k <= kmax; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double a = std::fabs((*v)[g.id(i, j, k)]);
        if (!(a <= mx)) mx = a;  // NaN propagates
      }
  return mx;
}

void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_get_threads(void) { return omp_get_max_threads(); }

double orc_iface5(double qm2, double qm1, double q0, double q1, double q2, double q3,
                  double vel) {
  return iface5(qm2, qm1, q0, q1, q2, q3, vel);
}

void orc_thomas(int n, const double* a, const double* b, const double* c, const double* d,
                double* x) {
  std::vector<double> cs(n), ds(n);
  for (int k = 0; k < n; ++k) {
    if (k == 0) {
      cs[k] = c[k] / b[k];
      ds[k] = d[k] / b[k];
    } else {
      const double mm = b[k] - a[k] * cs[k - 1];
      cs[k] = c[k] / mm;
      ds[k] = (d[k] - a[k] * ds[k - 1]) / mm;
    }
  }
  double above = 0;
  for (int k = n - 1; k >= 0; --k) {
    x[k] = ds[k] - cs[k] * above;
    above = x[k];
  }
}

double orc_qsat_w(double p, double T) { return qsat_tetens(p, T); }
double orc_sfc_cosz(int doy, double utc, double lat, double lon) {
  return orc_model::solar_cosz(doy, utc, lat, lon);
}
void orc_sfc_louis(double Rib, double z1, double z0, double* cd_ch) {
  orc_model::louis(Rib, z1, z0, &cd_ch[0], &cd_ch[1]);
}
void orc_sfc_diag(double V1, double Rib, double z1, double z0, double ths, double th1,
                  double* u10_th2) {
  orc_model::sfc_diag(V1, Rib, z1, z0, ths, th1, &u10_th2[0], &u10_th2[1]);
}
void orc_sfc_lw(int nz, const double* T, const double* eps, double Ts, double* Fup,
                double* Fdn) {
  orc_model::lw_fluxes(nz, T, eps, Ts, Fup, Fdn);
}
double orc_sfc_sw(int nz, double cosz, double lwp, const double* du, const double* dl,
                  double* swabs) {
  return orc_model::sw_column(nz, cosz, lwp, du, dl, swabs);
}
void orc_sfc_soil(double dt, double w, double G, double tdeep, double* Tso) {
  orc_model::soil_step(dt, w, G, tdeep, Tso);
}
double orc_sfc_soil_heat(double w, int which) {
  return which == 0 ? orc_model::soil_C(w) : orc_model::soil_K(w);
}

double orc_sfc_psi(int heat, double zeta) {
  return heat ? orc_model::psi_h(zeta) : orc_model::psi_m(zeta);
}
void orc_sfc_diffuse(int nz, double dt, const double* rho, const double* rhow, const double* dzc,
                     const double* dzw, const double* Kw, double src0, double* phi,
                     const double* fcg) {
  orc_model::diffuse_column(nz, dt, rho, rhow, dzc, dzw, Kw, src0, phi, fcg);
}
double orc_qsat_i(double p, double T) { return qsat_ice(p, T); }

}  // extern "C"
