#pragma once
// Surface-layer / PBL / soil / radiation column physics applied after each
// large step (SfcPBL, src/physics/surface.cu:87-514): Louis (1979) bulk
// surface fluxes with Businger-Dyer 2 m / 10 m diagnostics, a two-stream
// broadband radiation column, a four-layer implicit soil, a local-Ri or
// nonlocal (Troen-Mahrt / Hong-Pan) diffusivity profile, and implicit
// vertical mixing of theta, qv, u, v. One thread per column (cells for the
// scalars, u and v faces for the momentum); a column's working arrays live
// in the thread's private memory (KMAX levels). Arithmetic is FP64 in both
// precision modes; the state keeps its storage type.
#include "hip/device_common.hpp"

namespace wfe::dev {

// Physics fields: FP64, 2-D on the padded plane (id2) unless noted.
struct SfcArgs {
  double *tsk, *tdeep, *land, *lat, *lon, *soilw;
  double *cdv, *pblh, *t2m, *u10;
  double* km;     // 3-D, w levels 1..nz-1 of every interior column
  double* soilt;  // four soil layers, layer L at L * plane
  const double* utc;  // hour of day at the end of the step (device scalar)
  double dt;
  int doy, nonlocal;
};

template <class T>
struct SfcState {
  T *thp, *qv, *u, *v;
  const T* qc;
  const double* pip;
};

inline constexpr double kKarman = 0.4, kStefan = 5.67e-8, kSolar = 1361.0;
inline constexpr double kKmMax = 1000.0, kKmMin = 0.1;

// Businger-Dyer stability functions of zeta = z/L (surface.cu:31-42).
__device__ __forceinline__ double sfc_psi_m(double zeta) {
  if (zeta >= 0) return -5.0 * fmin(zeta, 2.0);
  const double x = pow(1.0 - 16.0 * zeta, 0.25);
  return 2.0 * log((1.0 + x) * 0.5) + log((1.0 + x * x) * 0.5) - 2.0 * atan(x) +
         1.5707963267948966;
}
__device__ __forceinline__ double sfc_psi_h(double zeta) {
  if (zeta >= 0) return -5.0 * fmin(zeta, 2.0);
  const double x = pow(1.0 - 16.0 * zeta, 0.25);
  return 2.0 * log((1.0 + x * x) * 0.5);
}

// Implicit vertical diffusion of a column (surface.cu:61-84): rho-weighted
// fluxes K dphi/dz at w levels 1..nz-1, surface source src0 into level 0,
// optional upward counter-gradient flux fcg at w levels 0..nz; Thomas
// elimination in place of phi. The sub- and super-diagonals are -k_w of the
// w level between two cells, k_w = rhow K / dzw, so only the diagonal B and
// the right-hand side D need storage.
template <int KMAX>
__device__ void sfc_diffuse(int nz, double dt, const double* rho, const double* rhow,
                            const double* dzc, const double* dzw, const double* Kw, double src0,
                            double* phi, const double* fcg, double* B, double* D) {
  auto kw = [&](int k) { return rhow[k] * Kw[k] / dzw[k]; };  // w level k, 1 <= k < nz
  for (int k = 0; k < nz; ++k) {
    const double below = k > 0 ? kw(k) : 0.0;
    const double above = k < nz - 1 ? kw(k + 1) : 0.0;
    const double mass = rho[k] * dzc[k] / dt;
    B[k] = mass + below + above;
    D[k] = mass * phi[k];
    if (fcg) D[k] += fcg[k] - fcg[k + 1];
  }
  D[0] += src0;
  for (int k = 1; k < nz; ++k) {
    const double c = kw(k);
    const double f = -c / B[k - 1];
    B[k] -= f * -c;
    D[k] -= f * D[k - 1];
  }
  phi[nz - 1] = D[nz - 1] / B[nz - 1];
  for (int k = nz - 2; k >= 0; --k) phi[k] = (D[k] + kw(k + 1) * phi[k + 1]) / B[k];
}

// k_sfc_scalar (surface.cu:87-387): one thread per interior column.
template <class T, int KMAX>
__global__ void __launch_bounds__(64) k_sfc_scalar(Grid g, Base<T> b, Metric<T> m,
                                                   SfcState<T> s, SfcArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int nz = g.nz, PL = g.plane;
  const int c2 = g.id2(i, j), c0 = g.id(i, j, 0);
  const double J = (double)m.jac[c2];
  const double dt = a.dt;
  double rho[KMAX], rhow[KMAX], dzc[KMAX], dzw[KMAX], uc[KMAX], vc[KMAX], thv[KMAX];
  double thcol[KMAX], qvcol[KMAX], Kw[KMAX + 1], fth[KMAX + 1], fqv[KMAX + 1];
  double lwp = 0;
  for (int k = 0; k < nz; ++k) {  // :101-114
    const int kk = k + g.ng, c = c0 + k * PL;
    rho[k] = (double)b.rhob[c];
    rhow[k] = (double)b.rhobw[c];
    dzc[k] = (double)m.dzeta_c[kk] * J;
    dzw[k] = (double)m.dzeta_w[kk] * J;
    uc[k] = 0.5 * ((double)s.u[c] + (double)s.u[c + 1]);
    vc[k] = 0.5 * ((double)s.v[c] + (double)s.v[c + g.NX]);
    thcol[k] = (double)b.thb[c] + (double)s.thp[c];
    qvcol[k] = (double)s.qv[c];
    thv[k] = thcol[k] * (1.0 + kEps61 * qvcol[k]);
    const double qc = (double)s.qc[c];
    lwp += rho[k] * (qc > 0 ? qc : 0.0) * dzc[k] * 1000.0;
  }

  // surface layer (Louis 1979), :117-159
  const bool onland = a.land[c2] > 0.5;
  const double z0 = onland ? 0.1 : 2e-4;
  const double alb = onland ? 0.2 : 0.08;
  double beta = 1.0;
  if (onland) {  // quadratic soil-moisture stress (wilting 0.08, field capacity 0.32)
    const double f = fmin(1.0, fmax(0.0, (a.soilw[c2] - 0.08) / (0.32 - 0.08)));
    beta = fmax(0.05, f * f);
  }
  const double z1 = 0.5 * dzc[0];
  double V1 = sqrt(uc[0] * uc[0] + vc[0] * vc[0]);
  if (V1 < 0.5) V1 = 0.5;
  const double pi1 = b.pib[c0] + s.pip[c0];
  const double p1 = kP00 * pow(pi1, kCp / kRd);
  const double Ts = a.tsk[c2];
  const double pis = pi1 + kGrav * z1 / (kCp * thv[0]);
  const double ths = Ts / pis;
  const double es = 610.78 * exp(17.269 * (Ts - 273.16) / (Ts - 35.86));  // thermo.hpp:11-18
  const double qs_s = 0.622 * es / fmax(p1 - es, 1.0);
  const double thvs = ths * (1.0 + kEps61 * (beta * qs_s + (1.0 - beta) * qvcol[0]));
  const double Rib = kGrav * z1 * (thv[0] - thvs) / (thvs * V1 * V1);
  double a2 = kKarman / log(z1 / z0);
  a2 *= a2;
  double Fm, Fh;
  if (Rib < 0) {
    const double sq = sqrt(-Rib * z1 / z0);
    Fm = 1.0 - 9.4 * Rib / (1.0 + 7.4 * 9.4 * a2 * sq);
    Fh = 1.0 - 9.4 * Rib / (1.0 + 5.3 * 9.4 * a2 * sq);
  } else {
    const double d = 1.0 + 4.7 * Rib;
    Fm = Fh = 1.0 / (d * d);
  }
  const double Cd = a2 * Fm, Ch = a2 * Fh;
  a.cdv[c2] = Cd * V1;

  // 2 m temperature and 10 m wind from flux-profile relations (:166-184)
  const double z0h = z0 * 0.1;
  const double lnz1z0 = log(z1 / z0);
  double zeta1 = Rib >= 0 ? Rib * lnz1z0 / (1.0 - 5.0 * fmin(Rib, 0.19)) : Rib * lnz1z0;
  zeta1 = fmax(-8.0, fmin(2.0, zeta1));
  const double zol = zeta1 / z1;
  const double den_m = lnz1z0 - sfc_psi_m(zeta1) + sfc_psi_m(z0 * zol);
  const double num_m = log(10.0 / z0) - sfc_psi_m(10.0 * zol) + sfc_psi_m(z0 * zol);
  a.u10[c2] = V1 * fmax(0.0, fmin(1.2, num_m / fmax(den_m, 0.1)));
  const double den_h = log(z1 / z0h) - sfc_psi_h(zeta1) + sfc_psi_h(z0h * zol);
  const double num_h = log(2.0 / z0h) - sfc_psi_h(2.0 * zol) + sfc_psi_h(z0h * zol);
  const double r2 = fmax(0.0, fmin(1.2, num_h / fmax(den_h, 0.1)));
  a.t2m[c2] = (ths + (thcol[0] - ths) * r2) * pi1;

  const double shf = Ch * V1 * (ths - thcol[0]);  // kinematic [K m/s]
  double qflx = Ch * V1 * beta * (qs_s - qvcol[0]);
  if (!onland && qflx < 0) qflx = 0;  // no dew over the sea

  // radiation column (:188-244): solar geometry, LW two-stream, SW beam
  const double utc = *a.utc;
  const double decl = 0.40928 * sin(6.2832 * (284 + a.doy) / 365.0);
  const double latr = a.lat[c2] * 0.0174533;
  const double H = 0.2618 * (utc + a.lon[c2] / 15 - 12);
  double cosz = sin(latr) * sin(decl) + cos(latr) * cos(decl) * cos(H);
  if (cosz < 0) cosz = 0;
  // emissivities into Kw (scratch until the diffusivities), layer blackbody
  // emission into fth; downward LW into fqv (levels 0..nz)
  double* eps = Kw;
  double* emis = fth;
  double* Fdn = fqv;
  for (int k = 0; k < nz; ++k) {
    const int c = c0 + k * PL;
    const double T4 = thcol[k] * (b.pib[c] + s.pip[c]);
    const double qc = (double)s.qc[c];
    const double du = rho[k] * qvcol[k] * dzc[k];
    const double dl = rho[k] * (qc > 0 ? qc : 0.0) * dzc[k] * 1000.0;
    eps[k] = 1.0 - exp(-(0.22 * du + 0.30 * dl));
    emis[k] = eps[k] * kStefan * T4 * T4 * T4 * T4;
  }
  Fdn[nz] = 0;
  for (int k = nz - 1; k >= 0; --k) Fdn[k] = Fdn[k + 1] * (1.0 - eps[k]) + emis[k];
  const double lwd = Fdn[0];
  // heating per layer: LW convergence (Fup recomputed upward on the fly)
  // plus SW absorption (downward beam)
  double radht[KMAX];
  {
    double Fup = 0.97 * kStefan * Ts * Ts * Ts * Ts + 0.03 * Fdn[0];
    for (int k = 0; k < nz; ++k) {
      const double Fup1 = Fup * (1.0 - eps[k]) + emis[k];
      radht[k] = (Fdn[k + 1] - Fdn[k]) + (Fup - Fup1);
      Fup = Fup1;
    }
  }
  const double mu = fmax(cosz, 0.05);
  const double cldalb = lwp / (lwp + 60.0);
  double Sdn = kSolar * cosz * 0.95 * (1.0 - cldalb);
  for (int k = nz - 1; k >= 0; --k) {
    const int c = c0 + k * PL;
    const double qc = (double)s.qc[c];
    const double du = rho[k] * qvcol[k] * dzc[k];
    const double dl = rho[k] * (qc > 0 ? qc : 0.0) * dzc[k] * 1000.0;
    const double tau = exp(-(0.008 * du + 0.15 * dl) / mu);
    radht[k] += Sdn * (1.0 - tau);
    Sdn *= tau;
  }
  const double swn = Sdn * (1.0 - alb);
  for (int k = 0; k < nz; ++k) {
    const int c = c0 + k * PL;
    radht[k] = radht[k] / (rho[k] * kCp * dzc[k]) / (b.pib[c] + s.pip[c]);
  }

  // four-layer implicit soil over land (:247-289)
  const double lwu = 0.97 * kStefan * Ts * Ts * Ts * Ts;
  const double G = swn + 0.97 * lwd - lwu - rho[0] * kCp * shf * pi1 - rho[0] * kLv * qflx;
  if (onland) {
    const double sdz[4] = {0.1, 0.3, 0.6, 1.0};
    const double w = fmin(fmax(a.soilw[c2], 0.0), 0.4);
    const double Cso = 1.2e6 + 3.2e6 * w, Kso = 0.25 + 3.0 * w;
    double Tso[4], A[4], Bs[4], C[4], Ds[4];
    for (int L = 0; L < 4; ++L) Tso[L] = a.soilt[L * PL + c2];
    for (int L = 0; L < 4; ++L) {
      const double cap = Cso * sdz[L] / dt;
      const double ku = L > 0 ? Kso / (0.5 * (sdz[L - 1] + sdz[L])) : 0.0;
      const double kl = L < 3 ? Kso / (0.5 * (sdz[L] + sdz[L + 1])) : Kso / sdz[L];
      A[L] = -ku;
      C[L] = L < 3 ? -kl : 0.0;
      Bs[L] = cap + ku + kl;
      Ds[L] = cap * Tso[L];
    }
    Ds[0] += G;
    Ds[3] += (Kso / sdz[3]) * a.tdeep[c2];
    for (int L = 1; L < 4; ++L) {
      const double f = A[L] / Bs[L - 1];
      Bs[L] -= f * C[L - 1];
      Ds[L] -= f * Ds[L - 1];
    }
    Tso[3] = Ds[3] / Bs[3];
    for (int L = 2; L >= 0; --L) Tso[L] = (Ds[L] - C[L] * Tso[L + 1]) / Bs[L];
    for (int L = 0; L < 4; ++L) {
      Tso[L] = fmin(fmax(Tso[L], 200.0), 350.0);
      a.soilt[L * PL + c2] = Tso[L];
    }
    a.tsk[c2] = Tso[0];
  }

  // diffusivity at w levels 1..nz-1 (:292-382)
  auto local_K = [&](int k) -> double {
    const double dz = dzw[k];
    const double du = (uc[k] - uc[k - 1]) / dz;
    const double dv = (vc[k] - vc[k - 1]) / dz;
    const double S2 = du * du + dv * dv + 1e-8;
    const double thvw = 0.5 * (thv[k - 1] + thv[k]);
    const double N2 = kGrav / thvw * (thv[k] - thv[k - 1]) / dz;
    const double Ri = N2 / S2;
    const double zw = (double)m.zeta_w[k + g.ng] * J;
    const double l = kKarman * zw / (1.0 + kKarman * zw / 150.0);
    if (Ri < 0) return l * l * sqrt(S2) * sqrt(1.0 - 16.0 * Ri);
    const double d = 1.0 + 5.0 * Ri;
    return l * l * sqrt(S2) / (d * d);
  };
  for (int k = 0; k <= nz; ++k) fth[k] = fqv[k] = 0.0;
  Kw[0] = 0;
  if (a.nonlocal) {
    const double ust = sqrt(Cd) * V1;
    const double thv1 = thv[0];
    const double wthv0 = shf * (1.0 + kEps61 * qvcol[0]) + kEps61 * thcol[0] * qflx;
    constexpr double Ric = 0.25;
    const double ztop = (double)m.zeta_c[(nz - 1) + g.ng] * J;
    double hpbl = ztop;
    double ws = ust > 0.1 ? ust : 0.1;
    for (int it = 0; it < 3; ++it) {  // boundary-layer height <-> w* iteration
      const double thv_s = thv1 + (wthv0 > 0 ? 6.8 * wthv0 / ws : 0.0);
      double prevRb = 0;
      hpbl = ztop;
      for (int k = 1; k < nz; ++k) {
        const double zc = (double)m.zeta_c[k + g.ng] * J;
        const double shear = uc[k] * uc[k] + vc[k] * vc[k] + 0.1;
        const double Rb = kGrav * (thv[k] - thv_s) * zc / (thv1 * shear);
        if (Rb >= Ric) {
          const double zcm = (double)m.zeta_c[(k - 1) + g.ng] * J;
          const double f = fmin(1.0, fmax(0.0, (Ric - prevRb) / (Rb - prevRb + 1e-9)));
          hpbl = zcm + f * (zc - zcm);
          break;
        }
        prevRb = Rb;
      }
      const double zc0 = (double)m.zeta_c[g.ng] * J;
      if (hpbl < zc0) hpbl = zc0;
      const double wstar = wthv0 > 0 ? cbrt(kGrav / thv1 * wthv0 * hpbl) : 0.0;
      ws = cbrt(ust * ust * ust + 0.6 * wstar * wstar * wstar);
      if (ws < 0.1) ws = 0.1;
    }
    a.pblh[c2] = hpbl;
    const double gth = wthv0 > 0 ? 6.8 * shf / (ws * hpbl) : 0.0;
    const double gqv = wthv0 > 0 ? 6.8 * qflx / (ws * hpbl) : 0.0;
    for (int k = 1; k < nz; ++k) {
      const double zw = (double)m.zeta_w[k + g.ng] * J;
      double Kv;
      if (wthv0 > 0 && zw < hpbl) {  // convective boundary layer: nonlocal profile
        const double zf = zw / hpbl;
        Kv = kKarman * ws * zw * (1 - zf) * (1 - zf);
        fth[k] = rhow[k] * Kv * gth;
        fqv[k] = rhow[k] * Kv * gqv;
      } else {
        Kv = local_K(k);
      }
      Kv = fmin(fmax(Kv, kKmMin), kKmMax);
      Kw[k] = Kv;
      a.km[c0 + k * PL] = Kv;
    }
  } else {
    a.pblh[c2] = 0;
    for (int k = 1; k < nz; ++k) {
      const double Kv = fmin(fmax(local_K(k), kKmMin), kKmMax);
      Kw[k] = Kv;
      a.km[c0 + k * PL] = Kv;
    }
  }

  // implicit mixing of theta and qv with the surface fluxes (:385-386),
  // then the radiative heating, limited to 0.02 K per step (:389-397)
  double Bw[KMAX], Dw[KMAX];
  sfc_diffuse<KMAX>(nz, dt, rho, rhow, dzc, dzw, Kw, rho[0] * shf, thcol, fth, Bw, Dw);
  sfc_diffuse<KMAX>(nz, dt, rho, rhow, dzc, dzw, Kw, rho[0] * qflx, qvcol, fqv, Bw, Dw);
  for (int k = 0; k < nz; ++k) {
    const int c = c0 + k * PL;
    const double dth = fmin(fmax(dt * radht[k], -0.02), 0.02);
    s.thp[c] = (T)(thcol[k] - (double)b.thb[c] + dth);
    s.qv[c] = (T)(qvcol[k] > 0 ? qvcol[k] : 0.0);
  }
}

// k_sfc_u / k_sfc_v (surface.cu:390-437): momentum columns at the inner u
// faces i = 1..nx-1 (DIR 0) or v faces j = 1..ny-1 (DIR 1), averaged from the
// two cells sharing the face: explicit surface drag, implicit mixing with
// the mean diffusivity of the two cells.
template <class T, int KMAX, int DIR>
__global__ void __launch_bounds__(64) k_sfc_face(Grid g, Base<T> b, Metric<T> m, T* f,
                                                 SfcArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + (DIR == 0 ? 1 : 0);
  const int j = blockIdx.y + (DIR == 1 ? 1 : 0);
  if (i >= g.nx || j >= g.ny) return;
  const int nz = g.nz, PL = g.plane;
  const int sh = DIR == 0 ? 1 : g.NX;  // offset to the other cell of the face
  const int c2 = g.id2(i, j), c0 = g.id(i, j, 0);
  const double J = 0.5 * ((double)m.jac[c2 - sh] + (double)m.jac[c2]);
  double rho[KMAX], rhow[KMAX], dzc[KMAX], dzw[KMAX], Kw[KMAX], phi[KMAX], Bw[KMAX], Dw[KMAX];
  for (int k = 0; k < nz; ++k) {
    const int kk = k + g.ng, c = c0 + k * PL;
    rho[k] = 0.5 * ((double)b.rhob[c - sh] + (double)b.rhob[c]);
    rhow[k] = 0.5 * ((double)b.rhobw[c - sh] + (double)b.rhobw[c]);
    dzc[k] = (double)m.dzeta_c[kk] * J;
    dzw[k] = (double)m.dzeta_w[kk] * J;
    Kw[k] = k > 0 ? 0.5 * (a.km[c - sh] + a.km[c]) : 0.0;
    phi[k] = (double)f[c];
  }
  const double cdva = 0.5 * (a.cdv[c2 - sh] + a.cdv[c2]);
  sfc_diffuse<KMAX>(nz, a.dt, rho, rhow, dzc, dzw, Kw, -rho[0] * cdva * phi[0], phi, nullptr,
                    Bw, Dw);
  for (int k = 0; k < nz; ++k) f[c0 + k * PL] = (T)phi[k];
}

}  // namespace wfe::dev
