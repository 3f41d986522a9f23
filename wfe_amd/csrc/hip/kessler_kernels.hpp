#pragma once
// Kessler (1969) warm rain with the reference's mixed-phase blend, applied
// after the large step by operator splitting (k_kessler,
// src/physics/kessler.cu:22-139): CFL-subcycled upwind rain sedimentation
// with surface accumulation, autoconversion, accretion, rain evaporation,
// one-Newton-step saturation adjustment. One thread per column; the
// sedimentation is serial in k with a data-dependent subcycle count.
#include "hip/device_common.hpp"

namespace wfe::dev {

// Tetens saturation mixing ratio over water / ice with the 1 Pa floor on
// p - e_s (src/core/thermo.hpp:11-26).
template <class T>
__device__ __forceinline__ T qsat_w(T p, T temp) {
  const T es = (T)610.78 * exp((T)17.269 * (temp - (T)273.16) / (temp - (T)35.86));
  T den = p - es;
  if (den < (T)1) den = (T)1;
  return (T)0.622 * es / den;
}
template <class T>
__device__ __forceinline__ T qsat_i(T p, T temp) {
  const T es = (T)610.78 * exp((T)21.875 * (temp - (T)273.16) / (temp - (T)7.66));
  T den = p - es;
  if (den < (T)1) den = (T)1;
  return (T)0.622 * es / den;
}

template <class T>
struct KesArgs {
  T* thp;
  const double* pip;
  T* qv;
  T* qc;
  T* qr;
  T* rain;  // 2-D
  T dt;
};

template <class T, int MAXL>
__global__ void __launch_bounds__(64) k_kessler_col(Grid g, Base<T> b, Metric<T> m,
                                                    KesArgs<T> a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const T J = m.jac[g.id2(i, j)];
  const T k1 = (T)0.001, qc0 = (T)0.001;
  const T dt = a.dt;
  const int PL = g.plane;
  const int c0 = g.id(i, j, 0);
  T qrl[MAXL], vr[MAXL];

  // Falling hydrometeor below freezing falls ~3x slower (snow-like).
  auto fallfac = [&](int c) -> T {
    const T temp = (T)(((double)b.thb[c] + (double)a.thp[c]) * (b.pib[c] + a.pip[c]));
    return temp < (T)kTfrz ? (T)0.35 : (T)1;
  };
  auto vterm = [&](int c, T rho, T q) -> T {
    const T rq = rho * q;
    return rq > (T)1e-9 ? (T)36.34 * pow(rq, (T)0.1364) * sqrt(b.rhob[c0] / rho) * fallfac(c)
                        : (T)0;
  };

  // ---- 1) sedimentation (first-order upwind, CFL-subcycled)
  T vrmax = 0, dzmin = (T)1e30;
  for (int k = 0; k < g.nz; ++k) {
    const int c = c0 + k * PL;
    const T q = a.qr[c];
    qrl[k] = q > (T)0 ? q : (T)0;
    vr[k] = vterm(c, b.rhob[c], qrl[k]);
    const T dzl = m.dzeta_c[k + g.ng] * J;
    if (vr[k] > vrmax) vrmax = vr[k];
    if (dzl < dzmin) dzmin = dzl;
  }
  const int nsub = 1 + (int)(vrmax * dt / ((T)0.8 * dzmin));
  const T dts = dt / (T)nsub;
  T acc = 0;
  for (int s = 0; s < nsub; ++s) {
    acc += dts * (b.rhob[c0] * qrl[0] * vr[0]);
    for (int k = 0; k < g.nz; ++k) {
      const int c = c0 + k * PL;
      const T ftop = (k + 1 < g.nz) ? b.rhob[c + PL] * qrl[k + 1] * vr[k + 1] : (T)0;
      const T fbot = b.rhob[c] * qrl[k] * vr[k];
      qrl[k] += dts * (ftop - fbot) / (b.rhob[c] * (m.dzeta_c[k + g.ng] * J));
    }
    if (s + 1 < nsub)
      for (int k = 0; k < g.nz; ++k) {
        const int c = c0 + k * PL;
        vr[k] = vterm(c, b.rhob[c], qrl[k] > (T)0 ? qrl[k] : (T)0);
      }
  }
  a.rain[g.id2(i, j)] += acc;

  // ---- 2+3) conversions + saturation adjustment, level by level
  for (int k = 0; k < g.nz; ++k) {
    const int c = c0 + k * PL;
    const T rhol = b.rhob[c];
    T qvk = a.qv[c] > (T)0 ? a.qv[c] : (T)0;
    T qck = a.qc[c] > (T)0 ? a.qc[c] : (T)0;
    T qrk = qrl[k] > (T)0 ? qrl[k] : (T)0;
    const double pid = b.pib[c] + a.pip[c];  // pibar + pi' in FP64
    const T pi = (T)pid;
    T th = b.thb[c] + a.thp[c];
    T temp = th * pi;
    const T pres = (T)(kP00 * pow(pid, kCp / kRd));
    T fl = (temp - (T)233.15) / (T)40;
    fl = fmax((T)0, fmin((T)1, fl));
    const T qsw = qsat_w(pres, temp);
    T qs = fl * qsw + ((T)1 - fl) * qsat_i(pres, temp);
    const T Leff = fl * (T)kLv + ((T)1 - fl) * (T)kLs;

    const T dqauto = k1 * (qck > qc0 ? qck - qc0 : (T)0);
    const T dqacc = (T)2.2 * qck * pow(qrk > (T)1e-9 ? qrk : (T)1e-9, (T)0.875);
    T dqr = dt * (dqauto + dqacc);
    if (dqr > qck) dqr = qck;
    qck -= dqr;
    qrk += dqr;

    T thpk = a.thp[c];
    if (qvk < qs && qrk > (T)0) {
      const T rq = rhol * qrk;
      const T vent = (T)1.6 + (T)124.9 * pow(rq, (T)0.2046);
      const T ern = vent * pow(rq, (T)0.525) / ((T)2.55e8 / (pres * qs) + (T)5.4e5) *
                    (qs - qvk) / (rhol * qs);
      T dqe = dt * ern;
      if (dqe > qrk) dqe = qrk;
      if (dqe > qs - qvk) dqe = qs - qvk;
      qrk -= dqe;
      qvk += dqe;
      thpk -= Leff * dqe / ((T)kCp * pi);
      th = b.thb[c] + thpk;
      temp = th * pi;
      const T fl2 = fmax((T)0, fmin((T)1, (temp - (T)233.15) / (T)40));
      qs = fl2 * qsat_w(pres, temp) + ((T)1 - fl2) * qsat_i(pres, temp);
    }
    const T gam = Leff * Leff * qs / ((T)kCp * (T)kRv * temp * temp);
    T dq = (qvk - qs) / ((T)1 + gam);
    if (dq < (T)0) {
      T evap = -dq;
      if (evap > qck) evap = qck;
      dq = -evap;
    }
    qvk -= dq;
    qck += dq;
    thpk += Leff * dq / ((T)kCp * pi);
    a.thp[c] = thpk;
    a.qv[c] = qvk > (T)0 ? qvk : (T)0;
    a.qc[c] = qck > (T)0 ? qck : (T)0;
    a.qr[c] = qrk > (T)0 ? qrk : (T)0;
  }
}

}  // namespace wfe::dev
