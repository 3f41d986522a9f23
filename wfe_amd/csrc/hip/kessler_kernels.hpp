#pragma once
// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.
// Kessler (1969) warm rain with the reference's mixed-phase blend, applied
// after the large step by operator splitting (k_kessler,
// src/physics/kessler.cu:22-139): CFL-subcycled upwind rain sedimentation
// with surface accumulation, autoconversion, accretion, rain evaporation,
// one-Newton-step saturation adjustment.
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

// x^y for x > 0 as exp(y ln x) in FP64 (the library pow carries extra
// precision for all x and y that these rates do not need; the difference is
// a few ulps).
__device__ __forceinline__ double pow_pos(double x, double y) { return exp(y * log(x)); }

// Autoconversion, accretion, rain evaporation and the one-Newton-step
// saturation adjustment of one point (kessler.cu:99-158), in FP64 whatever
// the storage type: theta = thetabar + theta' and the saturation deficit
// q_v - q_s are small differences of large numbers.
template <class T>
__device__ __forceinline__ void kessler_convert(T& thp_io, T& qv_io, T& qc_io, T& qr_io,
                                                double pid, T thb_t, T rho_t, T dt_t) {
  using W = double;
  const W k1 = 0.001, qc0 = 0.001;
  const W dt = (W)dt_t, rhol = (W)rho_t, thb = (W)thb_t;
  W qvk = qv_io > (T)0 ? (W)qv_io : 0.0;
  W qck = qc_io > (T)0 ? (W)qc_io : 0.0;
  W qrk = qr_io > (T)0 ? (W)qr_io : 0.0;
  const W pi = pid;
  W th = thb + (W)thp_io;
  W temp = th * pi;
  const W pres = kP00 * pow_pos(pid, kCp / kRd);
  W fl = (temp - 233.15) / 40.0;
  fl = fmax(0.0, fmin(1.0, fl));
  const W qsw = qsat_w(pres, temp);
  W qs = fl * qsw + (1.0 - fl) * qsat_i(pres, temp);
  const W Leff = fl * kLv + (1.0 - fl) * kLs;
  const W dqauto = k1 * (qck > qc0 ? qck - qc0 : 0.0);
  const W dqacc = 2.2 * qck * pow_pos(qrk > 1e-9 ? qrk : 1e-9, 0.875);
  W dqr = dt * (dqauto + dqacc);
  if (dqr > qck) dqr = qck;
  qck -= dqr;
  qrk += dqr;
  W thpk = (W)thp_io;
  if (qvk < qs && qrk > 0.0) {
    const W rq = rhol * qrk;
    const W lrq = log(rq);
    const W vent = 1.6 + 124.9 * exp(0.2046 * lrq);
    const W ern =
        vent * exp(0.525 * lrq) / (2.55e8 / (pres * qs) + 5.4e5) * (qs - qvk) / (rhol * qs);
    W dqe = dt * ern;
    if (dqe > qrk) dqe = qrk;
    if (dqe > qs - qvk) dqe = qs - qvk;
    qrk -= dqe;
    qvk += dqe;
    thpk -= Leff * dqe / (kCp * pi);
    th = thb + thpk;
    temp = th * pi;
    const W fl2 = fmax(0.0, fmin(1.0, (temp - 233.15) / 40.0));
    qs = fl2 * qsat_w(pres, temp) + (1.0 - fl2) * qsat_i(pres, temp);
  }
  const W gam = Leff * Leff * qs / (kCp * kRv * temp * temp);
  W dq = (qvk - qs) / (1.0 + gam);
  if (dq < 0.0) {
    W evap = -dq;
    if (evap > qck) evap = qck;
    dq = -evap;
  }
  qvk -= dq;
  qck += dq;
  thpk += Leff * dq / (kCp * pi);
  thp_io = (T)thpk;
  qv_io = (T)(qvk > 0.0 ? qvk : 0.0);
  qc_io = (T)(qck > 0.0 ? qck : 0.0);
  qr_io = (T)(qrk > 0.0 ? qrk : 0.0);
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

// One thread per column (any nz): the sedimentation runs serially in k.
template <class T, int MAXL>
__global__ void __launch_bounds__(64) k_kessler_col(Grid g, Base<T> b, Metric<T> m,
                                                    KesArgs<T> a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const T J = m.jac[g.id2(i, j)];
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
    T thp = a.thp[c], qv = a.qv[c], qc = a.qc[c], qr = qrl[k];
    kessler_convert(thp, qv, qc, qr, b.pib[c] + a.pip[c], b.thb[c], b.rhob[c], dt);
    a.thp[c] = thp;
    a.qv[c] = qv;
    a.qc[c] = qc;
    a.qr[c] = qr;
  }
}

// k-parallel form of the same column physics. A workgroup owns TX columns of
// one row and NW thread rows (levels k = r, r + NW, ...). The reference's
// sedimentation update reads the level above before that level is updated
// (kessler.cu:79-86), so every subcycle is a parallel sweep over k: fluxes
// F_k = rho_k q_k V_k to LDS, barrier, q_k += dts (F_k+1 - F_k)/(rho_k dz_k).
// The subcycle count is per column; the workgroup loops to its maximum.
template <class T, int TX, int NW, int KPT>
__global__ void __launch_bounds__(TX * NW) k_kessler_tile(Grid g, Base<T> b, Metric<T> m,
                                                          KesArgs<T> a) {
  extern __shared__ double kes_lds[];
  T* sF = reinterpret_cast<T*>(kes_lds);  // [nz + 1][TX]
  T* sR1 = sF + (g.nz + 1) * TX;          // [NW][TX] reduction scratch (max V)
  T* sR2 = sR1 + NW * TX;                 // [NW][TX] reduction scratch (min dz)
  __shared__ int s_nsub[TX];
  __shared__ int s_nmax;
  const int lx = threadIdx.x, ry = threadIdx.y;
  const BlockId bid = xcd_block();
  const int i = bid.x * TX + lx;
  const int j = bid.y;
  const bool col = i < g.nx;
  const int PL = g.plane;
  const int c0 = g.id(col ? i : 0, j, 0);
  const T J = m.jac[g.id2(col ? i : 0, j)];
  const T dt = a.dt;
  const int cb0 = b.hom ? g.id(0, 0, 0) : c0;  // base-state column
  const T rho_sfc = b.rhob[cb0];

  T qrl[KPT], vr[KPT], rho[KPT], dzl[KPT], ff[KPT];
  T vmax = 0, dzmin = (T)1e30;
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    qrl[t] = 0;
    vr[t] = 0;
    rho[t] = 1;
    dzl[t] = 1;
    ff[t] = 1;
    if (!col || k >= g.nz) continue;
    const int c = c0 + k * PL;
    const T q = a.qr[c];
    qrl[t] = q > (T)0 ? q : (T)0;
    const int cb = cb0 + k * PL;
    rho[t] = b.rhob[cb];
    dzl[t] = m.dzeta_c[k + g.ng] * J;
    const T temp = (T)(((double)b.thb[cb] + (double)a.thp[c]) * (b.pib[cb] + a.pip[c]));
    ff[t] = temp < (T)kTfrz ? (T)0.35 : (T)1;
    const T rq = rho[t] * qrl[t];
    vr[t] = rq > (T)1e-9 ? (T)36.34 * pow(rq, (T)0.1364) * sqrt(rho_sfc / rho[t]) * ff[t] : (T)0;
    if (vr[t] > vmax) vmax = vr[t];
    if (dzl[t] < dzmin) dzmin = dzl[t];
  }
  sR1[ry * TX + lx] = vmax;
  sR2[ry * TX + lx] = dzmin;
  __syncthreads();
  if (ry == 0) {
    for (int r = 1; r < NW; ++r) {
      const T v = sR1[r * TX + lx], d = sR2[r * TX + lx];
      if (v > vmax) vmax = v;
      if (d < dzmin) dzmin = d;
    }
    s_nsub[lx] = col ? 1 + (int)(vmax * dt / ((T)0.8 * dzmin)) : 0;
  }
  __syncthreads();
  if (ry == 0 && lx == 0) {
    int mx = 0;
    for (int q = 0; q < TX; ++q) mx = s_nsub[q] > mx ? s_nsub[q] : mx;
    s_nmax = mx;
  }
  __syncthreads();
  const int nsub = s_nsub[lx];
  const int nmax = s_nmax;
  const T dts = col ? dt / (T)nsub : (T)0;
  T acc = 0;
  for (int sc = 0; sc < nmax; ++sc) {
    const bool act = sc < nsub;
#pragma unroll
    for (int t = 0; t < KPT; ++t) {
      const int k = ry + NW * t;
      if (act && k < g.nz) sF[k * TX + lx] = rho[t] * qrl[t] * vr[t];
    }
    if (ry == 0) sF[g.nz * TX + lx] = (T)0;
    __syncthreads();
    if (act) {
      if (ry == 0) acc += dts * sF[lx];
#pragma unroll
      for (int t = 0; t < KPT; ++t) {
        const int k = ry + NW * t;
        if (k >= g.nz) continue;
        const T ftop = sF[(k + 1) * TX + lx];
        const T fbot = sF[k * TX + lx];
        qrl[t] += dts * (ftop - fbot) / (rho[t] * dzl[t]);
        if (sc + 1 < nsub) {
          const T rq = rho[t] * (qrl[t] > (T)0 ? qrl[t] : (T)0);
          vr[t] = rq > (T)1e-9
                      ? (T)36.34 * pow(rq, (T)0.1364) * sqrt(rho_sfc / rho[t]) * ff[t]
                      : (T)0;
        }
      }
    }
    __syncthreads();
  }
  if (col && ry == 0) a.rain[g.id2(i, j)] += acc;

  // conversions + saturation adjustment, level by level (kessler.cu:99-158)
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    if (!col || k >= g.nz) continue;
    const int c = c0 + k * PL;
    T thp = a.thp[c], qv = a.qv[c], qc = a.qc[c], qr = qrl[t];
    const int cb = cb0 + k * PL;
    kessler_convert(thp, qv, qc, qr, b.pib[cb] + a.pip[c], b.thb[cb], rho[t], dt);
    a.thp[c] = thp;
    a.qv[c] = qv;
    a.qc[c] = qc;
    a.qr[c] = qr;
  }
}

template <int TX, int NW, class T>
inline size_t kessler_tile_lds(int nz) {
  return (size_t)((nz + 1) * TX + 2 * NW * TX) * sizeof(T);
}

}  // namespace wfe::dev
