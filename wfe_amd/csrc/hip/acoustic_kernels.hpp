#pragma once
// Split-explicit acoustic substep (acoustic_substep, kernels.cu:929-965):
// a horizontal forward step of u, v and a vertically implicit w-pi' column
// solve. The fast system restarts from time n each stage; instead of the
// reference's full s_work copy, the first substep reads the time-n buffers
// directly and later substeps read the work buffers. Ghost reads follow the
// values the reference's fill schedule leaves there:
//  - pi' (current substep)   : composed lateral/vertical map of the interior;
//  - pi'prev, theta', qv     : the time-n ghost cells (never refreshed in the
//                              reference's substeps).
#include "hip/device_common.hpp"

namespace wfe::dev {

template <class T>
struct UVArgs {
  const T* u_src;
  const T* v_src;
  T* u_dst;
  T* v_dst;
  const double* pcur;   // pi' entering the substep
  const double* pprev;  // pi' of the previous substep (time n at substep 0)
  const double* pn;     // time-n pi' (ghost cells filled)
  const T* thp_cur;
  const T* thp_n;  // time-n theta' (ghost cells filled)
  const T* qv;     // stage-constant qv (time n, ghosts filled)
  const T* tu;
  const T* tv;
  T dtau;
};

template <class T>
struct UVDev {
  const Grid& g;
  const Params& p;
  const Base<T>& b;
  const Metric<T>& m;
  const UVArgs<T>& a;

  // pi* = pi' + smdiv (pi' - pi'prev) at any point of the padded stencil, FP64.
  __device__ __forceinline__ double pstar(int ii, int jj, int kz) const {
    const int c = g.id(ii, jj, kz);
    double pc, pp;
    if ((unsigned)ii < (unsigned)g.nx && (unsigned)jj < (unsigned)g.ny &&
        (unsigned)kz < (unsigned)g.nz) {
      pc = a.pcur[c];
      pp = a.pprev[c];
    } else {
      const int im = map_lat(ii, g.nx, g.nx - 1, p.bcx);
      const int jm = map_lat(jj, g.ny, g.ny - 1, p.bcy);
      const int km = kz < 0 ? 0 : (kz > g.nz - 1 ? g.nz - 1 : kz);
      pc = a.pcur[g.id(im, jm, km)];
      pp = a.pn[c];
    }
    return pc + p.smdiv * (pc - pp);
  }
  // Full virtual potential temperature theta_v at cell (ii, jj, k).
  __device__ __forceinline__ T thv(int ii, int jj, int k) const {
    const int c = g.id(ii, jj, k);
    const bool in = (unsigned)ii < (unsigned)g.nx && (unsigned)jj < (unsigned)g.ny;
    T t = b.thb[c] + (in ? a.thp_cur[c] : a.thp_n[c]);
    if (p.moisture) t *= (T)1 + (T)kEps61 * a.qv[c];
    return t;
  }
  // u after the horizontal PGF step at face (i, j, k) (k_acou_uv :496-512).
  __device__ T u_step(int i, int j, int k) const {
    const int kk = k + g.ng;
    const int c = g.id(i, j, k);
    const int c2 = g.id2(i, j);
    const double ps0 = pstar(i, j, k);
    const T thbu = (T)0.5 * (thv(i - 1, j, k) + thv(i, j, k));
    const T mu = (T)0.5 * (m.mapf[c2 - 1] + m.mapf[c2]);
    double grad = (ps0 - pstar(i - 1, j, k)) / g.dx;
    const T zxu = m.hxu[c2] * ((T)1 - m.zeta_c[kk] / m.zt);
    if (zxu != (T)0) {
      const T dzc2 = m.zeta_c[kk + 1] - m.zeta_c[kk - 1];
      const T ju = (T)0.5 * (m.jac[c2 - 1] + m.jac[c2]);
      const double dpdz = 0.5 *
                          ((pstar(i - 1, j, k + 1) - pstar(i - 1, j, k - 1)) +
                           (pstar(i, j, k + 1) - pstar(i, j, k - 1))) /
                          (double)dzc2;
      grad -= (double)(zxu / ju) * dpdz;
    }
    return a.u_src[c] + a.dtau * (-(T)kCp * thbu * mu * (T)grad + a.tu[c]);
  }
  __device__ T v_step(int i, int j, int k) const {
    const int kk = k + g.ng;
    const int c = g.id(i, j, k);
    const int c2 = g.id2(i, j);
    const double ps0 = pstar(i, j, k);
    const T thbv = (T)0.5 * (thv(i, j - 1, k) + thv(i, j, k));
    const T mv = (T)0.5 * (m.mapf[c2 - g.NX] + m.mapf[c2]);
    double grad = (ps0 - pstar(i, j - 1, k)) / g.dy;
    const T zyv = m.hyv[c2] * ((T)1 - m.zeta_c[kk] / m.zt);
    if (zyv != (T)0) {
      const T dzc2 = m.zeta_c[kk + 1] - m.zeta_c[kk - 1];
      const T jv = (T)0.5 * (m.jac[c2 - g.NX] + m.jac[c2]);
      const double dpdz = 0.5 *
                          ((pstar(i, j - 1, k + 1) - pstar(i, j - 1, k - 1)) +
                           (pstar(i, j, k + 1) - pstar(i, j, k - 1))) /
                          (double)dzc2;
      grad -= (double)(zyv / jv) * dpdz;
    }
    return a.v_src[c] + a.dtau * (-(T)kCp * thbv * mv * (T)grad + a.tv[c]);
  }
};

// Horizontal step of u on faces i = 0..nx and v on faces j = 0..ny
// (k_acou_uv kernels.cu:480-527) fused with what the reference does next
// before the column solve: Klemp-Wilhelmson radiation on open edges
// (k_bc_rad_u_x / k_bc_rad_v_y kernels.cu:649-688, inflow -> u_b / v_b) and
// the periodic ghost u(nx) = u(0), v(ny) = v(0) (bc_lateral_x/y :947-950).
// Neighbour values the radiation needs (u at 1 and nx-1) are recomputed.
template <class T>
__global__ void k_acou_uv(Grid g, Params p, Base<T> b, Metric<T> m, UVArgs<T> a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i > g.nx || j > g.ny || k >= g.nz) return;
  const UVDev<T> d{g, p, b, m, a};
  const int kk = k + g.ng;
  const int c = g.id(i, j, k);
  const T dtau = a.dtau;
  if (j < g.ny) {
    T un;
    if (!p.bcx) {
      un = d.u_step(i < g.nx ? i : 0, j, k);
    } else if (i == 0) {
      const T u0 = a.u_src[c];
      if (u0 >= (T)0) {
        un = b.ub[kk];
      } else {
        const T cb = u0 - (T)p.cstar;
        un = u0 - dtau * cb * (d.u_step(1, j, k) - u0) / (T)g.dx;
      }
    } else if (i == g.nx) {
      const T u1 = a.u_src[c];
      if (u1 <= (T)0) {
        un = b.ub[kk];
      } else {
        const T cb = u1 + (T)p.cstar;
        un = u1 - dtau * cb * (u1 - d.u_step(g.nx - 1, j, k)) / (T)g.dx;
      }
    } else {
      un = d.u_step(i, j, k);
    }
    a.u_dst[c] = un;
  }
  if (i < g.nx) {
    T vn;
    if (!p.bcy) {
      vn = d.v_step(i, j < g.ny ? j : 0, k);
    } else if (j == 0) {
      const T v0 = a.v_src[c];
      if (v0 >= (T)0) {
        vn = b.vb[kk];
      } else {
        const T cb = v0 - (T)p.cstar;
        vn = v0 - dtau * cb * (d.v_step(i, 1, k) - v0) / (T)g.dy;
      }
    } else if (j == g.ny) {
      const T v1 = a.v_src[c];
      if (v1 <= (T)0) {
        vn = b.vb[kk];
      } else {
        const T cb = v1 + (T)p.cstar;
        vn = v1 - dtau * cb * (v1 - d.v_step(i, g.ny - 1, k)) / (T)g.dy;
      }
    } else {
      vn = d.v_step(i, j, k);
    }
    a.v_dst[c] = vn;
  }
}

template <class T>
struct WPArgs {
  const T* u;  // after this substep's horizontal step (faces 0..nx)
  const T* v;
  const T* w_src;
  const double* pip_src;
  const T* thp_src;
  T* w_dst;
  double* pip_dst;
  double* piprev_dst;
  T* thp_dst;
  const T* tw;
  const T* tth;
  const T* tpi;  // nullptr when the pi' slow tendency is zero
  const T* qv;
  const T* qc;
  const T* qr;
  double dtau;
};

// Vertically implicit w-pi' solve, one thread per column (k_acou_wpi
// kernels.cu:531-645), in FP64 in every precision mode:
//   P_k = pi' + dtau T_pi - dtau K_t (D_h - d(S_v)/dzeta + c2 d(A w)/dzeta)
//   E_k = dtau K_t c1 / dzeta
//   -G E_{k-1} A_{k-1} w_{k-1} + (1 + G A_k (E_k + E_{k-1})) w_k
//       - G E_k A_{k+1} w_{k+1} = RHS_k - G (P_k - P_{k-1})
// Thomas elimination bottom-up, back substitution top-down, diagnostic
// surface w(0) = u h_x + v h_y, then pi', pi'prev and theta' updates.
template <class T, int MAXL>
__global__ void __launch_bounds__(64) k_acou_wpi_col(Grid g, Params p, Base<T> b, Metric<T> m,
                                                     Coef<T> cf, WPArgs<T> a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const double c1 = (1.0 + p.beta) * 0.5;
  const double c2 = (1.0 - p.beta) * 0.5;
  const int c2d = g.id2(i, j);
  const double J = (double)m.jac[c2d];
  const double hxc = (double)m.hx[c2d];
  const double hyc = (double)m.hy[c2d];
  const double mc = (double)m.mapf[c2d];
  const double dtau = a.dtau;
  const double zt = (double)m.zt;
  const int NX = g.NX, PL = g.plane;
  const int c0 = g.id(i, j, 0);

  double P[MAXL], E[MAXL], cs[MAXL], ds[MAXL];

  // Sv(k): explicit cross-metric vertical mass flux at w level k.
  auto Sv_at = [&](int k) -> double {
    if (k <= 0 || k >= g.nz) return 0.0;
    const int c = c0 + k * PL;
    const double av = (double)cf.aw[c];
    const double fac = 1.0 - (double)m.zeta_w[k + g.ng] / zt;
    const double uw = 0.25 * ((double)a.u[c - PL] + (double)a.u[c - PL + 1] + (double)a.u[c] +
                              (double)a.u[c + 1]);
    const double vw = 0.25 * ((double)a.v[c - PL] + (double)a.v[c - PL + NX] +
                              (double)a.v[c] + (double)a.v[c + NX]);
    return av * (uw * hxc + vw * hyc) * fac;
  };
  auto Av_at = [&](int k) -> double {
    if (k <= 0 || k >= g.nz) return 0.0;
    return (double)cf.aw[c0 + k * PL];
  };

  double sv_lo = 0.0, av_lo = 0.0;
  for (int k = 0; k < g.nz; ++k) {
    const int c = c0 + k * PL;
    const double sv_hi = Sv_at(k + 1), av_hi = Av_at(k + 1);
    const double Kt = (double)cf.kt3[c];
    const double Dh =
        mc * (((double)cf.rtjx[c + 1] * (double)a.u[c + 1] - (double)cf.rtjx[c] * (double)a.u[c]) /
                  g.dx +
              ((double)cf.rtjy[c + NX] * (double)a.v[c + NX] -
               (double)cf.rtjy[c] * (double)a.v[c]) /
                  g.dy);
    const double dzc = (double)m.dzeta_c[k + g.ng];
    const double wk = (double)a.w_src[c];
    const double wkp = (double)a.w_src[c + PL];
    const double tpi = a.tpi ? (double)a.tpi[c] : 0.0;
    P[k] = a.pip_src[c] + dtau * tpi -
           dtau * Kt * (Dh - (sv_hi - sv_lo) / dzc + c2 * (av_hi * wkp - av_lo * wk) / dzc);
    E[k] = dtau * Kt * c1 / dzc;
    sv_lo = sv_hi;
    av_lo = av_hi;
  }

  for (int k = 1; k < g.nz; ++k) {
    const int kk = k + g.ng;
    const int c = c0 + k * PL;
    const int cm = c - PL;
    const double thp_m = (double)a.thp_src[cm], thp_c = (double)a.thp_src[c];
    double thvw = (double)b.thbw[c] + 0.5 * (thp_m + thp_c);
    if (p.moisture) thvw *= 1.0 + kEps61 * 0.5 * ((double)a.qv[cm] + (double)a.qv[c]);
    const double dzw = (double)m.dzeta_w[kk];
    const double G = dtau * kCp * thvw * c1 / (J * dzw);
    const double A = -G * E[k - 1] * Av_at(k - 1);
    const double B = 1.0 + G * Av_at(k) * (E[k] + E[k - 1]);
    const double C = -G * E[k] * Av_at(k + 1);
    double buoy = kGrav * 0.5 * (thp_m + thp_c) / (double)b.thbw[c];
    if (p.moisture) {
      const double dqv = 0.5 * ((double)a.qv[cm] + (double)a.qv[c]) -
                         0.5 * ((double)b.qvb[cm] + (double)b.qvb[c]);
      const double qcr = 0.5 * ((double)a.qc[cm] + (double)a.qc[c]) +
                         0.5 * ((double)a.qr[cm] + (double)a.qr[c]);
      buoy += kGrav * (kEps61 * dqv - qcr);
    }
    const double RHS = (double)a.w_src[c] + dtau * (buoy + (double)a.tw[c]) -
                       dtau * kCp * thvw * c2 * (a.pip_src[c] - a.pip_src[cm]) / (J * dzw);
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
  // Back substitution; ds[] now holds W_k.
  double wabove = 0.0;
  for (int k = g.nz - 1; k >= 1; --k) {
    const double wk = ds[k] - cs[k] * wabove;
    ds[k] = wk;
    wabove = wk;
  }
  {
    const double us = 0.5 * ((double)a.u[c0] + (double)a.u[c0 + 1]);
    const double vs = 0.5 * ((double)a.v[c0] + (double)a.v[c0 + NX]);
    ds[0] = us * hxc + vs * hyc;
  }
  ds[g.nz] = 0.0;
  for (int k = 0; k < g.nz; ++k) {
    const int c = c0 + k * PL;
    const double oldpi = a.pip_src[c];
    a.pip_dst[c] = P[k] - E[k] * (Av_at(k + 1) * ds[k + 1] - Av_at(k) * ds[k]);
    a.piprev_dst[c] = oldpi;
    const double wc = 0.5 * (ds[k] + ds[k + 1]);
    a.thp_dst[c] = (T)((double)a.thp_src[c] +
                       dtau * ((double)a.tth[c] - wc * (double)b.dthbdz[c]));
    a.w_dst[c] = (T)ds[k];
  }
}

}  // namespace wfe::dev
