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
  CArr<T> u_src;
  CArr<T> v_src;
  Arr<T> u_dst;
  Arr<T> v_dst;
  CArr<double> pcur;   // pi' entering the substep
  CArr<double> pprev;  // pi' of the previous substep (time n at substep 0)
  CArr<double> pn;     // time-n pi' (ghost cells filled)
  CArr<T> thp_cur;
  CArr<T> thp_n;  // time-n theta' (ghost cells filled)
  CArr<T> qv;     // stage-constant qv (time n, ghosts filled)
  CArr<T> tu;
  CArr<T> tv;
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
    const T t = b.thb[p.hom ? g.id(0, 0, k) : c] + (in ? a.thp_cur[c] : a.thp_n[c]);
    const T q = a.qv[c];  // allocated in every mode; zero weight when dry
    return t * ((T)1 + (T)kEps61 * (p.moisture ? q : (T)0));
  }
  // u after the horizontal PGF step at face (i, j, k) (k_acou_uv :496-512).
  __device__ T u_step(int i, int j, int k) const {
    const int kk = k + g.ng;
    const int c = g.id(i, j, k);
    const int c2 = g.id2(i, j);
    const double ps0 = pstar(i, j, k);
    const T thbu = (T)0.5 * (thv(i - 1, j, k) + thv(i, j, k));
    const T mu = (T)0.5 * (m.mapf[c2 - 1] + m.mapf[c2]);
    double grad = (ps0 - pstar(i - 1, j, k)) * g.rdx;
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
    double grad = (ps0 - pstar(i, j - 1, k)) * g.rdy;
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

// LDS-tiled form of k_acou_uv: a BX x BY block of faces (256 threads, BX =
// 64 or 32) at one level stages pi* (at k and, over terrain, k +- 1) and
// theta_v once per cell (block + one-cell west/south halo), instead of
// recomputing them for every face that touches the cell. The radiation
// edges and the periodic wrap u(nx) = u(0) go through the direct UVDev
// path, which evaluates the same expressions.
// The tile of logical block bid (x: BX faces along x, y: BY rows, z: level).
template <class T, int BX>
__device__ __forceinline__ void uv_tile(const Grid& g, const Params& p, const Base<T>& b,
                                        const Metric<T>& m, const UVArgs<T>& a, BlockId bid) {
  constexpr int BY = 256 / BX;
  __shared__ double sP[3][BY + 1][BX + 1];  // pi* at k-1, k, k+1
  __shared__ T sT[BY + 1][BX + 1];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = bid.x * BX, j0 = bid.y * BY;
  const int k = bid.z;
  const UVDev<T> d{g, p, b, m, a};
  // This thread's faces: u at (i, j) for j < ny, v at (i, j) for i < nx.
  // Their own operands are loaded before the staging barrier so that both
  // sets of loads are in flight together.
  const int i = i0 + tx, j = j0 + ty;
  const bool live = i <= g.nx && j <= g.ny;
  const int c = g.id(live ? i : 0, live ? j : 0, k);
  const int c2 = g.id2(live ? i : 0, live ? j : 0);
  T usrc = 0, tus = 0, vsrc = 0, tvs = 0, mapc = 0, mapw = 0, maps = 0, hxu = 0, hyv = 0;
  if (live) {
    usrc = a.u_src[c];
    tus = a.tu[c];
    vsrc = a.v_src[c];
    tvs = a.tv[c];
    mapc = m.mapf[c2];
    mapw = m.mapf[c2 - 1];
    maps = m.mapf[c2 - g.NX];
    hxu = m.hxu[c2];
    hyv = m.hyv[c2];
  }
  // Stage cells i in [i0-1, i0+BX-1], j in [j0-1, j0+BY-1] that exist
  // (cells up to nx-1 / ny-1 plus the halo, which may be a ghost).
  for (int e = ty * BX + tx; e < (BX + 1) * (BY + 1); e += BX * BY) {
    const int r = e / (BX + 1), q = e % (BX + 1);
    const int ii = i0 - 1 + q, jj = j0 - 1 + r;
    if (ii <= g.nx - 1 && jj <= g.ny - 1) {
      sP[1][r][q] = d.pstar(ii, jj, k);
      if (p.terrain) {
        sP[0][r][q] = d.pstar(ii, jj, k - 1);
        sP[2][r][q] = d.pstar(ii, jj, k + 1);
      }
      sT[r][q] = d.thv(ii, jj, k);
    }
  }
  __syncthreads();
  if (!live) return;
  const int kk = k + g.ng;
  const T dtau = a.dtau;
  if (j < g.ny) {
    T un;
    if (i < g.nx && !(p.bcx && i == 0)) {
      const T thbu = (T)0.5 * (sT[ty + 1][tx] + sT[ty + 1][tx + 1]);
      const T mu = (T)0.5 * (mapw + mapc);
      double grad = (sP[1][ty + 1][tx + 1] - sP[1][ty + 1][tx]) * g.rdx;
      // terrain cross term -(z_x / J) dpi*/dzeta where the slope is non-zero
      if (hxu != (T)0) {
        const T zxu = hxu * ((T)1 - m.zeta_c[kk] / m.zt);
        const T dzc2 = m.zeta_c[kk + 1] - m.zeta_c[kk - 1];
        const T ju = (T)0.5 * (m.jac[c2 - 1] + m.jac[c2]);
        const double dpdz = 0.5 *
                            ((sP[2][ty + 1][tx] - sP[0][ty + 1][tx]) +
                             (sP[2][ty + 1][tx + 1] - sP[0][ty + 1][tx + 1])) /
                            (double)dzc2;
        grad -= (double)(zxu / ju) * dpdz;
      }
      un = usrc + dtau * (-(T)kCp * thbu * mu * (T)grad + tus);
    } else if (!p.bcx) {
      un = d.u_step(0, j, k);
    } else if (i == 0) {
      const T u0 = usrc;
      if (u0 >= (T)0) {
        un = b.ub[kk];
      } else {
        const T cb = u0 - (T)p.cstar;
        un = u0 - dtau * cb * (d.u_step(1, j, k) - u0) / (T)g.dx;
      }
    } else {
      const T u1 = usrc;
      if (u1 <= (T)0) {
        un = b.ub[kk];
      } else {
        const T cb = u1 + (T)p.cstar;
        un = u1 - dtau * cb * (u1 - d.u_step(g.nx - 1, j, k)) / (T)g.dx;
      }
    }
    a.u_dst[c] = un;
  }
  if (i < g.nx) {
    T vn;
    if (j < g.ny && !(p.bcy && j == 0)) {
      const T thbv = (T)0.5 * (sT[ty][tx + 1] + sT[ty + 1][tx + 1]);
      const T mv = (T)0.5 * (maps + mapc);
      double grad = (sP[1][ty + 1][tx + 1] - sP[1][ty][tx + 1]) * g.rdy;
      if (hyv != (T)0) {
        const T zyv = hyv * ((T)1 - m.zeta_c[kk] / m.zt);
        const T dzc2 = m.zeta_c[kk + 1] - m.zeta_c[kk - 1];
        const T jv = (T)0.5 * (m.jac[c2 - g.NX] + m.jac[c2]);
        const double dpdz = 0.5 *
                            ((sP[2][ty][tx + 1] - sP[0][ty][tx + 1]) +
                             (sP[2][ty + 1][tx + 1] - sP[0][ty + 1][tx + 1])) /
                            (double)dzc2;
        grad -= (double)(zyv / jv) * dpdz;
      }
      vn = vsrc + dtau * (-(T)kCp * thbv * mv * (T)grad + tvs);
    } else if (!p.bcy) {
      vn = d.v_step(i, 0, k);
    } else if (j == 0) {
      const T v0 = vsrc;
      if (v0 >= (T)0) {
        vn = b.vb[kk];
      } else {
        const T cb = v0 - (T)p.cstar;
        vn = v0 - dtau * cb * (d.v_step(i, 1, k) - v0) / (T)g.dy;
      }
    } else {
      const T v1 = vsrc;
      if (v1 <= (T)0) {
        vn = b.vb[kk];
      } else {
        const T cb = v1 + (T)p.cstar;
        vn = v1 - dtau * cb * (v1 - d.v_step(i, g.ny - 1, k)) / (T)g.dy;
      }
    }
    a.v_dst[c] = vn;
  }
}

template <class T, int BX = 64>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? 8 : 1) k_acou_uv_tile(Grid g, Params p, Base<T> b, Metric<T> m,
                                                      UVArgs<T> a) {
  uv_tile<T, BX>(g, p, b, m, a, xcd_block());
}

template <class T>
struct WPArgs {
  CArr<T> u;  // after this substep's horizontal step (faces 0..nx)
  CArr<T> v;
  CArr<T> w_src;
  CArr<double> pip_src;
  CArr<T> thp_src;
  Arr<T> w_dst;
  Arr<double> pip_dst;
  Arr<T> thp_dst;
  CArr<T> tw;   // w tendency including the moist buoyancy (see k_tendencies)
  CArr<T> tth;
  CArr<T> tpi;  // nullptr when the pi' slow tendency is zero
  // Davies rim weight (2-D) when tpi is set: the pi' tendency is the rim
  // relaxation term alone, zero where the weight is zero, so it is read only
  // in columns with a positive weight.
  CArr<T> tpw;
  CArr<T> qv;
  double dtau;
};

// Vertically implicit w-pi' solve, one thread per column (k_acou_wpi
// kernels.cu:531-645), in FP64 in every precision mode:
//   P_k = pi' + dtau T_pi - dtau K_t (D_h - d(S_v)/dzeta + c2 d(A w)/dzeta)
//   E_k = dtau K_t c1 / dzeta
//   -G E_{k-1} A_{k-1} w_{k-1} + (1 + G A_k (E_k + E_{k-1})) w_k
//       - G E_k A_{k+1} w_{k+1} = RHS_k - G (P_k - P_{k-1})
// with RHS_k = w + dtau (g theta'_w / thbar_w + T_w) - dtau c_p theta_v,w c2
// d(pi')/dzeta / J. The moist part of the reference's buoyancy,
// g (0.61 (q_v - qbar_v) - q_c - q_r) at the w level from the time-n moisture
// (kernels.cu:560-585), is constant over the substeps of a step and arrives
// already added to T_w (k_tendencies, or k_fold_moist_buoyancy on the
// phase-level path), so the substeps read only q_v (for theta_v).
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
    const double tpi = a.tpi && a.tpw[c2d] > (T)0 ? (double)a.tpi[c] : 0.0;
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
    const double buoy = kGrav * 0.5 * (thp_m + thp_c) / (double)b.thbw[c];
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
    a.pip_dst[c] = P[k] - E[k] * (Av_at(k + 1) * ds[k + 1] - Av_at(k) * ds[k]);
    const double wc = 0.5 * (ds[k] + ds[k + 1]);
    a.thp_dst[c] = (T)((double)a.thp_src[c] +
                       dtau * ((double)a.tth[c] - wc * (double)b.dthbdz[c]));
    a.w_dst[c] = (T)ds[k];
  }
}

// Rows read ahead of each serial chain (two chunk buffers of this many rows
// stay in registers).
inline constexpr int kTwistChunk = 4;
// LDS rows kept free before the first and after the last LDS array of the
// column kernels, for the read-ahead of the solve.
inline constexpr int kTwistPad = 2 * kTwistChunk;

// Tiled, k-parallel form of the same column solve. A workgroup owns TX
// consecutive interior columns in row-major order (column q = j nx + i, so a
// tile may continue on the next row and no lane idles at the end of a row)
// and NW thread rows; thread row r handles the levels k = r, r + NW, ... .
// Phases, separated by barriers:
//   1. every thread assembles, for its levels, P_k, E_k, Av_k, G_k, RHS_k
//      in registers (all loads of the column are issued here, in parallel
//      over k) and publishes P_k, E_k, Av_k in the D, C, B rows of LDS;
//   2. every thread forms the tridiagonal row k (A_k, B_k, C_k, D_k of
//      kernels.cu:589-628) from its registers and P_{k-1}, E_{k-1},
//      Av_{k-1} read from LDS, and writes it after a barrier, so the serial
//      part carries only the recurrence;
//   3. the twisted tridiagonal solve (twist_sweep_top / _bottom,
//      twist_row_coefs, twist_meet, twist_back): thread rows 0 and 1 (one
//      lane per column each) eliminate from the top and the bottom;
//   4. every thread writes w, pi' and theta' of its levels.
// LDS: 4 (nz + 1) + 2 kTwistPad rows of TX doubles. Levels per thread are
// bounded by KPT.
template <int TX, int NW, int KPT>
struct WpiTile {
  static constexpr int kThreads = TX * NW;
  static size_t lds_bytes(int nz) {
    return ((size_t)4 * (nz + 1) + 2 * kTwistPad) * TX * sizeof(double);
  }
};


// WPE: minimum waves per SIMD the register allocation must allow (1 = no
// constraint); WPE = 4 caps VGPRs at 128 so two 512-thread blocks share a CU
// and one block's serial elimination overlaps the other's loads.
// Tridiagonal solve of the rows k = 1 .. nz-1 of TX columns held in LDS
// (row k of lane lx at k * TX + lx; A_k x_{k-1} + B_k x_k + C_k x_{k+1} = D_k)
// by a twisted factorization: two lanes per column, in two wavefronts,
// eliminate at the same time from the top (rows 1 .. h-1) and from the
// bottom (rows n1 .. h+1, n1 = nz-1), meet at row h, and substitute back
// outwards, so each serial chain is half the column. No division sits on a
// serial chain:
//   1. sweeps over principal minors. Top: p_k = B_k p_{k-1} - A_k C_{k-1}
//      p_{k-2} (p_0 = 1, p_{-1} = 0) and e_k = D_k p_{k-1} - A_k e_{k-1}
//      (e_0 = 0), stored as p_{k-1}, p_k, e_k over A_k, B_k, D_k. Bottom:
//      q_k = B_k q_{k+1} - C_k A_{k+1} q_{k+2} (q_{n1+1} = 1, q_{n1+2} = 0)
//      and f_k = D_k q_{k+1} - C_k f_{k+1} (f_{n1+1} = 0), stored as
//      q_{k+1}, q_k, f_k over C_k, B_k, D_k. One dependent FMA per row and
//      chain; after every chunk of rows the running minors are rescaled by
//      the same power of two (exact), each row keeping its own scale;
//   2. (all threads, parallel over k) the eliminated rows. Top: x_k = d_k -
//      c_k x_{k+1} with c_k = C_k p_{k-1} / p_k, d_k = e_k / p_k (the
//      reference's Thomas coefficients, kernels.cu:589-628). Bottom: x_k =
//      g_k - a_k x_{k-1} with a_k = A_k q_{k+1} / q_k, g_k = f_k / q_k;
//   3. x_h = (D_h - A_h d_{h-1} - C_h g_{h+1}) / (B_h - A_h c_{h-1} - C_h
//      a_{h+1}) in both lanes, then (after a barrier) back substitution
//      upwards (top lane) and downwards (bottom lane). sD ends as W_k,
//      W_0 = w0, W_nz = 0.
// The phases are separated by workgroup barriers in the caller. Rows are
// read in chunks of kTwistChunk, one chunk ahead of the recurrence, at fixed
// offsets: the reads run up to 2 kTwistChunk - 2 rows past either end of a
// column, which the caller's LDS layout covers (kTwistPad rows before the
// first row array and after the last LDS array; a row read past the end of
// one array lands in the next one), and such values are never used. Whole
// chunks run without per-row tests; the running minors are rescaled after
// each of them.
__host__ __device__ constexpr int twist_row(int nz) { return nz / 2; }  // h, 1 <= h <= n1

template <int TX>
__device__ __forceinline__ void twist_sweep_top(double* sA, double* sB, const double* sC,
                                                double* sD, int lx, int nz) {
  constexpr int U = kTwistChunk;
  const int last = twist_row(nz) - 1;  // rows 1 .. h-1
  auto at = [&](int k) { return k * TX + lx; };
  double A[U], B[U], Cm[U], D[U], nA[U], nB[U], nCm[U], nD[U];  // Cm = C_{k-1}
#pragma unroll
  for (int u = 0; u < U; ++u) {
    A[u] = sA[at(1 + u)];
    B[u] = sB[at(1 + u)];
    Cm[u] = u == 0 ? 0.0 : sC[at(u)];
    D[u] = sD[at(1 + u)];
  }
  double p1 = 1.0, p2 = 0.0, e1 = 0.0;  // p_{k-1}, p_{k-2}, e_{k-1}
  auto row = [&](int k, int u) {
    const double pk = fma(B[u], p1, -(A[u] * Cm[u]) * p2);
    const double ek = fma(-A[u], e1, D[u] * p1);
    sA[at(k)] = p1;
    sB[at(k)] = pk;
    sD[at(k)] = ek;
    p2 = p1;
    p1 = pk;
    e1 = ek;
  };
  int k0 = 1;
  for (; k0 + U - 1 <= last; k0 += U) {  // whole chunks
#pragma unroll
    for (int u = 0; u < U; ++u) {
      nA[u] = sA[at(k0 + U + u)];
      nB[u] = sB[at(k0 + U + u)];
      nCm[u] = sC[at(k0 + U + u - 1)];
      nD[u] = sD[at(k0 + U + u)];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) row(k0 + u, u);
    const int ex = __builtin_amdgcn_frexp_exp(p1);
    p1 = ldexp(p1, -ex);
    p2 = ldexp(p2, -ex);
    e1 = ldexp(e1, -ex);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      A[u] = nA[u];
      B[u] = nB[u];
      Cm[u] = nCm[u];
      D[u] = nD[u];
    }
  }
#pragma unroll
  for (int u = 0; u < U - 1; ++u)  // the remaining rows of a partial chunk
    if (k0 + u <= last) row(k0 + u, u);
}

template <int TX>
__device__ __forceinline__ void twist_sweep_bottom(const double* sA, double* sB, double* sC,
                                                   double* sD, int lx, int nz) {
  constexpr int U = kTwistChunk;
  const int n1 = nz - 1, first = twist_row(nz) + 1;  // rows n1 .. h+1
  auto at = [&](int k) { return k * TX + lx; };
  double Ap[U], B[U], C[U], D[U], nAp[U], nB[U], nC[U], nD[U];  // Ap = A_{k+1}
#pragma unroll
  for (int u = 0; u < U; ++u) {
    Ap[u] = u == 0 ? 0.0 : sA[at(n1 - u + 1)];
    B[u] = sB[at(n1 - u)];
    C[u] = sC[at(n1 - u)];
    D[u] = sD[at(n1 - u)];
  }
  double q1 = 1.0, q2 = 0.0, f1 = 0.0;  // q_{k+1}, q_{k+2}, f_{k+1}
  auto row = [&](int k, int u) {
    const double qk = fma(B[u], q1, -(C[u] * Ap[u]) * q2);
    const double fk = fma(-C[u], f1, D[u] * q1);
    sC[at(k)] = q1;
    sB[at(k)] = qk;
    sD[at(k)] = fk;
    q2 = q1;
    q1 = qk;
    f1 = fk;
  };
  int k0 = n1;
  for (; k0 - U + 1 >= first; k0 -= U) {  // whole chunks
#pragma unroll
    for (int u = 0; u < U; ++u) {
      nAp[u] = sA[at(k0 - U - u + 1)];
      nB[u] = sB[at(k0 - U - u)];
      nC[u] = sC[at(k0 - U - u)];
      nD[u] = sD[at(k0 - U - u)];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) row(k0 - u, u);
    const int ex = __builtin_amdgcn_frexp_exp(q1);
    q1 = ldexp(q1, -ex);
    q2 = ldexp(q2, -ex);
    f1 = ldexp(f1, -ex);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      Ap[u] = nAp[u];
      B[u] = nB[u];
      C[u] = nC[u];
      D[u] = nD[u];
    }
  }
#pragma unroll
  for (int u = 0; u < U - 1; ++u)
    if (k0 - u >= first) row(k0 - u, u);
}

// Phase 2 for row k (any thread): the eliminated row from its minors.
__device__ __forceinline__ void twist_row_coefs(double* sA, const double* sB, double* sC,
                                                double* sD, int at, int k, int nz) {
  const int h = twist_row(nz);
  if (k < h) {  // c_k over C_k, d_k over D_k
    const double r = rcp_nr(sB[at]);
    sC[at] = sC[at] * sA[at] * r;
    sD[at] = sD[at] * r;
  } else if (k > h) {  // a_k over A_k, g_k over D_k
    const double r = rcp_nr(sB[at]);
    sA[at] = sA[at] * sC[at] * r;
    sD[at] = sD[at] * r;
  }
}

// Phase 3a: the meeting row's unknown x_h (read-only; the caller holds a
// barrier between this and the writes of phase 3b).
template <int TX>
__device__ __forceinline__ double twist_meet(const double* sA, const double* sB,
                                             const double* sC, const double* sD, int lx,
                                             int nz) {
  const int n1 = nz - 1, h = twist_row(nz);
  auto at = [&](int k) { return k * TX + lx; };
  double den = sB[at(h)], num = sD[at(h)];
  if (h - 1 >= 1) {
    den -= sA[at(h)] * sC[at(h - 1)];
    num -= sA[at(h)] * sD[at(h - 1)];
  }
  if (h + 1 <= n1) {
    den -= sC[at(h)] * sA[at(h + 1)];
    num -= sC[at(h)] * sD[at(h + 1)];
  }
  return num / den;
}

// Phase 3b: back substitution from x_h upwards (top) or downwards (bottom).
template <int TX>
__device__ __forceinline__ void twist_back(const double* sA, const double* sC, double* sD,
                                           int lx, int nz, bool top, double xh, double w0) {
  constexpr int U = kTwistChunk;
  const int n1 = nz - 1, h = twist_row(nz);
  auto at = [&](int k) { return k * TX + lx; };
  double x = xh;
  if (top) {  // x_k = d_k - c_k x_{k+1}, k = h-1 .. 1
    double C[U], D[U], nC[U], nD[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      C[u] = sC[at(h - 1 - u)];
      D[u] = sD[at(h - 1 - u)];
    }
    int k0 = h - 1;
    for (; k0 - U + 1 >= 1; k0 -= U) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        nC[u] = sC[at(k0 - U - u)];
        nD[u] = sD[at(k0 - U - u)];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        x = D[u] - C[u] * x;
        sD[at(k0 - u)] = x;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        C[u] = nC[u];
        D[u] = nD[u];
      }
    }
#pragma unroll
    for (int u = 0; u < U - 1; ++u)
      if (k0 - u >= 1) {
        x = D[u] - C[u] * x;
        sD[at(k0 - u)] = x;
      }
    sD[at(h)] = xh;
    sD[lx] = w0;
  } else {  // x_k = g_k - a_k x_{k-1}, k = h+1 .. n1
    double Av[U], D[U], nAv[U], nD[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      Av[u] = sA[at(h + 1 + u)];
      D[u] = sD[at(h + 1 + u)];
    }
    int k0 = h + 1;
    for (; k0 + U - 1 <= n1; k0 += U) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        nAv[u] = sA[at(k0 + U + u)];
        nD[u] = sD[at(k0 + U + u)];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        x = D[u] - Av[u] * x;
        sD[at(k0 + u)] = x;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        Av[u] = nAv[u];
        D[u] = nD[u];
      }
    }
#pragma unroll
    for (int u = 0; u < U - 1; ++u)
      if (k0 + u <= n1) {
        x = D[u] - Av[u] * x;
        sD[at(k0 + u)] = x;
      }
    sD[at(nz)] = 0.0;
  }
}

// Phases 1-4 of the split column kernel for the column (i, j) of lane lx
// (col: the lane owns a column; otherwise i, j name any valid column and
// nothing is stored), with u and v of the column read through `uv`:
// u0(k), u1(k) the west and east faces, v0(k), v1(k) the south and north
// faces at level k. `lds` holds the rows as k_acou_wpi_tile lays them out.
template <class T, int TX, int NW, int KPT, bool MOIST, bool TPI, class UV>
__device__ __forceinline__ void wpi_column(const Grid& g, const Params& p, const Base<T>& b,
                                           const Metric<T>& m, const Coef<T>& cf,
                                           const WPArgs<T>& a, double* lds, int lx, int ry,
                                           bool col, int i, int j, const UV& uv) {
  const int L = g.nz + 1;
  double* sD = lds + kTwistPad * TX;  // P_k, then D_k -> ds_k -> W_k (k in [0, nz])
  double* sA = sD + L * TX;  // A_k            k in [1, nz)
  double* sB = sA + L * TX;  // B_k            k in [1, nz)
  double* sC = sB + L * TX;  // C_k -> cs_k    k in [1, nz)
  const double c1 = (1.0 + p.beta) * 0.5;
  const double c2 = (1.0 - p.beta) * 0.5;
  const double dtau = a.dtau;
  const int NX = g.NX, PL = g.plane;
  const int c2d = g.id2(col ? i : 0, j);
  const double J = (double)m.jac[c2d];
  const double hxc = (double)m.hx[c2d];
  const double hyc = (double)m.hy[c2d];
  const double mc = (double)m.mapf[c2d];
  const double rJ = rcp_nr(J);
  // Over flat ground the cross-metric flux S_v is exactly zero.
  const bool flat = hxc == 0.0 && hyc == 0.0;
  const bool rim = TPI && a.tpw[c2d] > (T)0;  // columns with a pi' tendency
  const int c0 = g.id(col ? i : 0, j, 0);
  // Base-state and coefficient column: column (0, 0) when every column holds
  // the same profile (one cache line per level for the whole grid).
  const int cb = p.hom ? g.id(0, 0, 0) : c0;
  const int ox = p.hom ? 0 : 1, oy = p.hom ? 0 : NX;

  // Values kept in registers across the phases (storage-type values stay in
  // T; their FP64 conversion is exact).
  double Pk[KPT], Ek[KPT], Gk[KPT], Rk[KPT];
  T Avk[KPT], Avk1[KPT];
  double w0 = 0.0;

  auto sv = [&](int k) -> double {  // cross-metric vertical flux at w level k
    if (k <= 0 || k >= g.nz || flat) return 0.0;
    const int c = c0 + k * PL;
    const double fac = 1.0 - (double)m.zeta_w[k + g.ng] * m.rzt;
    const double uw = 0.25 * ((double)uv.u0(k - 1) + (double)uv.u1(k - 1) + (double)uv.u0(k) +
                              (double)uv.u1(k));
    const double vw = 0.25 * ((double)uv.v0(k - 1) + (double)uv.v1(k - 1) + (double)uv.v0(k) +
                              (double)uv.v1(k));
    return (double)cf.aw[cb + k * PL] * (uw * hxc + vw * hyc) * fac;
  };

  // ---- phase 1: assembly, parallel over k. Every load of a level is
  // issued unconditionally (levels clamped into [0, nz), the padded layout
  // holds k - 1 and k + 1), so the loads of all KPT levels are in flight
  // together; the level guards only select results and stores.
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    const bool act = col && k < g.nz;
    const int kc = k < g.nz ? k : g.nz - 1;
    const int c = c0 + kc * PL;
    const int cm = c - PL;
    const int kk = kc + g.ng;
    const int cc = cb + kc * PL;  // base-state / coefficient index
    const double Kt = (double)cf.kt3[cc];
    const double Dh = mc * (((double)cf.rtjx[cc + ox] * (double)uv.u1(kc) -
                             (double)cf.rtjx[cc] * (double)uv.u0(kc)) *
                                g.rdx +
                            ((double)cf.rtjy[cc + oy] * (double)uv.v1(kc) -
                             (double)cf.rtjy[cc] * (double)uv.v0(kc)) *
                                g.rdy);
    const double rdzc = m.rdzc[kk];
    const T awk = kc >= 1 ? cf.aw[cc] : (T)0;
    const T awk1 = kc + 1 < g.nz ? cf.aw[cc + PL] : (T)0;
    const double avk = (double)awk, avk1 = (double)awk1;
    const double wk = (double)a.w_src[c];
    const double wkp = (double)a.w_src[c + PL];
    const double tpi = rim ? (double)a.tpi[c] : 0.0;
    const double pk = a.pip_src[c];
    const double pkm = a.pip_src[cm];
    const T thkT = a.thp_src[c];
    const double thk = (double)thkT;
    const double thm = (double)a.thp_src[cm];
    const double thbw = (double)b.thbw[cc];
    const double twk = (double)a.tw[c];
    double mfac = 1.0;
    if constexpr (MOIST) mfac = 1.0 + kEps61 * 0.5 * ((double)a.qv[cm] + (double)a.qv[c]);
    Pk[t] = pk + dtau * tpi -
            dtau * Kt *
                (Dh - (sv(kc + 1) - sv(kc)) * rdzc + c2 * (avk1 * wkp - avk * wk) * rdzc);
    Ek[t] = dtau * Kt * c1 * rdzc;
    Avk[t] = awk;
    Avk1[t] = awk1;
    // w level k >= 1: G_k and the right-hand side (kernels.cu:560-585)
    const double thvw = (thbw + 0.5 * (thm + thk)) * mfac;
    const double rJdzw = rJ * m.rdzw[kk];
    const double buoy = kGrav * 0.5 * (thm + thk) * rcp_nr(thbw);
    Gk[t] = kc >= 1 ? dtau * kCp * thvw * c1 * rJdzw : 0.0;
    Rk[t] = kc >= 1 ? wk + dtau * (buoy + twk) - dtau * kCp * thvw * c2 * (pk - pkm) * rJdzw
                    : 0.0;
    if (k == 0) {  // diagnostic w(0) = u h_x + v h_y at the surface
      const double us = 0.5 * ((double)uv.u0(0) + (double)uv.u1(0));
      const double vs = 0.5 * ((double)uv.v0(0) + (double)uv.v1(0));
      w0 = us * hxc + vs * hyc;
    }
    if (act) {
      sD[k * TX + lx] = Pk[t];
      sC[k * TX + lx] = Ek[t];
      sB[k * TX + lx] = (double)Avk[t];
    }
  }
  __syncthreads();

  // ---- phase 2: tridiagonal rows (kernels.cu:596-603), parallel over k,
  // with E, A_v and P of the level below read from LDS
  double Ra[KPT], Rb[KPT], Rc[KPT], Dk[KPT];
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    Ra[t] = Rb[t] = Rc[t] = Dk[t] = 0.0;
    if (col && k >= 1 && k < g.nz) {
      const int bl = (k - 1) * TX + lx;
      const double e_prev = sC[bl];
      const double av_prev = sB[bl];
      const double G = Gk[t];
      Ra[t] = -G * e_prev * av_prev;
      Rb[t] = 1.0 + G * (double)Avk[t] * (Ek[t] + e_prev);
      Rc[t] = -G * Ek[t] * (double)Avk1[t];
      Dk[t] = Rk[t] - G * (Pk[t] - sD[bl]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    if (col && k >= 1 && k < g.nz) {
      sA[k * TX + lx] = Ra[t];
      sB[k * TX + lx] = Rb[t];
      sC[k * TX + lx] = Rc[t];
      sD[k * TX + lx] = Dk[t];
    }
  }
  __syncthreads();

  // ---- phase 3: tridiagonal solve, twisted at row nz/2 (thread rows 0 and
  // 1 take the top and the bottom of each column)
  if (ry == 0 && col) twist_sweep_top<TX>(sA, sB, sC, sD, lx, g.nz);
  if (ry == 1 && col) twist_sweep_bottom<TX>(sA, sB, sC, sD, lx, g.nz);
  __syncthreads();
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    if (col && k >= 1 && k < g.nz) twist_row_coefs(sA, sB, sC, sD, k * TX + lx, k, g.nz);
  }
  __syncthreads();
  double xh = 0.0;
  if (ry <= 1 && col) xh = twist_meet<TX>(sA, sB, sC, sD, lx, g.nz);
  __syncthreads();
  if (ry <= 1 && col) twist_back<TX>(sA, sC, sD, lx, g.nz, ry == 0, xh, w0);
  __syncthreads();

  // ---- phase 4: pi', theta', w (theta', T_theta and dtheta/dz re-read here
  // rather than held in registers across the solve)
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    if (col && k < g.nz) {
      const int c = c0 + k * PL;
      const double Wk = sD[k * TX + lx], Wk1 = sD[(k + 1) * TX + lx];
      a.pip_dst[c] = Pk[t] - Ek[t] * ((double)Avk1[t] * Wk1 - (double)Avk[t] * Wk);
      const double wc = 0.5 * (Wk + Wk1);
      a.thp_dst[c] = (T)((double)a.thp_src[c] +
                         dtau * ((double)a.tth[c] - wc * (double)b.dthbdz[cb + k * PL]));
      a.w_dst[c] = (T)Wk;
    }
  }
}

// u and v of one column read from the arrays (the split kernel).
template <class T>
struct UVFromArrays {
  CArr<T> u, v;
  int c0, PL, NX;
  __device__ __forceinline__ T u0(int k) const { return u[c0 + k * PL]; }
  __device__ __forceinline__ T u1(int k) const { return u[c0 + k * PL + 1]; }
  __device__ __forceinline__ T v0(int k) const { return v[c0 + k * PL]; }
  __device__ __forceinline__ T v1(int k) const { return v[c0 + k * PL + NX]; }
};

template <class T, int TX, int NW, int KPT, bool MOIST, bool TPI, int WPE = 1>
__global__ void __launch_bounds__(TX * NW, WPE) k_acou_wpi_tile(Grid g, Params p, Base<T> b,
                                                                Metric<T> m, Coef<T> cf,
                                                                WPArgs<T> a) {
  extern __shared__ double lds[];
  const int lx = threadIdx.x, ry = threadIdx.y;
  const BlockId bid = xcd_block();
  const int q = bid.x * TX + lx;  // interior column j nx + i
  const bool col = q < g.nx * g.ny;
  const int j = col ? q / g.nx : 0;
  const int i = col ? q - j * g.nx : 0;
  const UVFromArrays<T> uv{a.u, a.v, g.id(i, j, 0), g.plane, g.NX};
  wpi_column<T, TX, NW, KPT, MOIST, TPI>(g, p, b, m, cf, a, lds, lx, ry, col, i, j, uv);
}

// ---------------------------------------------------------------------------
// Member-batched substep pair: one launch of k_acou_uv_batch and one of
// k_acou_wpi_batch advance the same substep of every member of an ensemble,
// whatever the members' grid sizes. Each member's arguments and the range
// of logical blocks it owns sit in a device table, in ascending ranges; a
// workgroup finds its member from its logical block id (after the XCD remap
// over the whole launch) and runs the single-member tile on that member's
// arrays, so every cell sees exactly the arithmetic of k_acou_uv_tile and
// k_acou_wpi_tile.
template <class T>
struct BatchMember {
  Grid g;
  Params p;
  Base<T> b;
  Metric<T> m;
  Coef<T> cf;
  UVArgs<T> uv;
  WPArgs<T> wp;
  int uv_begin, uv_gx, uv_gz;  // u/v tiles: [uv_begin, uv_begin + gx * gy * gz)
  int wp_begin;                // column tiles: [wp_begin, next member's wp_begin)
};

// The member whose range holds logical block l: the last one starting at or
// before l (binary search over the ascending starts).
template <class T, bool UV>
__device__ __forceinline__ int batch_member(const BatchMember<T>* __restrict__ tab, int n, int l) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int start = UV ? tab[mid].uv_begin : tab[mid].wp_begin;
    if (start <= l)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <class T, int BX>
__global__ void __launch_bounds__(256) k_acou_uv_batch(const BatchMember<T>* __restrict__ tab,
                                                       int n) {
  const int l = xcd_remap(blockIdx.x, gridDim.x);
  const BatchMember<T>& M = tab[batch_member<T, true>(tab, n, l)];
  const int r = l - M.uv_begin;
  const int t = r / M.uv_gx;
  uv_tile<T, BX>(M.g, M.p, M.b, M.m, M.uv, BlockId{r - t * M.uv_gx, t / M.uv_gz, t % M.uv_gz});
}

template <class T, int TX, int NW, int KPT, bool MOIST, bool TPI>
__global__ void __launch_bounds__(TX * NW) k_acou_wpi_batch(const BatchMember<T>* __restrict__ tab,
                                                            int n) {
  extern __shared__ double lds[];
  const int l = xcd_remap(blockIdx.x, gridDim.x);
  const BatchMember<T>& M = tab[batch_member<T, false>(tab, n, l)];
  const Grid& g = M.g;
  const int lx = threadIdx.x, ry = threadIdx.y;
  const int q = (l - M.wp_begin) * TX + lx;  // interior column j nx + i
  const bool col = q < g.nx * g.ny;
  const int j = col ? q / g.nx : 0;
  const int i = col ? q - j * g.nx : 0;
  const UVFromArrays<T> uv{M.wp.u, M.wp.v, g.id(i, j, 0), g.plane, g.NX};
  wpi_column<T, TX, NW, KPT, MOIST, TPI>(g, M.p, M.b, M.m, M.cf, M.wp, lds, lx, ry, col, i, j, uv);
}

// ---------------------------------------------------------------------------
// Fused acoustic substep on 2-D tiles of columns, for grids whose columns
// differ (terrain). A workgroup owns CX x CY columns (CX CY = 64, one lane
// each) and all their levels (NW thread rows, KPT levels each):
//   0a. pi* (levels -1 .. nz) and theta_v (levels 0 .. nz-1) of the cells
//       i0-1 .. i0+CX, j0-1 .. j0+CY that lie below i = nx, j = ny (the
//       west/south halo may be a ghost cell) are staged in LDS once;
//   0b. u on the faces i0 .. i0+CX of rows j0 .. j0+CY-1 and v on the faces
//       j0 .. j0+CY of columns i0 .. i0+CX-1, at every level, from the
//       staged values with the expressions of k_acou_uv_tile (radiation
//       edges and the periodic wraps through the direct path), kept in LDS
//       and written out for the faces the workgroup owns (the west and
//       south faces of its columns, and faces nx and ny);
//   1-4. the column phases of the split kernel (wpi_column) with u and v
//       from LDS; the rows of the solve reuse the LDS of step 0a.
// Against the split pair this saves the second read of u, v, theta', qv and
// pi' and one launch per substep; the east and north halo faces are formed
// by both workgroups that touch them.
template <int CX, int CY>
struct Tile2D {
  static constexpr int TX = 64, PX = CX + 2, PY = CY + 2, NP = PX * PY;
  // doubles of the region shared by the staging (0a) and the solve rows
  template <class T>
  __host__ __device__ static size_t region_a(int nz) {
    const size_t rows = (size_t)(4 * (nz + 1) + 2 * kTwistPad) * TX;
    const size_t stage = (size_t)NP * (nz + 2) + ((size_t)NP * nz * sizeof(T) + 7) / 8;
    return rows > stage ? rows : stage;
  }
  template <class T>
  static size_t lds_bytes(int nz) {
    return region_a<T>(nz) * sizeof(double) +
           ((size_t)nz * CY * (CX + 1) + (size_t)nz * (CY + 1) * CX) * sizeof(T);
  }
};

// u, v of one column from the LDS face tiles of k_acou_tile2d.
template <class T, int CX, int CY>
struct UVFromTile {
  const T* sU;  // [k][cy][0 .. CX]
  const T* sV;  // [k][0 .. CY][cx]
  int cx, cy;
  __device__ __forceinline__ T u0(int k) const { return sU[(k * CY + cy) * (CX + 1) + cx]; }
  __device__ __forceinline__ T u1(int k) const { return sU[(k * CY + cy) * (CX + 1) + cx + 1]; }
  __device__ __forceinline__ T v0(int k) const { return sV[(k * (CY + 1) + cy) * CX + cx]; }
  __device__ __forceinline__ T v1(int k) const { return sV[(k * (CY + 1) + cy + 1) * CX + cx]; }
};

template <class T, int CX, int CY, int NW, int KPT, bool MOIST, bool TPI>
__global__ void __launch_bounds__(64 * NW) k_acou_tile2d(Grid g, Params p, Base<T> b,
                                                          Metric<T> m, Coef<T> cf, UVArgs<T> a,
                                                          WPArgs<T> w) {
  static_assert(CX * CY == 64, "one lane per column of the tile");
  using Sh = Tile2D<CX, CY>;
  constexpr int PX = Sh::PX, NP = Sh::NP, NT = 64 * NW;
  extern __shared__ double lds[];
  const int nz = g.nz;
  double* sPs = lds;                                        // pi* [k + 1][py][px]
  T* sTv = reinterpret_cast<T*>(lds + (size_t)NP * (nz + 2));  // theta_v [k][py][px]
  T* sU = reinterpret_cast<T*>(lds + Sh::template region_a<T>(nz));
  T* sV = sU + (size_t)nz * CY * (CX + 1);
  const int tid = threadIdx.y * 64 + threadIdx.x;
  const BlockId bid = xcd_block();
  const int i0 = bid.x * CX, j0 = bid.y * CY;
  const UVDev<T> d{g, p, b, m, a};

  // ---- 0a: pi* and theta_v of the tile's cells and their west/south halo
  for (int e = tid; e < NP * (nz + 2); e += NT) {
    const int lev = e / NP, r2 = e - lev * NP;
    const int py = r2 / PX, px = r2 - py * PX;
    const int ii = i0 - 1 + px, jj = j0 - 1 + py, k = lev - 1;
    if (ii <= g.nx - 1 && jj <= g.ny - 1) {
      sPs[e] = d.pstar(ii, jj, k);
      if (k >= 0 && k < nz) sTv[k * NP + r2] = d.thv(ii, jj, k);
    }
  }
  __syncthreads();
  auto P = [&](int ii, int jj, int k) { return sPs[((k + 1) * Sh::PY + (jj - j0 + 1)) * PX + (ii - i0 + 1)]; };
  auto TV = [&](int ii, int jj, int k) { return sTv[(k * Sh::PY + (jj - j0 + 1)) * PX + (ii - i0 + 1)]; };
  const T dtau = a.dtau;

  // ---- 0b: u faces (CX + 1 per row, CY rows) and v faces (CX per row,
  // CY + 1 rows) at every level
  for (int e = tid; e < nz * CY * (CX + 1); e += NT) {
    const int k = e / (CY * (CX + 1)), r2 = e - k * (CY * (CX + 1));
    const int fy = r2 / (CX + 1), fx = r2 - fy * (CX + 1);
    const int i = i0 + fx, j = j0 + fy;
    if (i > g.nx || j >= g.ny) continue;
    const int kk = k + g.ng;
    const int c = g.id(i, j, k);
    const int c2 = g.id2(i, j);
    const T usrc = a.u_src[c];
    T un;
    if (i < g.nx && !(p.bcx && i == 0)) {
      const T thbu = (T)0.5 * (TV(i - 1, j, k) + TV(i, j, k));
      const T mu = (T)0.5 * (m.mapf[c2 - 1] + m.mapf[c2]);
      double grad = (P(i, j, k) - P(i - 1, j, k)) * g.rdx;
      const T hxu = m.hxu[c2];
      if (hxu != (T)0) {
        const T zxu = hxu * ((T)1 - m.zeta_c[kk] / m.zt);
        const T dzc2 = m.zeta_c[kk + 1] - m.zeta_c[kk - 1];
        const T ju = (T)0.5 * (m.jac[c2 - 1] + m.jac[c2]);
        const double dpdz = 0.5 *
                            ((P(i - 1, j, k + 1) - P(i - 1, j, k - 1)) +
                             (P(i, j, k + 1) - P(i, j, k - 1))) /
                            (double)dzc2;
        grad -= (double)(zxu / ju) * dpdz;
      }
      un = usrc + dtau * (-(T)kCp * thbu * mu * (T)grad + a.tu[c]);
    } else if (!p.bcx) {
      un = d.u_step(0, j, k);
    } else if (i == 0) {
      if (usrc >= (T)0) {
        un = b.ub[kk];
      } else {
        const T cb = usrc - (T)p.cstar;
        un = usrc - dtau * cb * (d.u_step(1, j, k) - usrc) / (T)g.dx;
      }
    } else {
      if (usrc <= (T)0) {
        un = b.ub[kk];
      } else {
        const T cb = usrc + (T)p.cstar;
        un = usrc - dtau * cb * (usrc - d.u_step(g.nx - 1, j, k)) / (T)g.dx;
      }
    }
    sU[e] = un;
    if (fx < CX || i == g.nx) a.u_dst[c] = un;
  }
  for (int e = tid; e < nz * (CY + 1) * CX; e += NT) {
    const int k = e / ((CY + 1) * CX), r2 = e - k * ((CY + 1) * CX);
    const int fy = r2 / CX, fx = r2 - fy * CX;
    const int i = i0 + fx, j = j0 + fy;
    if (i >= g.nx || j > g.ny) continue;
    const int kk = k + g.ng;
    const int c = g.id(i, j, k);
    const int c2 = g.id2(i, j);
    const T vsrc = a.v_src[c];
    T vn;
    if (j < g.ny && !(p.bcy && j == 0)) {
      const T thbv = (T)0.5 * (TV(i, j - 1, k) + TV(i, j, k));
      const T mv = (T)0.5 * (m.mapf[c2 - g.NX] + m.mapf[c2]);
      double grad = (P(i, j, k) - P(i, j - 1, k)) * g.rdy;
      const T hyv = m.hyv[c2];
      if (hyv != (T)0) {
        const T zyv = hyv * ((T)1 - m.zeta_c[kk] / m.zt);
        const T dzc2 = m.zeta_c[kk + 1] - m.zeta_c[kk - 1];
        const T jv = (T)0.5 * (m.jac[c2 - g.NX] + m.jac[c2]);
        const double dpdz = 0.5 *
                            ((P(i, j - 1, k + 1) - P(i, j - 1, k - 1)) +
                             (P(i, j, k + 1) - P(i, j, k - 1))) /
                            (double)dzc2;
        grad -= (double)(zyv / jv) * dpdz;
      }
      vn = vsrc + dtau * (-(T)kCp * thbv * mv * (T)grad + a.tv[c]);
    } else if (!p.bcy) {
      vn = d.v_step(i, 0, k);
    } else if (j == 0) {
      if (vsrc >= (T)0) {
        vn = b.vb[kk];
      } else {
        const T cb = vsrc - (T)p.cstar;
        vn = vsrc - dtau * cb * (d.v_step(i, 1, k) - vsrc) / (T)g.dy;
      }
    } else {
      if (vsrc <= (T)0) {
        vn = b.vb[kk];
      } else {
        const T cb = vsrc + (T)p.cstar;
        vn = vsrc - dtau * cb * (vsrc - d.v_step(i, g.ny - 1, k)) / (T)g.dy;
      }
    }
    sV[e] = vn;
    if (fy < CY || j == g.ny) a.v_dst[c] = vn;
  }
  __syncthreads();

  // ---- 1-4: the column solve, u and v from the face tiles
  const int lx = threadIdx.x, ry = threadIdx.y;
  const int cx = lx % CX, cy = lx / CX;
  const int i = i0 + cx, j = j0 + cy;
  const bool col = i < g.nx && j < g.ny;
  const UVFromTile<T, CX, CY> uv{sU, sV, cx, cy};
  wpi_column<T, 64, NW, KPT, MOIST, TPI>(g, p, b, m, cf, w, lds, lx, ry, col, col ? i : 0,
                                         col ? j : 0, uv);
}

// ---------------------------------------------------------------------------
// Fused acoustic substep: the horizontal u, v step and the w-pi' column solve
// of one substep in a single launch, on rows of columns. A workgroup is TX
// lanes along x by NW thread rows (levels k = r, r + NW, ...) at one j. Lane
// l holds cell i = i0 - 1 + l: lanes 1 .. TX-2 own the columns of the block
// (TX - 2 of them), lane 0 is the west halo cell and lane TX-1 the cell whose
// west face is the east face of the last column.
//   0. every lane forms u at its west face and v at the faces j and j+1 of
//      its cell (radiation edges and the periodic wrap as k_acou_uv_tile);
//      u at the east face is the next lane's value (a lane shuffle), v(j+1)
//      is recomputed here rather than read from the block at j+1. The owned
//      faces are written out;
//   1. assembly of P, E, G and the right-hand side per level in FP64, as
//      k_acou_wpi_tile (u, v from registers; over terrain the vertical
//      cross-metric flux takes u, v at k-1 from LDS);
//   2. rows A..D of the tridiagonal system in LDS;
//   3. Thomas elimination, one lane per column, FP64;
//   4. w, pi', theta' written.
// Neighbour reads of pi' and theta' require that their outputs do not alias
// their inputs: pi' rotates over three buffers (the previous pi' is the
// input of the substep before, so pi'prev is never written) and theta'
// ping-pongs; u, v ping-pong; w is only read and written in its own column.
template <class T>
struct FusedArgs {
  UVArgs<T> uv;  // u/v sources and destinations, pi' cur/prev/n, theta' cur/n, qv, tu, tv
  CArr<T> w_src;
  Arr<T> w_dst;
  Arr<double> pip_dst;
  Arr<T> thp_dst;
  CArr<T> tw;
  CArr<T> tth;
  CArr<T> tpi;
  CArr<T> tpw;  // rim weight gating the tpi reads (as WPArgs::tpw)
};

template <int TX, int NW, int KPT>
struct RowShape {
  static constexpr int NC = TX - 2;  // columns per workgroup
  // Six FP64 arrays of (nz + 1) levels (rows A..D, E, A_v), plus u, v(j),
  // v(j+1) per level over terrain.
  static size_t lds_bytes(int nz, bool terrain, size_t tsize) {
    return ((size_t)6 * (nz + 1) + 2 * kTwistPad) * TX * sizeof(double) +
           (terrain ? (size_t)3 * nz * TX * tsize : 0);
  }
};

// WPE: minimum waves per SIMD the register allocation must allow (1 = no
// cap; 4 with 512-thread blocks caps VGPRs at 128 so two blocks share a CU).
template <class T, int TX, int NW, int KPT, bool MOIST, bool TPI, int WPE = 1>
__global__ void __launch_bounds__(TX * NW, WPE) k_acou_row(Grid g, Params p, Base<T> b,
                                                          Metric<T> m, Coef<T> cf,
                                                          FusedArgs<T> a) {
  static_assert(TX == 32 || TX == 64, "a row of lanes is half or all of a wavefront");
  extern __shared__ double lds[];
  const int L = g.nz + 1;
  double* sD = lds + kTwistPad * TX;  // P_k, then D_k -> ds_k -> W_k (k in [0, nz])
  double* sA = sD + L * TX;  // A_k            k in [1, nz)
  double* sB = sA + L * TX;  // B_k            k in [1, nz)
  double* sC = sB + L * TX;  // C_k -> cs_k    k in [1, nz)
  double* sE = sC + L * TX;  // E_k            k in [0, nz)
  double* sV = sE + L * TX;  // A_v at level k k in [0, nz)
  T* sU = reinterpret_cast<T*>(sV + L * TX);  // terrain: u, v(j), v(j+1) per level
  T* sV0 = sU + g.nz * TX;
  T* sV1 = sV0 + g.nz * TX;

  const int lx = threadIdx.x, ry = threadIdx.y;
  const BlockId bid = xcd_block();
  const int ci = bid.x * (TX - 2) - 1 + lx;  // this lane's cell; its u face is ci
  const int j = bid.y;
  const bool col = lx >= 1 && lx <= TX - 2 && ci < g.nx;  // owns column (ci, j)
  const int cx = ci < g.nx ? ci : g.nx;  // lanes past the first east ghost repeat it
  const int NX = g.NX, PL = g.plane;
  const UVArgs<T>& q = a.uv;
  const T dtau_t = q.dtau;
  const double dtau = (double)q.dtau;
  const int nxm = g.nx - 1, nym = g.ny - 1;

  // Lateral addressing of the three rows j-1, j, j+1 of this lane's cell:
  // level-0 index, level-0 index of the mapped interior cell (ghost reads of
  // the current pi') and whether the cell is interior.
  int crow[3], cmap[3];
  bool inrow[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int jj = j - 1 + r;
    crow[r] = g.id(cx, jj, 0);
    inrow[r] = cx >= 0 && cx < g.nx && jj >= 0 && jj < g.ny;
    cmap[r] = g.id(map_lat(cx, g.nx, nxm, p.bcx), map_lat(jj, g.ny, nym, p.bcy), 0);
  }
  // pi* = pi' + smdiv (pi' - pi'prev) of row r at level kz (UVDev::pstar).
  auto pstar = [&](int r, int kz) -> double {
    const bool in = inrow[r] && (unsigned)kz < (unsigned)g.nz;
    const int c = crow[r] + kz * PL;
    const int km = kz < 0 ? 0 : (kz > g.nz - 1 ? g.nz - 1 : kz);
    const int cm = in ? c : cmap[r] + km * PL;
    const double pc = q.pcur[cm];
    const double pp = (in ? q.pprev : q.pn)[c];
    return pc + p.smdiv * (pc - pp);
  };
  // Base-state and coefficient column: column (0, 0) when every column holds
  // the same profile.
  const int chom = g.id(0, 0, 0);
  // theta_v of row r at level k (UVDev::thv), and theta' itself.
  auto thv = [&](int r, int k, T& th) -> T {
    const int c = crow[r] + k * PL;
    th = (inrow[r] ? q.thp_cur : q.thp_n)[c];
    const T t = b.thb[p.hom ? chom + k * PL : c] + th;
    if constexpr (MOIST) return t * ((T)1 + (T)kEps61 * q.qv[c]);
    return t;
  };

  const int c2 = g.id2(cx, j);
  const T mapc = m.mapf[c2], mapw = m.mapf[c2 - 1], maps = m.mapf[c2 - NX],
          mapn = m.mapf[c2 + NX];
  const T hxu = m.hxu[c2], hyv0 = m.hyv[c2], hyv1 = m.hyv[c2 + NX];
  const T mu = (T)0.5 * (mapw + mapc);
  const T mv0 = (T)0.5 * (maps + mapc), mv1 = (T)0.5 * (mapc + mapn);
  const T ju = (T)0.5 * (m.jac[c2 - 1] + m.jac[c2]);
  const T jv0 = (T)0.5 * (m.jac[c2 - NX] + m.jac[c2]);
  const T jv1 = (T)0.5 * (m.jac[c2] + m.jac[c2 + NX]);
  // The face-nx lane of a periodic x row, and the faces of an open edge.
  const bool uwrap = !p.bcx && ci == g.nx;
  const bool urad_w = p.bcx && ci == 0, urad_e = p.bcx && ci == g.nx;
  const bool vwrap = !p.bcy && j == nym;  // v(j+1) = v(ny) = v(0)
  const bool vrad_s = p.bcy && j == 0, vrad_n = p.bcy && j == nym;
  const UVDev<T> d{g, p, b, m, q};

  // ---- phase 0: horizontal step
  T Un[KPT], Ue[KPT], V0[KPT], V1[KPT], Th[KPT];
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    const int kc = k < g.nz ? k : g.nz - 1;
    const int kk = kc + g.ng;
    T thm, th0, thp1;
    const double ps_m = pstar(0, kc), ps_0 = pstar(1, kc), ps_p = pstar(2, kc);
    const T tv_m = thv(0, kc, thm), tv_0 = thv(1, kc, th0), tv_p = thv(2, kc, thp1);
    Th[t] = th0;
    const int c = crow[1] + kc * PL;
    const T usrc = q.u_src[c], tus = q.tu[c];
    const T v0s = q.v_src[c], tv0s = q.tv[c];
    const T v1s = q.v_src[c + NX], tv1s = q.tv[c + NX];
    // west neighbour cell of this lane's u face
    const double ps_w = __shfl_up(ps_0, 1, TX);
    const T tv_w = __shfl_up(tv_0, 1, TX);
    double gu = (ps_0 - ps_w) * g.rdx;
    double gv0 = (ps_0 - ps_m) * g.rdy;
    double gv1 = (ps_p - ps_0) * g.rdy;
    if (p.terrain) {  // -(z_x / J) dpi*/dzeta on the faces with a slope
      const T fac = (T)1 - m.zeta_c[kk] / m.zt;
      const T dzc2 = m.zeta_c[kk + 1] - m.zeta_c[kk - 1];
      const double dm = pstar(0, kc + 1) - pstar(0, kc - 1);
      const double d0 = pstar(1, kc + 1) - pstar(1, kc - 1);
      const double dp = pstar(2, kc + 1) - pstar(2, kc - 1);
      const double dw = __shfl_up(d0, 1, TX);
      if (hxu != (T)0) gu -= (double)(hxu * fac / ju) * (0.5 * (dw + d0) / (double)dzc2);
      if (hyv0 != (T)0) gv0 -= (double)(hyv0 * fac / jv0) * (0.5 * (dm + d0) / (double)dzc2);
      if (hyv1 != (T)0) gv1 -= (double)(hyv1 * fac / jv1) * (0.5 * (d0 + dp) / (double)dzc2);
    }
    const T thbu = (T)0.5 * (tv_w + tv_0);
    const T thbv0 = (T)0.5 * (tv_m + tv_0);
    const T thbv1 = (T)0.5 * (tv_0 + tv_p);
    T un = usrc + dtau_t * (-(T)kCp * thbu * mu * (T)gu + tus);
    T vn0 = v0s + dtau_t * (-(T)kCp * thbv0 * mv0 * (T)gv0 + tv0s);
    T vn1 = v1s + dtau_t * (-(T)kCp * thbv1 * mv1 * (T)gv1 + tv1s);
    (void)thm;
    (void)thp1;
    // open x edges (k_bc_rad_u_x kernels.cu:649-667): u at face 1 / nx-1 is
    // the neighbour lane's interior value
    const T u_east = __shfl_down(un, 1, TX);
    const T u_west = __shfl_up(un, 1, TX);
    if (uwrap) un = d.u_step(0, j, kc);  // periodic u(nx) = u(0)
    if (urad_w) {
      if (usrc >= (T)0) {
        un = b.ub[kk];
      } else {
        const T cb = usrc - (T)p.cstar;
        un = usrc - dtau_t * cb * (u_east - usrc) / (T)g.dx;
      }
    }
    if (urad_e) {
      if (usrc <= (T)0) {
        un = b.ub[kk];
      } else {
        const T cb = usrc + (T)p.cstar;
        un = usrc - dtau_t * cb * (usrc - u_west) / (T)g.dx;
      }
    }
    // open y edges (k_bc_rad_v_y :670-688) and the periodic v(ny) = v(0)
    const T vg0 = vn0, vg1 = vn1;
    if (vrad_s) {
      if (v0s >= (T)0) {
        vn0 = b.vb[kk];
      } else {
        const T cb = v0s - (T)p.cstar;
        vn0 = v0s - dtau_t * cb * (vg1 - v0s) / (T)g.dy;
      }
    }
    if (vwrap) vn1 = d.v_step(cx, 0, kc);
    if (vrad_n) {
      if (v1s <= (T)0) {
        vn1 = b.vb[kk];
      } else {
        const T cb = v1s + (T)p.cstar;
        vn1 = v1s - dtau_t * cb * (v1s - vg0) / (T)g.dy;
      }
    }
    Un[t] = un;
    Ue[t] = __shfl_down(un, 1, TX);
    V0[t] = vn0;
    V1[t] = vn1;
    if (k < g.nz) {
      // owned faces: the west face of an owned column, and face nx (only
      // the last block of the row has a lane at cell nx)
      if (col || ci == g.nx) q.u_dst[c] = un;
      if (col) {
        q.v_dst[c] = vn0;
        if (j == nym) q.v_dst[c + NX] = vn1;
      }
      if (p.terrain) {
        sU[k * TX + lx] = un;
        sV0[k * TX + lx] = vn0;
        sV1[k * TX + lx] = vn1;
      }
    }
  }
  if (p.terrain) __syncthreads();

  // ---- phase 1: assembly (k_acou_wpi kernels.cu:552-585), FP64
  const double c1 = (1.0 + p.beta) * 0.5;
  const double c2w = (1.0 - p.beta) * 0.5;
  const int c2d = g.id2(col ? ci : 0, j);
  const double J = (double)m.jac[c2d];
  const double hxc = (double)m.hx[c2d];
  const double hyc = (double)m.hy[c2d];
  const double mc = (double)m.mapf[c2d];
  const double rJ = rcp_nr(J);
  const bool flat = hxc == 0.0 && hyc == 0.0;
  const bool rim = TPI && a.tpw[c2d] > (T)0;  // columns with a pi' tendency
  const int c0 = g.id(col ? ci : 0, j, 0);
  const int cb = p.hom ? chom : c0;
  const int ox = p.hom ? 0 : 1, oy = p.hom ? 0 : NX;
  const int lxe = lx < TX - 1 ? lx + 1 : lx;
  auto sv = [&](int k) -> double {  // cross-metric vertical flux at w level k
    if (k <= 0 || k >= g.nz || flat) return 0.0;
    const double fac = 1.0 - (double)m.zeta_w[k + g.ng] * m.rzt;
    const double uw = 0.25 * ((double)sU[(k - 1) * TX + lx] + (double)sU[(k - 1) * TX + lxe] +
                              (double)sU[k * TX + lx] + (double)sU[k * TX + lxe]);
    const double vw = 0.25 * ((double)sV0[(k - 1) * TX + lx] + (double)sV1[(k - 1) * TX + lx] +
                              (double)sV0[k * TX + lx] + (double)sV1[k * TX + lx]);
    return (double)cf.aw[cb + k * PL] * (uw * hxc + vw * hyc) * fac;
  };

  double Pk[KPT], Ek[KPT], Gk[KPT], Rk[KPT];
  T Avk[KPT], Avk1[KPT];
  double w0 = 0.0;
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    const bool act = col && k < g.nz;
    const int kc = k < g.nz ? k : g.nz - 1;
    const int c = c0 + kc * PL;
    const int cm = c - PL;
    const int kk = kc + g.ng;
    const int cc = cb + kc * PL;  // base-state / coefficient index
    const double Kt = (double)cf.kt3[cc];
    const double Dh = mc * (((double)cf.rtjx[cc + ox] * (double)Ue[t] -
                             (double)cf.rtjx[cc] * (double)Un[t]) *
                                g.rdx +
                            ((double)cf.rtjy[cc + oy] * (double)V1[t] -
                             (double)cf.rtjy[cc] * (double)V0[t]) *
                                g.rdy);
    const double rdzc = m.rdzc[kk];
    const T awk = kc >= 1 ? cf.aw[cc] : (T)0;
    const T awk1 = kc + 1 < g.nz ? cf.aw[cc + PL] : (T)0;
    const double avk = (double)awk, avk1 = (double)awk1;
    const double wk = (double)a.w_src[c];
    const double wkp = (double)a.w_src[c + PL];
    const double tpi = rim ? (double)a.tpi[c] : 0.0;
    const double pk = q.pcur[c];
    const double pkm = q.pcur[cm];
    const double thk = (double)Th[t];
    const double thm = (double)q.thp_cur[cm];
    const double thbw = (double)b.thbw[cc];
    const double twk = (double)a.tw[c];
    double mfac = 1.0;
    if constexpr (MOIST) mfac = 1.0 + kEps61 * 0.5 * ((double)q.qv[cm] + (double)q.qv[c]);
    Pk[t] = pk + dtau * tpi -
            dtau * Kt *
                (Dh - (sv(kc + 1) - sv(kc)) * rdzc + c2w * (avk1 * wkp - avk * wk) * rdzc);
    Ek[t] = dtau * Kt * c1 * rdzc;
    Avk[t] = awk;
    Avk1[t] = awk1;
    const double thvw = (thbw + 0.5 * (thm + thk)) * mfac;
    const double rJdzw = rJ * m.rdzw[kk];
    const double buoy = kGrav * 0.5 * (thm + thk) * rcp_nr(thbw);
    Gk[t] = kc >= 1 ? dtau * kCp * thvw * c1 * rJdzw : 0.0;
    Rk[t] = kc >= 1 ? wk + dtau * (buoy + twk) - dtau * kCp * thvw * c2w * (pk - pkm) * rJdzw
                    : 0.0;
    if (k == 0) {  // diagnostic w(0) = u h_x + v h_y at the surface
      const double us = 0.5 * ((double)Un[t] + (double)Ue[t]);
      const double vs = 0.5 * ((double)V0[t] + (double)V1[t]);
      w0 = us * hxc + vs * hyc;
    }
    if (act) {  // P, E and A_v of this level for the rows of the level above
      sD[k * TX + lx] = Pk[t];
      sE[k * TX + lx] = Ek[t];
      sV[k * TX + lx] = (double)Avk[t];
    }
  }
  __syncthreads();

  // ---- phase 2: tridiagonal rows (kernels.cu:596-603), with P, E and A_v
  // of the level below from LDS
  double Dk[KPT];
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    Dk[t] = 0.0;
    if (col && k >= 1 && k < g.nz) {
      const int bl = (k - 1) * TX + lx;
      const double e_prev = sE[bl];
      const double av_prev = sV[bl];
      const double G = Gk[t];
      sA[k * TX + lx] = -G * e_prev * av_prev;
      sB[k * TX + lx] = 1.0 + G * (double)Avk[t] * (Ek[t] + e_prev);
      sC[k * TX + lx] = -G * Ek[t] * (double)Avk1[t];
      Dk[t] = Rk[t] - G * (Pk[t] - sD[bl]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    if (col && k >= 1 && k < g.nz) sD[k * TX + lx] = Dk[t];
  }
  __syncthreads();

  // ---- phase 3: tridiagonal solve, twisted at row nz/2 (thread rows 0 and
  // 1 take the top and the bottom of each column)
  if (ry == 0 && col) twist_sweep_top<TX>(sA, sB, sC, sD, lx, g.nz);
  if (ry == 1 && col) twist_sweep_bottom<TX>(sA, sB, sC, sD, lx, g.nz);
  __syncthreads();
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    if (col && k >= 1 && k < g.nz) twist_row_coefs(sA, sB, sC, sD, k * TX + lx, k, g.nz);
  }
  __syncthreads();
  double xh = 0.0;
  if (ry <= 1 && col) xh = twist_meet<TX>(sA, sB, sC, sD, lx, g.nz);
  __syncthreads();
  if (ry <= 1 && col) twist_back<TX>(sA, sC, sD, lx, g.nz, ry == 0, xh, w0);
  __syncthreads();

  // ---- phase 4: pi', theta', w (theta', T_theta and dtheta/dz re-read
  // here rather than held in registers across the solve)
#pragma unroll
  for (int t = 0; t < KPT; ++t) {
    const int k = ry + NW * t;
    if (col && k < g.nz) {
      const int c = c0 + k * PL;
      const int cc = cb + k * PL;
      const double Wk = sD[k * TX + lx], Wk1 = sD[(k + 1) * TX + lx];
      a.pip_dst[c] = Pk[t] - Ek[t] * ((double)Avk1[t] * Wk1 - (double)Avk[t] * Wk);
      const double wc = 0.5 * (Wk + Wk1);
      a.thp_dst[c] = (T)((double)q.thp_cur[c] +
                         dtau * ((double)a.tth[c] - wc * (double)b.dthbdz[cc]));
      a.w_dst[c] = (T)Wk;
    }
  }
}

}  // namespace wfe::dev
