#pragma once
// Stage-level kernels: ghost fills, mass fluxes, slow tendencies,
// positive-definite moisture transport, boundary relaxation, moisture stage
// update, |f| max. Templated on the storage/arithmetic type T (double for
// the plain FP64 mode, float for the mixed mode).
#include "hip/device_common.hpp"

namespace wfe::dev {

// ------------------------------------------------------------ ghost fills
// apply_bcs (src/dynamics/kernels.cu:805-835): vertical zero-gradient (w:
// w(nz) = 0 and odd reflections, :695-723), then x (periodic or
// zero-gradient up to the field's last valid index, :725-758), then y
// (:735-770).

constexpr int kMaxFill = 9;
enum FillKind { kFillScalar = 0, kFillW = 1 };

// Up to kMaxFill fields of element type E plus per-field kind / last index.
template <class E>
struct FillList {
  E* f[kMaxFill];
  int kind[kMaxFill];
  int xlast[kMaxFill];
  int ylast[kMaxFill];
  int n;
};

// All ghost cells of a state set in one pass. The reference's sequence
// (vertical fill of every padded column, then x over every padded row, then
// y over every padded column) leaves each ghost cell equal to the vertical
// rule applied to the laterally mapped interior column, so every ghost value
// is computed directly from interior cells, which this kernel never writes
// (w(nz) = 0 aside, which no other ghost reads). One thread per ghost cell:
// the 2 ng + 1 full slabs k in [-ng, 0) and [nz, nz + ng], then the lateral
// ring of every interior level. All loads of a cell are issued before any
// store.
struct GhostCell {
  int i, j, k;
};

__device__ __forceinline__ GhostCell ghost_cell(const Grid& g, int idx) {
  const int hx = g.nx + 2 * g.ng, hy = g.ny + 2 * g.ng;
  const int plane = hx * hy;
  const int nslab = 2 * g.ng + 1;
  GhostCell q;
  if (idx < nslab * plane) {
    const int lev = idx / plane, r = idx - lev * plane;
    q.k = lev < g.ng ? lev - g.ng : g.nz + (lev - g.ng);
    q.j = r / hx - g.ng;
    q.i = r - (q.j + g.ng) * hx - g.ng;
    return q;
  }
  const int ring = plane - g.nx * g.ny;
  int r = idx - nslab * plane;
  q.k = r / ring;
  r -= q.k * ring;
  const int band = g.ng * hx;  // ng full rows below j = 0, and above j = ny - 1
  if (r < 2 * band) {
    const int top = r >= band;
    r -= top * band;
    const int jr = r / hx;
    q.j = top ? g.ny + jr : jr - g.ng;
    q.i = r - jr * hx - g.ng;
  } else {
    r -= 2 * band;  // rows j in [0, ny): ng columns west and east
    q.j = r / (2 * g.ng);
    const int c = r - q.j * 2 * g.ng;
    q.i = c < g.ng ? c - g.ng : g.nx + (c - g.ng);
  }
  return q;
}

__host__ __device__ inline int ghost_count(const Grid& g) {
  const int hx = g.nx + 2 * g.ng, hy = g.ny + 2 * g.ng;
  return (2 * g.ng + 1) * hx * hy + g.nz * (hx * hy - g.nx * g.ny);
}

// Periodic wrap of an index that lies a few periods at most outside [0, n),
// without an integer division (ghost indices are within ng of the interior).
__device__ __forceinline__ int wrap_near(int i, int n) {
  while (i < 0) i += n;
  while (i >= n) i -= n;
  return i;
}

// Value the fill leaves in ghost cell (., ., kk) of f, read from the
// laterally mapped interior column (im, jm).
template <class E>
__device__ __forceinline__ E ghost_value(const Grid& g, const E* f, int kind, int im, int jm,
                                         int kk) {
  if (kind == kFillW) {
    if (kk == g.nz) return (E)0;
    if (kk < 0) return (E)2 * f[g.id(im, jm, 0)] - f[g.id(im, jm, -kk)];
    if (kk > g.nz) return -f[g.id(im, jm, 2 * g.nz - kk)];
    return f[g.id(im, jm, kk)];
  }
  const int km = kk < 0 ? 0 : (kk > g.nz - 1 ? g.nz - 1 : kk);
  return f[g.id(im, jm, km)];
}

template <class T>
__global__ void __launch_bounds__(256) k_fill_all(Grid g, int bcx, int bcy, FillList<T> L,
                                                  FillList<double> Lp) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ghost_count(g)) return;
  const GhostCell q = ghost_cell(g, idx);
  // Lateral maps (map_lat): the periodic wrap is the same for every field,
  // the open-boundary clamp depends on the field's last index.
  const int ip = wrap_near(q.i, g.nx), jp = wrap_near(q.j, g.ny);
  auto mx = [&](int last) { return bcx ? (q.i < 0 ? 0 : (q.i > last ? last : q.i)) : ip; };
  auto my = [&](int last) { return bcy ? (q.j < 0 ? 0 : (q.j > last ? last : q.j)) : jp; };
  T v[kMaxFill];
  double vp[2];
#pragma unroll
  for (int f = 0; f < kMaxFill; ++f)
    if (f < L.n) v[f] = ghost_value(g, L.f[f], L.kind[f], mx(L.xlast[f]), my(L.ylast[f]), q.k);
#pragma unroll
  for (int f = 0; f < 2; ++f)
    if (f < Lp.n)
      vp[f] = ghost_value(g, Lp.f[f], Lp.kind[f], mx(Lp.xlast[f]), my(Lp.ylast[f]), q.k);
  const int c = g.id(q.i, q.j, q.k);
#pragma unroll
  for (int f = 0; f < kMaxFill; ++f)
    if (f < L.n) L.f[f][c] = v[f];
#pragma unroll
  for (int f = 0; f < 2; ++f)
    if (f < Lp.n) Lp.f[f][c] = vp[f];
}

// ------------------------------------------------------------ mass fluxes
// rho~-weighted contravariant mass fluxes on the ring the advection reads
// (k_massflux_xy / k_massflux_z, kernels.cu:68-97): mfx, mfy on
// i,j in [-2, n+2]; mfz = rho_w (w - (u^xz h_x + v^yz h_y)(1 - zeta_w/zt))
// on i,j in [-2, n+1], k in [1, nz) (k = 0 and nz stay 0: Omega = 0).
// The same launch forms the flux divergence for the advective-consistency
// term (k_divergence kernels.cu:99-110) on i,j in [-1, n], recomputing the
// three upper-face fluxes of the cell instead of reading them back.
template <class T>
struct MassFlux {
  const Grid& g;
  const Base<T>& b;
  const Metric<T>& m;
  CArr<T> u;
  CArr<T> v;
  CArr<T> w;
  int k;  // level of the faces fx, fy
  __device__ __forceinline__ T fx(int c, int c2) const {
    const T rj = b.rhob[b.at(g, c, k)] * m.jac[c2];
    return (T)0.5 * (b.rhob[b.at(g, c - 1, k)] * m.jac[c2 - 1] + rj) * u[c];
  }
  __device__ __forceinline__ T fy(int c, int c2) const {
    const T rj = b.rhob[b.at(g, c, k)] * m.jac[c2];
    return (T)0.5 * (b.rhob[b.at(g, c - g.NX, k)] * m.jac[c2 - g.NX] + rj) * v[c];
  }
  // k in [1, nz); 0 elsewhere
  __device__ __forceinline__ T fz(int c, int c2, int k) const {
    if (k < 1 || k >= g.nz) return (T)0;
    const T fac = (T)1 - m.zeta_w[k + g.ng] * (T)m.rzt;
    const int cm = c - g.plane;
    const T uw = (T)0.25 * (u[cm] + u[cm + 1] + u[c] + u[c + 1]);
    const T vw = (T)0.25 * (v[cm] + v[cm + g.NX] + v[c] + v[c + g.NX]);
    const T cross = (uw * m.hx[c2] + vw * m.hy[c2]) * fac;
    return b.rhobw[b.at(g, c, k)] * (w[c] - cross);
  }
};

template <class T>
__global__ void k_massflux(Grid g, Base<T> b, Metric<T> m, const T* __restrict__ u_,
                           const T* __restrict__ v_, const T* __restrict__ w_, T* mfx_,
                           T* mfy_, T* mfz_, T* div_) {
  const CArr<T> u{u_}, v{v_}, w{w_};
  const Arr<T> mfx{mfx_}, mfy{mfy_}, mfz{mfz_}, div{div_};
  const BlockId bid = xcd_block();
  const int i = (int)(bid.x * blockDim.x + threadIdx.x) - 2;
  const int j = (int)(bid.y * blockDim.y + threadIdx.y) - 2;
  const int k = bid.z;
  if (i > g.nx + 2 || j > g.ny + 2 || k >= g.nz) return;
  const MassFlux<T> F{g, b, m, u, v, w, k};
  const int c = g.id(i, j, k);
  const int c2 = g.id2(i, j);
  const T fxc = F.fx(c, c2), fyc = F.fy(c, c2);
  mfx[c] = fxc;
  mfy[c] = fyc;
  const T fzc = F.fz(c, c2, k);
  if (k >= 1 && i <= g.nx + 1 && j <= g.ny + 1) mfz[c] = fzc;
  if (i >= -1 && i <= g.nx && j >= -1 && j <= g.ny) {
    const T fxr = F.fx(c + 1, c2 + 1);
    const T fyr = F.fy(c + g.NX, c2 + g.NX);
    const T fzt = F.fz(c + g.plane, c2, k + 1);
    div[c] = (fxr - fxc) * (T)g.rdx + (fyr - fyc) * (T)g.rdy + (fzt - fzc) * (T)m.rdzc[k + g.ng];
  }
}

// Flux through the face whose upper cell is c (stride s along the normal):
// F = f * iface5(q[c-3s .. c+2s], f).
template <class A, class T>
__device__ __forceinline__ T face_flux(const A& q, int c, int s, T f) {
  return f * iface5(q[c - 3 * s], q[c - 2 * s], q[c - s], q[c], q[c + s], q[c + 2 * s], f);
}

// Reciprocal grid spacings of one point (1/dx, 1/dy, 1/dzeta).
template <class T>
struct Rsp {
  T x, y, z;
};

// Advective-consistent flux-form tendency of a cell-centred scalar at c:
// -(1/rho~)[div(MF q) - q div(MF)] (k_tend_thp / k_tend_q, kernels.cu:114-179),
// with the divisions by dx, dy, dzeta and rho~ taken as products with
// reciprocals (rrt = 1/rho~).
template <class A, class T>
__device__ __forceinline__ T scalar_adv(const Grid& g, const A& q, const A& mfx, const A& mfy,
                                        const A& mfz, const A& div, int c, Rsp<T> r, T rrt) {
  const T fxL = mfx[c], fxR = mfx[c + 1];
  const T fyL = mfy[c], fyR = mfy[c + g.NX];
  const T fzB = mfz[c], fzT = mfz[c + g.plane];
  const T FxL = face_flux(q, c, 1, fxL);
  const T FxR = face_flux(q, c + 1, 1, fxR);
  const T FyL = face_flux(q, c, g.NX, fyL);
  const T FyR = face_flux(q, c + g.NX, g.NX, fyR);
  const T FzB = face_flux(q, c, g.plane, fzB);
  const T FzT = face_flux(q, c + g.plane, g.plane, fzT);
  const T fdiv = (FxR - FxL) * r.x + (FyR - FyL) * r.y + (FzT - FzB) * r.z;
  return (-fdiv + q[c] * div[c]) * rrt;
}

// Constant-K Laplacian with the local physical vertical spacing
// (k_diffuse kernels.cu:430-443); rdz2 = 1 / (dzeta J)^2.
template <class A, class T>
__device__ __forceinline__ T laplacian(const Grid& g, const A& f, int c, T rdz2) {
  const T cc = f[c];
  return (f[c + 1] - (T)2 * cc + f[c - 1]) * (T)(g.rdx * g.rdx) +
         (f[c + g.NX] - (T)2 * cc + f[c - g.NX]) * (T)(g.rdy * g.rdy) +
         (f[c + g.plane] - (T)2 * cc + f[c - g.plane]) * rdz2;
}

// Moist part of the w-level buoyancy from the time-n moisture,
// g (0.61 (q_v - qbar_v) - q_c - q_r) averaged to the w level (the moisture
// terms of k_acou_wpi's right-hand side, kernels.cu:560-585); it is constant
// over the substeps of a step, so it is added to T_w once per stage.
template <class T>
struct MoistBuoy {
  CArr<T> qv, qc, qr;
  int on;
};

template <class T, class A>
__device__ __forceinline__ T moist_buoyancy_w(const Base<T>& b, const A& qv, const A& qc,
                                              const A& qr, int c, int cb, int PL) {
  const int cm = c - PL;
  const T qvw = (T)0.5 * (qv[cm] + qv[c]);
  const T dqv = qvw - (T)0.5 * (b.qvb[cb - PL] + b.qvb[cb]);
  const T qcr = (T)0.5 * (qc[cm] + qc[c]) + (T)0.5 * (qr[cm] + qr[c]);
  return (T)kGrav * ((T)kEps61 * dqv - qcr);
}

// Moisture stage update q_out = q_n + dt_stage T_q (k_stage_q,
// kernels.cu:268-273) applied where T_q is formed, so T_q is never stored;
// on = 0 stores T_q instead.
template <class T>
struct QStage {
  CArr<T> qn[3];
  Arr<T> qo[3];
  T dts;
  int on;
};

// All slow tendencies (compute_tendencies, kernels.cu:872-913) of KZ
// consecutive levels of one column per thread: u, v, w (k >= 1), theta'
// (+ Rayleigh), the three moisture species when the non-PD scheme is used,
// and the const-K diffusion of u, v, w, theta' when diff_K > 0. Every array
// pointer is restrict-qualified, so the overlapping vertical stencil reads
// of consecutive levels can be loaded once. Divisions by grid spacings and
// densities are products with reciprocals (rcp_nr for the densities).
//   tq0..2   : T_q, or with stage_on the stage's q = q_n + dts T_q
//              (k_stage_q, kernels.cu:268-273) so T_q is never stored;
//   qn0..2   : the time-n moisture (stage update and, with buoy_on, the
//              moist buoyancy added to T_w).
template <class T, int KZ>
__global__ void __launch_bounds__(256) k_tendencies(
    Grid g, Params p, Base<T> b, Metric<T> m, T dt, const T* __restrict__ u_,
    const T* __restrict__ v_, const T* __restrict__ w_, const T* __restrict__ thp_,
    const T* __restrict__ qv_, const T* __restrict__ qc_, const T* __restrict__ qr_,
    const T* __restrict__ mfx_, const T* __restrict__ mfy_, const T* __restrict__ mfz_,
    const T* __restrict__ div_, T* __restrict__ tu_, T* __restrict__ tv_, T* __restrict__ tw_,
    T* __restrict__ tth_, T* __restrict__ tq0_, T* __restrict__ tq1_, T* __restrict__ tq2_,
    const T* __restrict__ qn0_, const T* __restrict__ qn1_, const T* __restrict__ qn2_, int do_q,
    T dts, int stage_on, int buoy_on) {
  const CArr<T> u{u_}, v{v_}, w{w_}, thp{thp_}, qv{qv_}, qc{qc_}, qr{qr_};
  const CArr<T> mfx{mfx_}, mfy{mfy_}, mfz{mfz_}, div{div_};
  const CArr<T> qn0{qn0_}, qn1{qn1_}, qn2{qn2_};
  const Arr<T> tu{tu_}, tv{tv_}, tw{tw_}, tth{tth_}, tq0{tq0_}, tq1{tq1_}, tq2{tq2_};
  const BlockId bid = xcd_block();
  const int i = bid.x * blockDim.x + threadIdx.x;
  const int j = bid.y * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
#pragma unroll
  for (int kz = 0; kz < KZ; ++kz) {
    const int k = bid.z * KZ + kz;
    if (k >= g.nz) break;
    const int kk = k + g.ng;
    const int c = g.id(i, j, k);
    const int c2 = g.id2(i, j);
    const int NX = g.NX, PL = g.plane;
    const T rdzc = (T)m.rdzc[kk];
    const Rsp<T> rc{(T)g.rdx, (T)g.rdy, rdzc};
    const T jc = m.jac[c2];
    const int cb = b.at(g, c, k);  // base-state index of this point
    const T rt = b.rhob[cb] * jc;
    const T K = (T)p.diff_K;
    const T rjc = rcp_nr(jc);
    auto rdz2 = [&](T rdz) { return rdz * rdz * rjc * rjc; };  // 1 / (dzeta J)^2

    // ---- u (k_tend_u, kernels.cu:308-346)
    {
      const T fcL = (T)0.5 * (mfx[c - 1] + mfx[c]);
      const T fcR = (T)0.5 * (mfx[c] + mfx[c + 1]);
      const T FxL = face_flux(u, c, 1, fcL);
      const T FxR = face_flux(u, c + 1, 1, fcR);
      const T fyB = (T)0.5 * (mfy[c - 1] + mfy[c]);
      const T fyT = (T)0.5 * (mfy[c - 1 + NX] + mfy[c + NX]);
      const T FyB = face_flux(u, c, NX, fyB);
      const T FyT = face_flux(u, c + NX, NX, fyT);
      const T fzB = (T)0.5 * (mfz[c - 1] + mfz[c]);
      const T fzT = (T)0.5 * (mfz[c - 1 + PL] + mfz[c + PL]);
      const T FzB = face_flux(u, c, PL, fzB);
      const T FzT = face_flux(u, c + PL, PL, fzT);
      const T fdiv = (FxR - FxL) * rc.x + (FyT - FyB) * rc.y + (FzT - FzB) * rc.z;
      const T dv = (T)0.5 * (div[c - 1] + div[c]);
      const T ru = (T)0.5 * (b.rhob[b.at(g, c - 1, k)] * m.jac[c2 - 1] + rt);
      const T vavg = (T)0.25 * (v[c - 1] + v[c - 1 + NX] + v[c] + v[c + NX]);
      const T fc = (T)0.5 * (m.fcor[c2 - 1] + m.fcor[c2]);
      const T vref = p.coriolis_use_ub ? b.vb[kk] : (T)0;
      const T relax = -m.ray_c[kk] * (u[c] - b.ub[kk]);
      T t = (-fdiv + u[c] * dv) * rcp_nr(ru) + fc * (vavg - vref) + relax;
      if (K > (T)0) t += K * laplacian(g, u, c, rdz2(rdzc));
      tu[c] = t;
    }
    // ---- v (k_tend_v, kernels.cu:348-385)
    {
      const T fcB = (T)0.5 * (mfy[c - NX] + mfy[c]);
      const T fcT = (T)0.5 * (mfy[c] + mfy[c + NX]);
      const T FyB = face_flux(v, c, NX, fcB);
      const T FyT = face_flux(v, c + NX, NX, fcT);
      const T fxL = (T)0.5 * (mfx[c - NX] + mfx[c]);
      const T fxR = (T)0.5 * (mfx[c + 1 - NX] + mfx[c + 1]);
      const T FxL = face_flux(v, c, 1, fxL);
      const T FxR = face_flux(v, c + 1, 1, fxR);
      const T fzB = (T)0.5 * (mfz[c - NX] + mfz[c]);
      const T fzT = (T)0.5 * (mfz[c - NX + PL] + mfz[c + PL]);
      const T FzB = face_flux(v, c, PL, fzB);
      const T FzT = face_flux(v, c + PL, PL, fzT);
      const T fdiv = (FxR - FxL) * rc.x + (FyT - FyB) * rc.y + (FzT - FzB) * rc.z;
      const T dv = (T)0.5 * (div[c - NX] + div[c]);
      const T rv = (T)0.5 * (b.rhob[b.at(g, c - NX, k)] * m.jac[c2 - NX] + rt);
      const T uavg = (T)0.25 * (u[c - NX] + u[c + 1 - NX] + u[c] + u[c + 1]);
      const T fc = (T)0.5 * (m.fcor[c2 - NX] + m.fcor[c2]);
      const T uref = p.coriolis_use_ub ? b.ub[kk] : (T)0;
      const T relax = -m.ray_c[kk] * (v[c] - b.vb[kk]);
      T t = (-fdiv + v[c] * dv) * rcp_nr(rv) - fc * (uavg - uref) + relax;
      if (K > (T)0) t += K * laplacian(g, v, c, rdz2(rdzc));
      tv[c] = t;
    }
    // ---- w on k >= 1 (k_tend_w, kernels.cu:387-425)
    if (k >= 1) {
      const T rdzw = (T)m.rdzw[kk];
      const T fxL = (T)0.5 * (mfx[c - PL] + mfx[c]);
      const T fxR = (T)0.5 * (mfx[c + 1 - PL] + mfx[c + 1]);
      const T FxL = face_flux(w, c, 1, fxL);
      const T FxR = face_flux(w, c + 1, 1, fxR);
      const T fyB = (T)0.5 * (mfy[c - PL] + mfy[c]);
      const T fyT = (T)0.5 * (mfy[c + NX - PL] + mfy[c + NX]);
      const T FyB = face_flux(w, c, NX, fyB);
      const T FyT = face_flux(w, c + NX, NX, fyT);
      const T fzB = (T)0.5 * (mfz[c - PL] + mfz[c]);
      const T fzT = (T)0.5 * (mfz[c] + mfz[c + PL]);
      const T FzB = face_flux(w, c, PL, fzB);
      const T FzT = face_flux(w, c + PL, PL, fzT);
      const T fdiv = (FxR - FxL) * rc.x + (FyT - FyB) * rc.y + (FzT - FzB) * rdzw;
      const T dv = (T)0.5 * (div[c - PL] + div[c]);
      const T rw = b.rhobw[cb] * jc;
      T relax = -m.ray_w[kk] * w[c];
      const T wk = w[c];
      if (p.w_damping) {
        const T cr = fabs(wk) * dt * rdzw * rjc;
        if (cr > (T)1) relax -= (T)0.3 * (cr - (T)1) * wk * rcp_nr(dt);
      }
      T t = (-fdiv + wk * dv) * rcp_nr(rw) + relax;
      if (K > (T)0) t += K * laplacian(g, w, c, rdz2(rdzw));
      if (buoy_on) t += moist_buoyancy_w(b, qn0, qn1, qn2, c, cb, PL);
      tw[c] = t;
    }
    const T rrt = rcp_nr(rt);
    // ---- theta' (k_tend_thp, kernels.cu:114-146)
    {
      T t = scalar_adv(g, thp, mfx, mfy, mfz, div, c, rc, rrt) - m.ray_c[kk] * thp[c];
      if (K > (T)0) t += K * laplacian(g, thp, c, rdz2(rdzc));
      tth[c] = t;
    }
    // ---- moisture, non-PD scheme (k_tend_q, kernels.cu:149-179)
    if (do_q) {
      const T t0 = scalar_adv(g, qv, mfx, mfy, mfz, div, c, rc, rrt);
      const T t1 = scalar_adv(g, qc, mfx, mfy, mfz, div, c, rc, rrt);
      const T t2 = scalar_adv(g, qr, mfx, mfy, mfz, div, c, rc, rrt);
      if (stage_on) {  // the stage's q = q_n + dt_stage T_q
        tq0[c] = qn0[c] + dts * t0;
        tq1[c] = qn1[c] + dts * t1;
        tq2[c] = qn2[c] + dts * t2;
      } else {
        tq0[c] = t0;
        tq1[c] = t1;
        tq2[c] = t2;
      }
    }
  }
}

// T_w + the moist buoyancy into a separate buffer (phase-level substeps,
// whose T_w comes from the stand-alone tendency call without it).
template <class T>
__global__ void k_fold_moist_buoyancy(Grid g, Base<T> b, MoistBuoy<T> mb, const T* __restrict__ tw,
                                      T* __restrict__ out) {
  const BlockId bid = xcd_block();
  const int i = bid.x * blockDim.x + threadIdx.x;
  const int j = bid.y * blockDim.y + threadIdx.y;
  const int k = bid.z + 1;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const int c = g.id(i, j, k);
  out[c] = tw[c] + moist_buoyancy_w(b, mb.qv, mb.qc, mb.qr, c, b.at(g, c, k), g.plane);
}

// ------------------------------------------ positive-definite moisture (PD)
// Skamarock (2006) flux renormalisation (kernels.cu:187-264): two launches
// for the three species (limiter ratio, then the renormalised divergence).
template <class T>
struct PDSet {
  const T* q[3];
  T* ratio[3];
  T* tq[3];
};

// Flux of species q through the west / south / bottom face of cell c.
template <class T>
struct PDFlux {
  const Grid& g;
  const T* __restrict__ mfx;
  const T* __restrict__ mfy;
  const T* __restrict__ mfz;
  __device__ __forceinline__ T fx(const T* __restrict__ q, int c) const {
    return face_flux(q, c, 1, mfx[c]);
  }
  __device__ __forceinline__ T fy(const T* __restrict__ q, int c) const {
    return face_flux(q, c, g.NX, mfy[c]);
  }
  __device__ __forceinline__ T fz(const T* __restrict__ q, int c) const {
    return face_flux(q, c, g.plane, mfz[c]);
  }
};

// Limiter ratio of each species at interior cells (k_pd_ratio
// kernels.cu:221-238); the six face fluxes (k_qflux_x/y/z :187-218) are
// formed here and again in k_pd_apply instead of being stored.
template <class T>
__global__ void k_pd_ratio(Grid g, Base<T> b, Metric<T> m, T dt, const T* __restrict__ mfx,
                           const T* __restrict__ mfy, const T* __restrict__ mfz, PDSet<T> S) {
  const BlockId bid = xcd_block();
  const int i = bid.x * blockDim.x + threadIdx.x;
  const int j = bid.y * blockDim.y + threadIdx.y;
  const int k = bid.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const int c = g.id(i, j, k);
  const T dzc = m.dzeta_c[k + g.ng];
  const T dx = (T)g.dx, dy = (T)g.dy;
  const T rt = b.rhob[b.at(g, c, k)] * m.jac[g.id2(i, j)];
  const PDFlux<T> F{g, mfx, mfy, mfz};
  for (int s = 0; s < 3; ++s) {
    const T* q = S.q[s];
    const T out = fmax((T)0, F.fx(q, c + 1)) / dx + fmax((T)0, -F.fx(q, c)) / dx +
                  fmax((T)0, F.fy(q, c + g.NX)) / dy + fmax((T)0, -F.fy(q, c)) / dy +
                  fmax((T)0, F.fz(q, c + g.plane)) / dzc + fmax((T)0, -F.fz(q, c)) / dzc;
    const T avail = rt * fmax((T)0, q[c]);
    const T r = (dt * out > (T)1e-20) ? avail / (dt * out) : (T)1;
    S.ratio[s][c] = fmin((T)1, r);
  }
}

// Renormalised flux divergence + consistency term (k_pd_apply
// kernels.cu:241-264). Ratio reads outside the interior follow the
// reference's lateral fill of the ratio field (kernels.cu:865-866);
// vertically they only multiply the zero fluxes at k = 0 and nz.
template <class T>
__global__ void k_pd_apply(Grid g, Params p, Base<T> b, Metric<T> m, const T* __restrict__ mfx,
                           const T* __restrict__ mfy, const T* __restrict__ mfz,
                           const T* __restrict__ mfdiv, PDSet<T> S, QStage<T> qs) {
  const BlockId bid = xcd_block();
  const int i = bid.x * blockDim.x + threadIdx.x;
  const int j = bid.y * blockDim.y + threadIdx.y;
  const int k = bid.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const int c = g.id(i, j, k);
  const T rdzc = (T)m.rdzc[k + g.ng];
  const T rrt = rcp_nr(b.rhob[b.at(g, c, k)] * m.jac[g.id2(i, j)]);
  const PDFlux<T> F{g, mfx, mfy, mfz};
  auto ridx = [&](int ii, int jj, int kz) {
    ii = map_lat(ii, g.nx, g.nx - 1, p.bcx);
    jj = map_lat(jj, g.ny, g.ny - 1, p.bcy);
    kz = kz < 0 ? 0 : (kz > g.nz - 1 ? g.nz - 1 : kz);
    return g.id(ii, jj, kz);
  };
  for (int s = 0; s < 3; ++s) {
    const T* q = S.q[s];
    const T* R = S.ratio[s];
    auto sx = [&](int i0) {
      const T f = F.fx(q, g.id(i0, j, k));
      return f * R[ridx(f >= (T)0 ? i0 - 1 : i0, j, k)];
    };
    auto sy = [&](int j0) {
      const T f = F.fy(q, g.id(i, j0, k));
      return f * R[ridx(i, f >= (T)0 ? j0 - 1 : j0, k)];
    };
    auto sz = [&](int k0) {
      const T f = F.fz(q, g.id(i, j, k0));
      return f * R[ridx(i, j, f >= (T)0 ? k0 - 1 : k0)];
    };
    const T fdiv = (sx(i + 1) - sx(i)) * (T)g.rdx + (sy(j + 1) - sy(j)) * (T)g.rdy +
                   (sz(k + 1) - sz(k)) * rdzc;
    const T tq = (-fdiv + q[c] * mfdiv[c]) * rrt;
    if (qs.on)
      qs.qo[s][c] = qs.qn[s][c] + qs.dts * tq;
    else
      S.tq[s][c] = tq;
  }
}

// ------------------------------------------------- boundary relaxation
// Davies rim relaxation toward the time-interpolated targets
// (k_bdy_relax kernels.cu:292-304): T += w(i,j) [(1-tf) lo + tf hi - f] for
// u, v, theta', qv; the pi' tendency is set in the rim. Outside the rim
// (w = 0) nothing is written: there the pi' tendency is zero by definition
// and the substep kernels do not read it.
template <class T>
struct RelaxSet {
  const T* f[4];      // u v thp qv (estimate state)
  const double* pip;  // estimate pi'
  const T* lo[5];
  const T* hi[5];
  T* t[5];  // tu tv tth tpi tqv
  const T* wgt;
  const T* tf;  // time fraction between the target pair (device scalar)
};

template <class T>
__device__ __forceinline__ void bdy_relax_cell(const Grid& g, const RelaxSet<T>& R, int i, int j,
                                               int k) {
  const int c = g.id(i, j, k);
  const T w = R.wgt[g.id2(i, j)];
  const T tf = *R.tf;
  if (w <= (T)0) return;
  const int fi[4] = {0, 1, 2, 4};
  for (int q = 0; q < 4; ++q) {
    const int v = fi[q];
    R.t[v][c] += w * (((T)1 - tf) * R.lo[v][c] + tf * R.hi[v][c] - R.f[q][c]);
  }
  R.t[3][c] = w * (((T)1 - tf) * R.lo[3][c] + tf * R.hi[3][c] - (T)R.pip[c]);
}

template <class T>
__global__ void k_bdy_relax(Grid g, RelaxSet<T> R) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  bdy_relax_cell(g, R, i, j, k);
}

// The same over the rim only: the columns within rw cells of an edge (every
// column with a positive weight is), enumerated as the south and north
// strips of rw full rows, then the west and east strips of rw columns
// between them; blockIdx.y is the level.
template <class T>
__global__ void k_bdy_relax_rim(Grid g, RelaxSet<T> R, int rw) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  const int strip = rw * g.nx, side = rw * (g.ny - 2 * rw);
  if (q >= 2 * strip + 2 * side) return;
  int i, j;
  if (q < 2 * strip) {
    const int r = q / g.nx;
    i = q - r * g.nx;
    j = r < rw ? r : g.ny - 2 * rw + r;
  } else {
    q -= 2 * strip;
    const int east = q >= side;
    if (east) q -= side;
    const int r = q / rw;
    j = rw + r;
    const int ii = q - r * rw;
    i = east ? g.nx - rw + ii : ii;
  }
  bdy_relax_cell(g, R, i, j, k);
}

// One device scalar set in stream order (kept out of kernel arguments so a
// captured step graph stays valid when the value changes).
template <class T>
__global__ void k_set_scalar(T* dst, T v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *dst = v;
}

// ---------------------------------------------------- moisture stage update
// out = s_n + dt_stage * T (k_stage_q kernels.cu:268-273). The reference
// sweeps every padded point; its ghost results (T = 0 there) are replaced
// by the ghost fill at the start of the next stage before anything reads
// them, so only the interior is updated here.
template <class T>
__global__ void k_stage_q(Grid g, T dt, const T* __restrict__ q0, const T* __restrict__ q1,
                          const T* __restrict__ q2, const T* __restrict__ t0,
                          const T* __restrict__ t1, const T* __restrict__ t2, T* o0, T* o1,
                          T* o2) {
  const BlockId bid = xcd_block();
  const int i = bid.x * blockDim.x + threadIdx.x;
  const int j = bid.y * blockDim.y + threadIdx.y;
  const int k = bid.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const int c = g.id(i, j, k);
  o0[c] = q0[c] + dt * t0[c];
  o1[c] = q1[c] + dt * t1[c];
  o2[c] = q2[c] + dt * t2[c];
}

// ------------------------------------------------------------ |f| max
// NaN-safe max|f| over the interior (k in [0, kmax]) as IEEE bits of a
// non-negative float: NaN bit patterns order above every finite value
// (k_absmax kernels.cu:278-289).
template <class E>
__global__ void k_absmax(Grid g, const E* __restrict__ f, int kmax, unsigned int* result) {
  __shared__ unsigned int smax;
  if (threadIdx.x == 0 && threadIdx.y == 0) smax = 0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i < g.nx && j < g.ny && k <= kmax) {
    const float a = fabsf((float)f[g.id(i, j, k)]);
    atomicMax(&smax, __float_as_uint(a));
  }
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) atomicMax(result, smax);
}

// Streaming copy used to measure the achievable HBM bandwidth and to
// calibrate the FETCH_SIZE / WRITE_SIZE counters against a known byte count.
__global__ void k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

}  // namespace wfe::dev
