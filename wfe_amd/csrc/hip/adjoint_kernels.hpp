#pragma once
// Tangent-linear (TL) and adjoint (AD) kernels of the large step in FP64:
// ghost fill, mass fluxes and divergence, flux-form advection tendencies with
// the Rayleigh layer, the horizontal acoustic step and the vertically
// implicit w-pi' column solve. They linearise the forward kernels around a
// stored trajectory, with the upwind branch of every interface value frozen
// at the trajectory's mass flux. Terrain enters through the Jacobian J, the
// cross-metric vertical mass flux, the terrain pressure-gradient term of the
// horizontal acoustic step, S_v in the column's pi' equation and the
// diagnostic surface w(0) = u h_x + v h_y. Coriolis (any f(x, y)) and
// constant-K diffusion are linear and enter the TL as they are, the AD as
// their transposes. Open lateral boundaries take the radiation condition on
// their edge faces, with the inflow/outflow branch frozen at the trajectory;
// Davies relaxation toward fixed targets adds -w df to the tendencies and
// sets the pi' tendency to -w dpi' in the rim; the vertical-Courant damping
// of w enters on its active branch. With moisture on (no microphysics,
// q_c = q_r = 0), q_v is a sixth linearised field: its transport (flux form,
// or positive-definite with the limiter's branches frozen) and stage update,
// the moist buoyancy in T_w and the virtual-temperature factor
// (1 + 0.61 q_v) of the horizontal pressure gradient and of the column's
// theta_v at w levels.
//
// Every AD kernel is the statement-by-statement transpose of its TL kernel;
// contributions to shared inputs are scattered with FP64 atomics.
//
// The trajectory may be stored in FP32 (a mixed-precision forward): its
// arrays are read through Traj<S>, which widens every element to FP64 at the
// load, so all arithmetic -- trajectory averages and coefficients included --
// is FP64 and TL and AD use bit-identical coefficients.
#include "hip/device_common.hpp"
#include "hip/slow_kernels.hpp"

namespace wfe::dev {

// Read-only trajectory array of storage type S, read as FP64.
template <class S>
struct Traj {
  const S* p;
  __device__ __forceinline__ double operator[](int i) const { return (double)p[i]; }
};

// 5th-order upwind interface weights on q[-2..3] for the branch of vel
// (iface5: centred minus / plus the upwind term).
__device__ __forceinline__ void iface5_weights(double vel, double w[6]) {
  const double s = 1.0 / 60.0;
  if (vel >= 0.0) {
    w[0] = 2 * s;
    w[1] = -13 * s;
    w[2] = 47 * s;
    w[3] = 27 * s;
    w[4] = -3 * s;
    w[5] = 0.0;
  } else {
    w[0] = 0.0;
    w[1] = -3 * s;
    w[2] = 27 * s;
    w[3] = 47 * s;
    w[4] = -13 * s;
    w[5] = 2 * s;
  }
}

template <class Q>
__device__ __forceinline__ double iface_at(const Q& q, int c, int s, double vel) {
  return iface5<double>(q[c - 3 * s], q[c - 2 * s], q[c - s], q[c], q[c + s], q[c + 2 * s], vel);
}

// TL of F = f iface5(q, f) through the face whose upper cell is c.
template <class Q>
__device__ __forceinline__ double tl_face(const Q& q, const double* dq, int c, int s, double f,
                                          double df) {
  return df * iface_at(q, c, s, f) + f * iface_at(dq, c, s, f);
}

// AD of the same: scatters Fb f w_n into qb, returns the adjoint of f.
template <class Q>
__device__ __forceinline__ double ad_face(const Q& q, double* qb, int c, int s, double f,
                                          double Fb) {
  double w[6];
  iface5_weights(f, w);
  const double a = Fb * f;
#pragma unroll
  for (int n = 0; n < 6; ++n)
    if (w[n] != 0.0) atomicAdd(&qb[c + (n - 3) * s], a * w[n]);
  return Fb * iface_at(q, c, s, f);
}

// ------------------------------------------------------------ ghost fill
// Five fields (u v w theta' pi'); the fill is linear, so the TL applies the
// forward fill to the perturbation (k_fill_all) and the AD moves each ghost
// cell's adjoint onto the interior cells it was copied from.
struct AdFill {
  double* f[6];
  int kind[6];
  int xlast[6], ylast[6];
  int n;
};

__global__ void __launch_bounds__(256) k_ad_fill(Grid g, int bcx, int bcy, AdFill A) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ghost_count(g)) return;
  const GhostCell q = ghost_cell(g, idx);
  const int c = g.id(q.i, q.j, q.k);
#pragma unroll
  for (int fi = 0; fi < 6; ++fi) {
    if (fi >= A.n) break;
    double* f = A.f[fi];
    const bool wf = A.kind[fi] == kFillW;
    if (wf && q.k == g.nz) {  // the rigid lid w(nz) = 0 is not a variable
      f[c] = 0.0;
      continue;
    }
    const int im = map_lat(q.i, g.nx, A.xlast[fi], bcx);
    const int jm = map_lat(q.j, g.ny, A.ylast[fi], bcy);
    // A ring cell the fill maps onto itself (u at i = nx, v at j = ny with
    // open boundaries: the last face is a prognostic face) keeps its
    // adjoint; other ring cells add into it meanwhile.
    const bool kin = q.k >= 0 && q.k < g.nz;  // vertical ghosts fold onto other levels
    if (kin && im == q.i && jm == q.j) continue;
    const double gb = f[c];
    f[c] = 0.0;
    if (gb == 0.0) continue;
    if (wf) {
      if (q.k < 0) {
        atomicAdd(&f[g.id(im, jm, 0)], 2.0 * gb);
        atomicAdd(&f[g.id(im, jm, -q.k)], -gb);
      } else if (q.k > g.nz) {
        atomicAdd(&f[g.id(im, jm, 2 * g.nz - q.k)], -gb);
      } else {
        atomicAdd(&f[g.id(im, jm, q.k)], gb);
      }
    } else {
      const int km = q.k < 0 ? 0 : (q.k > g.nz - 1 ? g.nz - 1 : q.k);
      atomicAdd(&f[g.id(im, jm, km)], gb);
    }
  }
}

// ------------------------------------------------------------ mass fluxes
// Linear in (u, v, w): the TL is k_massflux on the perturbation. AD over the
// same ranges (mfx, mfy on i,j in [-2, n+2]; mfz on [-2, n+1], k in
// [1, nz); div on [-1, n]).
__global__ void k_ad_massflux(Grid g, Metric<double> m, double* mfxb, double* mfyb, double* mfzb,
                              const double* divb) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 2;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) - 2;
  const int k = blockIdx.z;
  if (i > g.nx + 2 || j > g.ny + 2 || k >= g.nz) return;
  const int c = g.id(i, j, k);
  // divergence -> fluxes: div = (fx(i+1)-fx(i)) rdx + (fy(j+1)-fy(j)) rdy +
  // (fz(k+1)-fz(k)) rdzc with the same FP64 reciprocals as k_massflux (the
  // TL), so the pair stays an exact transpose when dzeta_c is not
  // representable in the trajectory's storage type
  if (i >= -1 && i <= g.nx && j >= -1 && j <= g.ny) {
    const double d = divb[c];
    if (d != 0.0) {
      const double dx = d * g.rdx, dy = d * g.rdy, dz = d * m.rdzc[k + g.ng];
      atomicAdd(&mfxb[c + 1], dx);
      atomicAdd(&mfxb[c], -dx);
      atomicAdd(&mfyb[c + g.NX], dy);
      atomicAdd(&mfyb[c], -dy);
      if (k + 1 < g.nz) atomicAdd(&mfzb[c + g.plane], dz);
      if (k >= 1) atomicAdd(&mfzb[c], -dz);
    }
  }
}

// Fluxes -> velocities, after k_ad_massflux has gathered every divergence
// contribution (one thread per flux point). The vertical flux
// rho_w (w - (u^xz h_x + v^yz h_y)(1 - zeta_w / zt)) scatters onto the four
// u and four v faces around the w point, so u and v accumulate atomically.
__global__ void k_ad_massflux_uvw(Grid g, Base<double> b, Metric<double> m, double* ub,
                                  double* vb, double* wb, const double* mfxb,
                                  const double* mfyb, const double* mfzb) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 2;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) - 2;
  const int k = blockIdx.z;
  if (i > g.nx + 2 || j > g.ny + 2 || k >= g.nz) return;
  const int c = g.id(i, j, k);
  const int c2 = g.id2(i, j);
  const double rj = b.rhob[c] * m.jac[c2];
  const double rjx = 0.5 * (b.rhob[c - 1] * m.jac[c2 - 1] + rj);
  const double rjy = 0.5 * (b.rhob[c - g.NX] * m.jac[c2 - g.NX] + rj);
  atomicAdd(&ub[c], rjx * mfxb[c]);
  atomicAdd(&vb[c], rjy * mfyb[c]);
  if (k >= 1 && i <= g.nx + 1 && j <= g.ny + 1) {
    const double fb = mfzb[c];
    wb[c] += b.rhobw[c] * fb;
    const double hx = m.hx[c2], hy = m.hy[c2];
    if (fb != 0.0 && (hx != 0.0 || hy != 0.0)) {
      const double cb = -0.25 * b.rhobw[c] * fb * (1.0 - m.zeta_w[k + g.ng] * m.rzt);
      const int cm = c - g.plane;
      atomicAdd(&ub[cm], cb * hx);
      atomicAdd(&ub[cm + 1], cb * hx);
      atomicAdd(&ub[c], cb * hx);
      atomicAdd(&ub[c + 1], cb * hx);
      atomicAdd(&vb[cm], cb * hy);
      atomicAdd(&vb[cm + g.NX], cb * hy);
      atomicAdd(&vb[c], cb * hy);
      atomicAdd(&vb[c + g.NX], cb * hy);
    }
  }
}

// ------------------------------------------------------------ tendencies
template <class S>
struct TanTraj {  // trajectory of one stage (storage type of the forward)
  Traj<S> u, v, w, th, mfx, mfy, mfz, div;
  Traj<S> q;  // the stage estimate's q_v (moist)
};
struct TanPert {  // perturbation (TL) or adjoint (AD) arrays of one stage
  double *u, *v, *w, *th, *mfx, *mfy, *mfz, *div;
  double* q;
  double* pip;
};
struct TanTend {  // tendency perturbations (TL out) or adjoints (AD in)
  double *tu, *tv, *tw, *tth;
  // moist: q_v at time n (TL in / AD accumulated; ghost cells filled) and
  // the stage's q_v = q_v,n + dts T_q (TL out / AD in)
  double *qn, *qout;
  double dts;
  int moist;
  double K;     // diffusion coefficient (0: off)
  double* tpi;  // pi' tendency (relaxation only; rim columns)
  int pd;       // q_v transported by the PD kernels below (k_*_tend skip it)
  // vertical-Courant damping of w (k_tendencies): where C = |w| dt / (dzeta J)
  // exceeds 1, T_w gets -0.3 (C - 1) w / dt; its slope -0.3 (2 C - 1) / dt
  double dtl;   // the large step dt
  int wdamp;
};

__device__ __forceinline__ double wdamp_slope(double wk, double a, double dtl) {
  const double cr = fabs(wk) * a;  // a = dt / (dzeta_w J)
  return cr > 1.0 ? -0.3 * (2.0 * cr - 1.0) / dtl : 0.0;
}

// Positive-definite q_v transport (k_pd_ratio / k_pd_apply,
// kernels.cu:187-264): face fluxes F = mf iface5(q, mf); the limiter ratio
// R = min(1, rt max(0, q) / (dt out)) with out the sum of the outgoing
// fluxes max(0, +-F)/h of the cell; the renormalised flux F R(donor cell,
// upwind of F); T_q = (-div(F R) + q div(mf)) / rt and the stage q =
// q_n + dts T_q. Every branch (flux sign, outgoing or not, q > 0, R < 1,
// out above its floor) is frozen at the trajectory; R's branch is read from
// the recorded ratio field.
template <class S>
struct PDLin {
  const Grid& g;
  const TanTraj<S>& X;
  __device__ int stride(int dir) const { return dir == 0 ? 1 : (dir == 1 ? g.NX : g.plane); }
  __device__ double mf(int dir, int c) const {
    return dir == 0 ? X.mfx[c] : (dir == 1 ? X.mfy[c] : X.mfz[c]);
  }
  __device__ double flux(int dir, int c) const {
    const double f = mf(dir, c);
    return f * iface_at(X.q, c, stride(dir), f);
  }
  __device__ double dflux(const TanPert& D, int dir, int c) const {
    const double* dm = dir == 0 ? D.mfx : (dir == 1 ? D.mfy : D.mfz);
    return tl_face(X.q, D.q, c, stride(dir), mf(dir, c), dm[c]);
  }
  __device__ void ad_flux(const TanPert& A, int dir, int c, double Fb) const {
    double* mb = dir == 0 ? A.mfx : (dir == 1 ? A.mfy : A.mfz);
    atomicAdd(&mb[c], ad_face(X.q, A.q, c, stride(dir), mf(dir, c), Fb));
  }
};

// The recorded ratio of a donor cell (ii, jj, kz): lateral map and vertical
// clamp of the ratio field, as k_pd_apply reads it.
__device__ __forceinline__ int pd_ridx(const Grid& g, const Params& p, int ii, int jj, int kz) {
  ii = map_lat(ii, g.nx, g.nx - 1, p.bcx);
  jj = map_lat(jj, g.ny, g.ny - 1, p.bcy);
  kz = kz < 0 ? 0 : (kz > g.nz - 1 ? g.nz - 1 : kz);
  return g.id(ii, jj, kz);
}

template <class S>
__global__ void k_tl_pd_ratio(Grid g, Base<double> b, Metric<double> m, double dt,
                              TanTraj<S> X, TanPert D, Traj<S> R, double* dR) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const int c = g.id(i, j, k);
  const PDLin<S> P{g, X};
  const double h[3] = {g.dx, g.dy, m.dzeta_c[k + g.ng]};
  double out = 0.0, dout = 0.0;
  for (int dir = 0; dir < 3; ++dir) {
    const int st = P.stride(dir);
    const double Fh = P.flux(dir, c + st), Fl = P.flux(dir, c);
    if (Fh > 0.0) {
      out += Fh / h[dir];
      dout += P.dflux(D, dir, c + st) / h[dir];
    }
    if (Fl < 0.0) {
      out -= Fl / h[dir];
      dout -= P.dflux(D, dir, c) / h[dir];
    }
  }
  double d = 0.0;
  if (dt * out > 1e-20 && R[c] < 1.0) {
    const double rt = b.rhob[c] * m.jac[g.id2(i, j)];
    const double q = X.q[c];
    const double avail = rt * (q > 0.0 ? q : 0.0), davail = q > 0.0 ? rt * D.q[c] : 0.0;
    const double r = avail / (dt * out);
    d = (davail - r * dt * dout) / (dt * out);
  }
  dR[c] = d;
}

template <class S>
__global__ void k_tl_pd_apply(Grid g, Params p, Base<double> b, Metric<double> m, TanTraj<S> X,
                              TanPert D, Traj<S> R, const double* dR, TanTend out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const int c = g.id(i, j, k);
  const PDLin<S> P{g, X};
  const double rd[3] = {g.rdx, g.rdy, m.rdzc[k + g.ng]};
  double dfdiv = 0.0;
  for (int dir = 0; dir < 3; ++dir) {
    for (int side = 0; side < 2; ++side) {  // west/south/bottom face, then the opposite
      const int fi = i + (dir == 0 ? side : 0), fj = j + (dir == 1 ? side : 0),
                fk = k + (dir == 2 ? side : 0);
      const int cf = g.id(fi, fj, fk);
      const double F = P.flux(dir, cf);
      const int di = fi - (dir == 0 && F >= 0.0), dj = fj - (dir == 1 && F >= 0.0),
                dk = fk - (dir == 2 && F >= 0.0);
      const int r = pd_ridx(g, p, di, dj, dk);
      const double ds = P.dflux(D, dir, cf) * R[r] + F * dR[r];
      dfdiv += (side ? ds : -ds) * rd[dir];
    }
  }
  const double rt = b.rhob[c] * m.jac[g.id2(i, j)];
  const double dtq = (-dfdiv + D.q[c] * X.div[c] + X.q[c] * D.div[c]) / rt;
  out.qout[c] = out.qn[c] + out.dts * dtq;
}

template <class S>
__global__ void k_ad_pd_apply(Grid g, Params p, Base<double> b, Metric<double> m, TanTraj<S> X,
                              TanPert A, Traj<S> R, double* Rb, TanTend in) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const int c = g.id(i, j, k);
  const double Qb = in.qout[c];
  if (Qb == 0.0) return;
  atomicAdd(&in.qn[c], Qb);
  const PDLin<S> P{g, X};
  const double rt = b.rhob[c] * m.jac[g.id2(i, j)];
  const double a = in.dts * Qb / rt;
  atomicAdd(&A.q[c], a * X.div[c]);
  atomicAdd(&A.div[c], a * X.q[c]);
  const double rd[3] = {g.rdx, g.rdy, m.rdzc[k + g.ng]};
  for (int dir = 0; dir < 3; ++dir) {
    for (int side = 0; side < 2; ++side) {
      const int fi = i + (dir == 0 ? side : 0), fj = j + (dir == 1 ? side : 0),
                fk = k + (dir == 2 ? side : 0);
      const int cf = g.id(fi, fj, fk);
      const double F = P.flux(dir, cf);
      const int di = fi - (dir == 0 && F >= 0.0), dj = fj - (dir == 1 && F >= 0.0),
                dk = fk - (dir == 2 && F >= 0.0);
      const int r = pd_ridx(g, p, di, dj, dk);
      const double sb = (side ? -a : a) * rd[dir];  // fdiv enters T_q with a minus sign
      atomicAdd(&Rb[r], sb * F);
      P.ad_flux(A, dir, cf, sb * R[r]);
    }
  }
}

template <class S>
__global__ void k_ad_pd_ratio(Grid g, Base<double> b, Metric<double> m, double dt,
                              TanTraj<S> X, TanPert A, Traj<S> R, const double* Rb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const int c = g.id(i, j, k);
  const double rb = Rb[c];
  if (rb == 0.0 || !(R[c] < 1.0)) return;
  const PDLin<S> P{g, X};
  const double h[3] = {g.dx, g.dy, m.dzeta_c[k + g.ng]};
  double out = 0.0;
  for (int dir = 0; dir < 3; ++dir) {
    const int st = P.stride(dir);
    const double Fh = P.flux(dir, c + st), Fl = P.flux(dir, c);
    if (Fh > 0.0) out += Fh / h[dir];
    if (Fl < 0.0) out -= Fl / h[dir];
  }
  if (!(dt * out > 1e-20)) return;
  const double rt = b.rhob[c] * m.jac[g.id2(i, j)];
  const double q = X.q[c];
  const double avail = rt * (q > 0.0 ? q : 0.0);
  const double r = avail / (dt * out);
  // R = avail / (dt out): avail_b = rb / (dt out), out_b = -rb r / out
  if (q > 0.0) atomicAdd(&A.q[c], rt * rb / (dt * out));
  const double ob = -rb * r / out;
  for (int dir = 0; dir < 3; ++dir) {
    const int st = P.stride(dir);
    if (P.flux(dir, c + st) > 0.0) P.ad_flux(A, dir, c + st, ob / h[dir]);
    if (P.flux(dir, c) < 0.0) P.ad_flux(A, dir, c, -ob / h[dir]);
  }
}

// Davies relaxation (k_bdy_relax, kernels.cu:292-304) of the stage
// estimate toward fixed targets: T_f += w (target - f) for u, v, theta' and
// (through the stage update q = q_n + dts T_q) q_v, T_pi' = w (target - pi')
// where w > 0. TL: -w df; AD: the transposes, one cell per thread.
template <class S>
__global__ void k_tl_relax(Grid g, Traj<S> wgt, TanPert D, TanTend out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const double w = wgt[g.id2(i, j)];
  if (w <= 0.0) return;
  const int c = g.id(i, j, k);
  out.tu[c] -= w * D.u[c];
  out.tv[c] -= w * D.v[c];
  out.tth[c] -= w * D.th[c];
  out.tpi[c] = -w * D.pip[c];
  if (out.moist) out.qout[c] -= out.dts * w * D.q[c];
}

template <class S>
__global__ void k_ad_relax(Grid g, Traj<S> wgt, TanPert A, TanTend in) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const double w = wgt[g.id2(i, j)];
  if (w <= 0.0) return;
  const int c = g.id(i, j, k);
  A.u[c] -= w * in.tu[c];
  A.v[c] -= w * in.tv[c];
  A.th[c] -= w * in.tth[c];
  A.pip[c] -= w * in.tpi[c];
  if (in.moist) A.q[c] -= in.dts * w * in.qout[c];
}

// K times the 7-point Laplacian of the perturbation f at c (k_diffuse,
// flat terrain: 1/(dzeta J)^2 = rdz^2), and its transpose: T scattered onto
// the seven points.
__device__ __forceinline__ double tl_lap(const Grid& g, const double* f, int c, double rdz2) {
  const double cc = f[c];
  return (f[c + 1] - 2.0 * cc + f[c - 1]) * (g.rdx * g.rdx) +
         (f[c + g.NX] - 2.0 * cc + f[c - g.NX]) * (g.rdy * g.rdy) +
         (f[c + g.plane] - 2.0 * cc + f[c - g.plane]) * rdz2;
}
__device__ __forceinline__ void ad_lap(const Grid& g, double* fb, int c, double rdz2, double T) {
  const double ax = T * (g.rdx * g.rdx), ay = T * (g.rdy * g.rdy), az = T * rdz2;
  atomicAdd(&fb[c], -2.0 * (ax + ay + az));
  atomicAdd(&fb[c + 1], ax);
  atomicAdd(&fb[c - 1], ax);
  atomicAdd(&fb[c + g.NX], ay);
  atomicAdd(&fb[c - g.NX], ay);
  atomicAdd(&fb[c + g.plane], az);
  atomicAdd(&fb[c - g.plane], az);
}

template <class S>
__global__ void k_tl_tend(Grid g, Base<double> b, Metric<double> m, TanTraj<S> X, TanPert D,
                          TanTend out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const int kk = k + g.ng;
  const int c = g.id(i, j, k);
  const int c2 = g.id2(i, j);
  const int NX = g.NX, PL = g.plane;
  const double dx = g.dx, dy = g.dy;
  const double dzc = m.dzeta_c[kk];
  const double rt = b.rhob[c] * m.jac[c2];
  const double rj2 = 1.0 / (m.jac[c2] * m.jac[c2]);  // diffusion: 1 / (dzeta J)^2
  {  // u
    const double fcL = 0.5 * (X.mfx[c - 1] + X.mfx[c]), dfcL = 0.5 * (D.mfx[c - 1] + D.mfx[c]);
    const double fcR = 0.5 * (X.mfx[c] + X.mfx[c + 1]), dfcR = 0.5 * (D.mfx[c] + D.mfx[c + 1]);
    const double fyB = 0.5 * (X.mfy[c - 1] + X.mfy[c]), dfyB = 0.5 * (D.mfy[c - 1] + D.mfy[c]);
    const double fyT = 0.5 * (X.mfy[c - 1 + NX] + X.mfy[c + NX]),
                 dfyT = 0.5 * (D.mfy[c - 1 + NX] + D.mfy[c + NX]);
    const double fzB = 0.5 * (X.mfz[c - 1] + X.mfz[c]), dfzB = 0.5 * (D.mfz[c - 1] + D.mfz[c]);
    const double fzT = 0.5 * (X.mfz[c - 1 + PL] + X.mfz[c + PL]),
                 dfzT = 0.5 * (D.mfz[c - 1 + PL] + D.mfz[c + PL]);
    const double dfd = (tl_face(X.u, D.u, c + 1, 1, fcR, dfcR) - tl_face(X.u, D.u, c, 1, fcL, dfcL)) / dx +
                       (tl_face(X.u, D.u, c + NX, NX, fyT, dfyT) - tl_face(X.u, D.u, c, NX, fyB, dfyB)) / dy +
                       (tl_face(X.u, D.u, c + PL, PL, fzT, dfzT) - tl_face(X.u, D.u, c, PL, fzB, dfzB)) / dzc;
    const double dv = 0.5 * (X.div[c - 1] + X.div[c]), ddv = 0.5 * (D.div[c - 1] + D.div[c]);
    const double ru = 0.5 * (b.rhob[c - 1] * m.jac[c2 - 1] + rt);
    const double fc = 0.5 * (m.fcor[c2 - 1] + m.fcor[c2]);
    const double dvavg = 0.25 * (D.v[c - 1] + D.v[c - 1 + NX] + D.v[c] + D.v[c + NX]);
    double t = (-dfd + D.u[c] * dv + X.u[c] * ddv) / ru + fc * dvavg - m.ray_c[kk] * D.u[c];
    if (out.K > 0.0) t += out.K * tl_lap(g, D.u, c, m.rdzc[kk] * m.rdzc[kk] * rj2);
    out.tu[c] = t;
  }
  {  // v
    const double fcB = 0.5 * (X.mfy[c - NX] + X.mfy[c]), dfcB = 0.5 * (D.mfy[c - NX] + D.mfy[c]);
    const double fcT = 0.5 * (X.mfy[c] + X.mfy[c + NX]), dfcT = 0.5 * (D.mfy[c] + D.mfy[c + NX]);
    const double fxL = 0.5 * (X.mfx[c - NX] + X.mfx[c]), dfxL = 0.5 * (D.mfx[c - NX] + D.mfx[c]);
    const double fxR = 0.5 * (X.mfx[c + 1 - NX] + X.mfx[c + 1]),
                 dfxR = 0.5 * (D.mfx[c + 1 - NX] + D.mfx[c + 1]);
    const double fzB = 0.5 * (X.mfz[c - NX] + X.mfz[c]), dfzB = 0.5 * (D.mfz[c - NX] + D.mfz[c]);
    const double fzT = 0.5 * (X.mfz[c - NX + PL] + X.mfz[c + PL]),
                 dfzT = 0.5 * (D.mfz[c - NX + PL] + D.mfz[c + PL]);
    const double dfd = (tl_face(X.v, D.v, c + 1, 1, fxR, dfxR) - tl_face(X.v, D.v, c, 1, fxL, dfxL)) / dx +
                       (tl_face(X.v, D.v, c + NX, NX, fcT, dfcT) - tl_face(X.v, D.v, c, NX, fcB, dfcB)) / dy +
                       (tl_face(X.v, D.v, c + PL, PL, fzT, dfzT) - tl_face(X.v, D.v, c, PL, fzB, dfzB)) / dzc;
    const double dv = 0.5 * (X.div[c - NX] + X.div[c]), ddv = 0.5 * (D.div[c - NX] + D.div[c]);
    const double rv = 0.5 * (b.rhob[c - NX] * m.jac[c2 - NX] + rt);
    const double fc = 0.5 * (m.fcor[c2 - NX] + m.fcor[c2]);
    const double duavg = 0.25 * (D.u[c - NX] + D.u[c + 1 - NX] + D.u[c] + D.u[c + 1]);
    double t = (-dfd + D.v[c] * dv + X.v[c] * ddv) / rv - fc * duavg - m.ray_c[kk] * D.v[c];
    if (out.K > 0.0) t += out.K * tl_lap(g, D.v, c, m.rdzc[kk] * m.rdzc[kk] * rj2);
    out.tv[c] = t;
  }
  if (k >= 1) {  // w
    const double dzw = m.dzeta_w[kk];
    const double fxL = 0.5 * (X.mfx[c - PL] + X.mfx[c]), dfxL = 0.5 * (D.mfx[c - PL] + D.mfx[c]);
    const double fxR = 0.5 * (X.mfx[c + 1 - PL] + X.mfx[c + 1]),
                 dfxR = 0.5 * (D.mfx[c + 1 - PL] + D.mfx[c + 1]);
    const double fyB = 0.5 * (X.mfy[c - PL] + X.mfy[c]), dfyB = 0.5 * (D.mfy[c - PL] + D.mfy[c]);
    const double fyT = 0.5 * (X.mfy[c + NX - PL] + X.mfy[c + NX]),
                 dfyT = 0.5 * (D.mfy[c + NX - PL] + D.mfy[c + NX]);
    const double fzB = 0.5 * (X.mfz[c - PL] + X.mfz[c]), dfzB = 0.5 * (D.mfz[c - PL] + D.mfz[c]);
    const double fzT = 0.5 * (X.mfz[c] + X.mfz[c + PL]), dfzT = 0.5 * (D.mfz[c] + D.mfz[c + PL]);
    const double dfd = (tl_face(X.w, D.w, c + 1, 1, fxR, dfxR) - tl_face(X.w, D.w, c, 1, fxL, dfxL)) / dx +
                       (tl_face(X.w, D.w, c + NX, NX, fyT, dfyT) - tl_face(X.w, D.w, c, NX, fyB, dfyB)) / dy +
                       (tl_face(X.w, D.w, c + PL, PL, fzT, dfzT) - tl_face(X.w, D.w, c, PL, fzB, dfzB)) / dzw;
    const double dv = 0.5 * (X.div[c - PL] + X.div[c]), ddv = 0.5 * (D.div[c - PL] + D.div[c]);
    const double rw = b.rhobw[c] * m.jac[c2];
    double t = (-dfd + D.w[c] * dv + X.w[c] * ddv) / rw - m.ray_w[kk] * D.w[c];
    if (out.wdamp) t += wdamp_slope(X.w[c], out.dtl * m.rdzw[kk] / m.jac[c2], out.dtl) * D.w[c];
    if (out.K > 0.0) t += out.K * tl_lap(g, D.w, c, m.rdzw[kk] * m.rdzw[kk] * rj2);
    // moist buoyancy g 0.61 (q_v at the w level - base) from time-n q_v
    if (out.moist) t += kGrav * kEps61 * 0.5 * (out.qn[c - PL] + out.qn[c]);
    out.tw[c] = t;
  }
  {  // theta'
    const double dfd =
        (tl_face(X.th, D.th, c + 1, 1, X.mfx[c + 1], D.mfx[c + 1]) -
         tl_face(X.th, D.th, c, 1, X.mfx[c], D.mfx[c])) / dx +
        (tl_face(X.th, D.th, c + NX, NX, X.mfy[c + NX], D.mfy[c + NX]) -
         tl_face(X.th, D.th, c, NX, X.mfy[c], D.mfy[c])) / dy +
        (tl_face(X.th, D.th, c + PL, PL, X.mfz[c + PL], D.mfz[c + PL]) -
         tl_face(X.th, D.th, c, PL, X.mfz[c], D.mfz[c])) / dzc;
    double t = (-dfd + D.th[c] * X.div[c] + X.th[c] * D.div[c]) / rt - m.ray_c[kk] * D.th[c];
    if (out.K > 0.0) t += out.K * tl_lap(g, D.th, c, m.rdzc[kk] * m.rdzc[kk] * rj2);
    out.tth[c] = t;
  }
  if (out.moist && !out.pd) {  // q_v: flux-form transport, then the stage update
    const double dfd =
        (tl_face(X.q, D.q, c + 1, 1, X.mfx[c + 1], D.mfx[c + 1]) -
         tl_face(X.q, D.q, c, 1, X.mfx[c], D.mfx[c])) / dx +
        (tl_face(X.q, D.q, c + NX, NX, X.mfy[c + NX], D.mfy[c + NX]) -
         tl_face(X.q, D.q, c, NX, X.mfy[c], D.mfy[c])) / dy +
        (tl_face(X.q, D.q, c + PL, PL, X.mfz[c + PL], D.mfz[c + PL]) -
         tl_face(X.q, D.q, c, PL, X.mfz[c], D.mfz[c])) / dzc;
    const double tq = (-dfd + D.q[c] * X.div[c] + X.q[c] * D.div[c]) / rt;
    out.qout[c] = out.qn[c] + out.dts * tq;
  }
}

template <class S>
__global__ void k_ad_tend(Grid g, Base<double> b, Metric<double> m, TanTraj<S> X, TanPert A,
                          TanTend in) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const int kk = k + g.ng;
  const int c = g.id(i, j, k);
  const int c2 = g.id2(i, j);
  const int NX = g.NX, PL = g.plane;
  const double dx = g.dx, dy = g.dy;
  const double dzc = m.dzeta_c[kk];
  const double rt = b.rhob[c] * m.jac[c2];
  const double rj2 = 1.0 / (m.jac[c2] * m.jac[c2]);
  // Adjoint of a face-averaged flux f = 0.5 (mf[a] + mf[b]).
  auto avg_b = [](double* mfb, int a, int bb, double fb) {
    atomicAdd(&mfb[a], 0.5 * fb);
    atomicAdd(&mfb[bb], 0.5 * fb);
  };
  {  // u
    const double T = in.tu[c];
    if (T != 0.0) {
      const double ru = 0.5 * (b.rhob[c - 1] * m.jac[c2 - 1] + rt);
      const double cb = 0.25 * 0.5 * (m.fcor[c2 - 1] + m.fcor[c2]) * T;  // Coriolis
      atomicAdd(&A.v[c - 1], cb);
      atomicAdd(&A.v[c - 1 + NX], cb);
      atomicAdd(&A.v[c], cb);
      atomicAdd(&A.v[c + NX], cb);
      if (in.K > 0.0) ad_lap(g, A.u, c, m.rdzc[kk] * m.rdzc[kk] * rj2, in.K * T);
      const double a = T / ru;
      const double fdb = -a;
      const double dv = 0.5 * (X.div[c - 1] + X.div[c]);
      atomicAdd(&A.u[c], a * dv - m.ray_c[kk] * T);
      const double dvb = a * X.u[c];
      atomicAdd(&A.div[c - 1], 0.5 * dvb);
      atomicAdd(&A.div[c], 0.5 * dvb);
      const double fcL = 0.5 * (X.mfx[c - 1] + X.mfx[c]);
      const double fcR = 0.5 * (X.mfx[c] + X.mfx[c + 1]);
      const double fyB = 0.5 * (X.mfy[c - 1] + X.mfy[c]);
      const double fyT = 0.5 * (X.mfy[c - 1 + NX] + X.mfy[c + NX]);
      const double fzB = 0.5 * (X.mfz[c - 1] + X.mfz[c]);
      const double fzT = 0.5 * (X.mfz[c - 1 + PL] + X.mfz[c + PL]);
      avg_b(A.mfx, c, c + 1, ad_face(X.u, A.u, c + 1, 1, fcR, fdb / dx));
      avg_b(A.mfx, c - 1, c, ad_face(X.u, A.u, c, 1, fcL, -fdb / dx));
      avg_b(A.mfy, c - 1 + NX, c + NX, ad_face(X.u, A.u, c + NX, NX, fyT, fdb / dy));
      avg_b(A.mfy, c - 1, c, ad_face(X.u, A.u, c, NX, fyB, -fdb / dy));
      avg_b(A.mfz, c - 1 + PL, c + PL, ad_face(X.u, A.u, c + PL, PL, fzT, fdb / dzc));
      avg_b(A.mfz, c - 1, c, ad_face(X.u, A.u, c, PL, fzB, -fdb / dzc));
    }
  }
  {  // v
    const double T = in.tv[c];
    if (T != 0.0) {
      const double rv = 0.5 * (b.rhob[c - NX] * m.jac[c2 - NX] + rt);
      const double cb = -0.25 * 0.5 * (m.fcor[c2 - NX] + m.fcor[c2]) * T;  // Coriolis
      atomicAdd(&A.u[c - NX], cb);
      atomicAdd(&A.u[c + 1 - NX], cb);
      atomicAdd(&A.u[c], cb);
      atomicAdd(&A.u[c + 1], cb);
      if (in.K > 0.0) ad_lap(g, A.v, c, m.rdzc[kk] * m.rdzc[kk] * rj2, in.K * T);
      const double a = T / rv;
      const double fdb = -a;
      const double dv = 0.5 * (X.div[c - NX] + X.div[c]);
      atomicAdd(&A.v[c], a * dv - m.ray_c[kk] * T);
      const double dvb = a * X.v[c];
      atomicAdd(&A.div[c - NX], 0.5 * dvb);
      atomicAdd(&A.div[c], 0.5 * dvb);
      const double fcB = 0.5 * (X.mfy[c - NX] + X.mfy[c]);
      const double fcT = 0.5 * (X.mfy[c] + X.mfy[c + NX]);
      const double fxL = 0.5 * (X.mfx[c - NX] + X.mfx[c]);
      const double fxR = 0.5 * (X.mfx[c + 1 - NX] + X.mfx[c + 1]);
      const double fzB = 0.5 * (X.mfz[c - NX] + X.mfz[c]);
      const double fzT = 0.5 * (X.mfz[c - NX + PL] + X.mfz[c + PL]);
      avg_b(A.mfx, c + 1 - NX, c + 1, ad_face(X.v, A.v, c + 1, 1, fxR, fdb / dx));
      avg_b(A.mfx, c - NX, c, ad_face(X.v, A.v, c, 1, fxL, -fdb / dx));
      avg_b(A.mfy, c, c + NX, ad_face(X.v, A.v, c + NX, NX, fcT, fdb / dy));
      avg_b(A.mfy, c - NX, c, ad_face(X.v, A.v, c, NX, fcB, -fdb / dy));
      avg_b(A.mfz, c - NX + PL, c + PL, ad_face(X.v, A.v, c + PL, PL, fzT, fdb / dzc));
      avg_b(A.mfz, c - NX, c, ad_face(X.v, A.v, c, PL, fzB, -fdb / dzc));
    }
  }
  if (k >= 1) {  // w
    const double T = in.tw[c];
    if (T != 0.0) {
      const double dzw = m.dzeta_w[kk];
      const double rw = b.rhobw[c] * m.jac[c2];
      if (in.K > 0.0) ad_lap(g, A.w, c, m.rdzw[kk] * m.rdzw[kk] * rj2, in.K * T);
      if (in.wdamp)
        atomicAdd(&A.w[c], wdamp_slope(X.w[c], in.dtl * m.rdzw[kk] / m.jac[c2], in.dtl) * T);
      if (in.moist) {
        const double qb = kGrav * kEps61 * 0.5 * T;
        atomicAdd(&in.qn[c - PL], qb);
        atomicAdd(&in.qn[c], qb);
      }
      const double a = T / rw;
      const double fdb = -a;
      const double dv = 0.5 * (X.div[c - PL] + X.div[c]);
      atomicAdd(&A.w[c], a * dv - m.ray_w[kk] * T);
      const double dvb = a * X.w[c];
      atomicAdd(&A.div[c - PL], 0.5 * dvb);
      atomicAdd(&A.div[c], 0.5 * dvb);
      const double fxL = 0.5 * (X.mfx[c - PL] + X.mfx[c]);
      const double fxR = 0.5 * (X.mfx[c + 1 - PL] + X.mfx[c + 1]);
      const double fyB = 0.5 * (X.mfy[c - PL] + X.mfy[c]);
      const double fyT = 0.5 * (X.mfy[c + NX - PL] + X.mfy[c + NX]);
      const double fzB = 0.5 * (X.mfz[c - PL] + X.mfz[c]);
      const double fzT = 0.5 * (X.mfz[c] + X.mfz[c + PL]);
      avg_b(A.mfx, c + 1 - PL, c + 1, ad_face(X.w, A.w, c + 1, 1, fxR, fdb / dx));
      avg_b(A.mfx, c - PL, c, ad_face(X.w, A.w, c, 1, fxL, -fdb / dx));
      avg_b(A.mfy, c + NX - PL, c + NX, ad_face(X.w, A.w, c + NX, NX, fyT, fdb / dy));
      avg_b(A.mfy, c - PL, c, ad_face(X.w, A.w, c, NX, fyB, -fdb / dy));
      avg_b(A.mfz, c, c + PL, ad_face(X.w, A.w, c + PL, PL, fzT, fdb / dzw));
      avg_b(A.mfz, c - PL, c, ad_face(X.w, A.w, c, PL, fzB, -fdb / dzw));
    }
  }
  {  // theta'
    const double T = in.tth[c];
    if (T != 0.0) {
      if (in.K > 0.0) ad_lap(g, A.th, c, m.rdzc[kk] * m.rdzc[kk] * rj2, in.K * T);
      const double a = T / rt;
      const double fdb = -a;
      atomicAdd(&A.th[c], a * X.div[c] - m.ray_c[kk] * T);
      atomicAdd(&A.div[c], a * X.th[c]);
      atomicAdd(&A.mfx[c + 1], ad_face(X.th, A.th, c + 1, 1, X.mfx[c + 1], fdb / dx));
      atomicAdd(&A.mfx[c], ad_face(X.th, A.th, c, 1, X.mfx[c], -fdb / dx));
      atomicAdd(&A.mfy[c + NX], ad_face(X.th, A.th, c + NX, NX, X.mfy[c + NX], fdb / dy));
      atomicAdd(&A.mfy[c], ad_face(X.th, A.th, c, NX, X.mfy[c], -fdb / dy));
      atomicAdd(&A.mfz[c + PL], ad_face(X.th, A.th, c + PL, PL, X.mfz[c + PL], fdb / dzc));
      atomicAdd(&A.mfz[c], ad_face(X.th, A.th, c, PL, X.mfz[c], -fdb / dzc));
    }
  }
  if (in.moist && !in.pd) {  // q_v stage update and transport
    const double Qb = in.qout[c];
    if (Qb != 0.0) {
      atomicAdd(&in.qn[c], Qb);
      const double a = in.dts * Qb / rt;
      const double fdb = -a;
      atomicAdd(&A.q[c], a * X.div[c]);
      atomicAdd(&A.div[c], a * X.q[c]);
      atomicAdd(&A.mfx[c + 1], ad_face(X.q, A.q, c + 1, 1, X.mfx[c + 1], fdb / dx));
      atomicAdd(&A.mfx[c], ad_face(X.q, A.q, c, 1, X.mfx[c], -fdb / dx));
      atomicAdd(&A.mfy[c + NX], ad_face(X.q, A.q, c + NX, NX, X.mfy[c + NX], fdb / dy));
      atomicAdd(&A.mfy[c], ad_face(X.q, A.q, c, NX, X.mfy[c], -fdb / dy));
      atomicAdd(&A.mfz[c + PL], ad_face(X.q, A.q, c + PL, PL, X.mfz[c + PL], fdb / dzc));
      atomicAdd(&A.mfz[c], ad_face(X.q, A.q, c, PL, X.mfz[c], -fdb / dzc));
    }
  }
}

// ------------------------------------------------- horizontal acoustic step
// u_new = u + dtau (-cp thv_u m_u grad pi* + Tu) on faces i = 0..nx (with
// periodic x, u(nx) is the image of u(0)), the same for v. pi* = pi' +
// smdiv (pi' - pi'prev) with the substep ghost rule: outside the interior,
// pi' is read from the mapped interior of the current substep and pi'prev
// from the time-n ghost cell; theta' outside the interior from the time-n
// ghost cell. Over terrain the gradient has the term -(z_x / J) dpi*/dzeta.
// On open edges the boundary face takes the Klemp-Wilhelmson radiation
// value (k_bc_rad_u_x / k_bc_rad_v_y, kernels.cu:649-688): the base wind on
// inflow; on outflow u - dtau (u - c*) (u_step(1) - u) / dx at i = 0 and
// u - dtau (u + c*) (u - u_step(nx-1)) / dx at i = nx. The inflow/outflow
// branch is frozen at the trajectory's edge velocity.
template <class S>
struct UVTraj {
  const double *pcur, *pprev, *pn;
  Traj<S> thc, thn;
  Traj<S> qn;          // time-n q_v, ghost cells filled (read when p.moisture)
  Traj<S> us, vs, tu, tv;  // u, v entering the substep and their tendencies
};
struct UVTan {  // TL: perturbations in, du/dv out; AD: adjoints
  double *u_src, *v_src, *u_dst, *v_dst, *pcur, *pprev, *pn, *thc, *thn, *tu, *tv;
  double* qn;
};

template <class S>
struct UVLin {
  const Grid& g;
  const Params& p;
  const Base<double>& b;
  const Metric<double>& m;
  const UVTraj<S>& X;
  double dtau;
  __device__ bool inside(int ii, int jj) const {
    return (unsigned)ii < (unsigned)g.nx && (unsigned)jj < (unsigned)g.ny;
  }
  __device__ bool inside3(int ii, int jj, int kz) const {
    return inside(ii, jj) && (unsigned)kz < (unsigned)g.nz;
  }
  // the interior point a stencil point outside reads pi' from (lateral map,
  // vertical clamp), as UVDev::pstar
  __device__ int mapped(int ii, int jj, int kz) const {
    const int km = kz < 0 ? 0 : (kz > g.nz - 1 ? g.nz - 1 : kz);
    return g.id(map_lat(ii, g.nx, g.nx - 1, p.bcx), map_lat(jj, g.ny, g.ny - 1, p.bcy), km);
  }
  __device__ double pstar(int ii, int jj, int k) const {
    const int c = g.id(ii, jj, k);
    if (inside3(ii, jj, k)) return X.pcur[c] + p.smdiv * (X.pcur[c] - X.pprev[c]);
    const double pc = X.pcur[mapped(ii, jj, k)];
    return pc + p.smdiv * (pc - X.pn[c]);
  }
  // theta_v = (theta_bar + theta') (1 + 0.61 q_v,n), the factor when moist
  __device__ double dry(int ii, int jj, int k) const {
    const int c = g.id(ii, jj, k);
    return b.thb[c] + (inside(ii, jj) ? X.thc[c] : X.thn[c]);
  }
  __device__ double fq(int c) const { return p.moisture ? 1.0 + kEps61 * X.qn[c] : 1.0; }
  __device__ double thv(int ii, int jj, int k) const {
    return dry(ii, jj, k) * fq(g.id(ii, jj, k));
  }
  __device__ double dpstar(const UVTan& D, int ii, int jj, int k) const {
    const int c = g.id(ii, jj, k);
    if (inside3(ii, jj, k)) return D.pcur[c] + p.smdiv * (D.pcur[c] - D.pprev[c]);
    const double pc = D.pcur[mapped(ii, jj, k)];
    return pc + p.smdiv * (pc - D.pn[c]);
  }
  __device__ double dthv(const UVTan& D, int ii, int jj, int k) const {
    const int c = g.id(ii, jj, k);
    const double dt = (inside(ii, jj) ? D.thc[c] : D.thn[c]) * fq(c);
    return p.moisture ? dt + dry(ii, jj, k) * kEps61 * D.qn[c] : dt;
  }
  __device__ void ad_pstar(const UVTan& A, int ii, int jj, int k, double pb) const {
    const int c = g.id(ii, jj, k);
    if (inside3(ii, jj, k)) {
      atomicAdd(&A.pcur[c], (1.0 + p.smdiv) * pb);
      atomicAdd(&A.pprev[c], -p.smdiv * pb);
    } else {
      atomicAdd(&A.pcur[mapped(ii, jj, k)], (1.0 + p.smdiv) * pb);
      atomicAdd(&A.pn[c], -p.smdiv * pb);
    }
  }
  __device__ void ad_thv(const UVTan& A, int ii, int jj, int k, double tb) const {
    const int c = g.id(ii, jj, k);
    atomicAdd(inside(ii, jj) ? &A.thc[c] : &A.thn[c], tb * fq(c));
    if (p.moisture) atomicAdd(&A.qn[c], tb * dry(ii, jj, k) * kEps61);
  }
  // Terrain term of the pi* gradient at a u face (dir 0) or v face (dir 1):
  // the coefficient (z_s / J_face) / (2 dzeta_2) of the centred vertical
  // differences of pi* in the two cells beside the face (zero on flat
  // ground).
  __device__ double tcoef(int c2, int k, int dir) const {
    const double h = dir == 0 ? m.hxu[c2] : m.hyv[c2];
    if (h == 0.0) return 0.0;
    const int kk = k + g.ng;
    const double zs = h * (1.0 - m.zeta_c[kk] / m.zt);
    const double dz2 = m.zeta_c[kk + 1] - m.zeta_c[kk - 1];
    const int o = dir == 0 ? 1 : g.NX;
    const double jf = 0.5 * (m.jac[c2 - o] + m.jac[c2]);
    return (zs / jf) * 0.5 / dz2;
  }
  // Face (i, j, k) in direction dir: the cells beside it are (i-1, j) and
  // (i, j) for u, (i, j-1) and (i, j) for v.
  struct Face {
    int c, c2, i0, j0, i1, j1;
    double mf, rd;
  };
  __device__ Face face(int i, int j, int k, int dir) const {
    Face f;
    f.c = g.id(i, j, k);
    f.c2 = g.id2(i, j);
    f.i1 = i;
    f.j1 = j;
    f.i0 = dir == 0 ? i - 1 : i;
    f.j0 = dir == 0 ? j : j - 1;
    const int o = dir == 0 ? 1 : g.NX;
    f.mf = 0.5 * (m.mapf[f.c2 - o] + m.mapf[f.c2]);
    f.rd = dir == 0 ? g.rdx : g.rdy;
    return f;
  }
  __device__ double tdiff(const Face& f, int k) const {
    return (pstar(f.i0, f.j0, k + 1) - pstar(f.i0, f.j0, k - 1)) +
           (pstar(f.i1, f.j1, k + 1) - pstar(f.i1, f.j1, k - 1));
  }
  __device__ double grad(const Face& f, int k, double tc) const {
    double gr = (pstar(f.i1, f.j1, k) - pstar(f.i0, f.j0, k)) * f.rd;
    if (tc != 0.0) gr -= tc * tdiff(f, k);
    return gr;
  }
  // Interior step of the face: trajectory value, TL, AD (adjoint ob of the
  // new value scattered onto the inputs).
  __device__ double step(int i, int j, int k, int dir) const {
    const Face f = face(i, j, k, dir);
    const double th = 0.5 * (thv(f.i0, f.j0, k) + thv(f.i1, f.j1, k));
    const double src = dir == 0 ? X.us[f.c] : X.vs[f.c];
    const double t = dir == 0 ? X.tu[f.c] : X.tv[f.c];
    return src + dtau * (-kCp * th * f.mf * grad(f, k, tcoef(f.c2, k, dir)) + t);
  }
  __device__ double dstep(const UVTan& D, int i, int j, int k, int dir) const {
    const Face f = face(i, j, k, dir);
    const double tc = tcoef(f.c2, k, dir);
    const double th = 0.5 * (thv(f.i0, f.j0, k) + thv(f.i1, f.j1, k));
    const double dth = 0.5 * (dthv(D, f.i0, f.j0, k) + dthv(D, f.i1, f.j1, k));
    double dgr = (dpstar(D, f.i1, f.j1, k) - dpstar(D, f.i0, f.j0, k)) * f.rd;
    if (tc != 0.0)
      dgr -= tc * ((dpstar(D, f.i0, f.j0, k + 1) - dpstar(D, f.i0, f.j0, k - 1)) +
                   (dpstar(D, f.i1, f.j1, k + 1) - dpstar(D, f.i1, f.j1, k - 1)));
    const double src = dir == 0 ? D.u_src[f.c] : D.v_src[f.c];
    const double t = dir == 0 ? D.tu[f.c] : D.tv[f.c];
    return src + dtau * (-kCp * f.mf * (dth * grad(f, k, tc) + th * dgr) + t);
  }
  __device__ void ad_step(const UVTan& A, int i, int j, int k, int dir, double ob) const {
    const Face f = face(i, j, k, dir);
    const double tc = tcoef(f.c2, k, dir);
    const double th = 0.5 * (thv(f.i0, f.j0, k) + thv(f.i1, f.j1, k));
    atomicAdd(dir == 0 ? &A.u_src[f.c] : &A.v_src[f.c], ob);
    atomicAdd(dir == 0 ? &A.tu[f.c] : &A.tv[f.c], dtau * ob);
    const double x = -dtau * kCp * f.mf * ob;
    const double thb = 0.5 * x * grad(f, k, tc);
    ad_thv(A, f.i0, f.j0, k, thb);
    ad_thv(A, f.i1, f.j1, k, thb);
    const double gb = x * th;
    ad_pstar(A, f.i1, f.j1, k, gb * f.rd);
    ad_pstar(A, f.i0, f.j0, k, -gb * f.rd);
    if (tc != 0.0) {
      const double tb = -tc * gb;
      ad_pstar(A, f.i0, f.j0, k + 1, tb);
      ad_pstar(A, f.i0, f.j0, k - 1, -tb);
      ad_pstar(A, f.i1, f.j1, k + 1, tb);
      ad_pstar(A, f.i1, f.j1, k - 1, -tb);
    }
  }
};

// Which formula face n (0..nn) of a direction takes: 0 the interior step,
// 1 the periodic image of face 0, 2 the west/south radiation edge, 3 the
// east/north one.
__device__ __forceinline__ int face_kind(int n, int nn, int open) {
  if (n < nn && !(open && n == 0)) return 0;
  if (!open) return 1;
  return n == 0 ? 2 : 3;
}

template <class S>
__global__ void k_tl_uv(Grid g, Params p, Base<double> b, Metric<double> m, UVTraj<S> X, UVTan D,
                        double dtau) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i > g.nx || j > g.ny || k >= g.nz) return;
  const UVLin<S> L{g, p, b, m, X, dtau};
  const int c = g.id(i, j, k);
  for (int dir = 0; dir < 2; ++dir) {
    if (dir == 0 ? j >= g.ny : i >= g.nx) continue;
    const int n = dir == 0 ? i : j, nn = dir == 0 ? g.nx : g.ny;
    const int open = dir == 0 ? p.bcx : p.bcy;
    const double* src = dir == 0 ? D.u_src : D.v_src;
    double d;
    switch (face_kind(n, nn, open)) {
      case 0: d = L.dstep(D, i, j, k, dir); break;
      case 1: d = dir == 0 ? L.dstep(D, 0, j, k, 0) : L.dstep(D, i, 0, k, 1); break;
      case 2: {  // west/south edge: outflow when the edge velocity is negative
        const double u0 = dir == 0 ? X.us[c] : X.vs[c];
        if (u0 >= 0.0) { d = 0.0; break; }
        const double cb = u0 - p.cstar, h = dir == 0 ? g.dx : g.dy;
        const int i1 = dir == 0 ? 1 : i, j1 = dir == 0 ? j : 1;
        const double U1 = L.step(i1, j1, k, dir), dU1 = L.dstep(D, i1, j1, k, dir);
        d = src[c] - dtau * (src[c] * (U1 - u0) + cb * (dU1 - src[c])) / h;
        break;
      }
      default: {  // east/north edge: outflow when the edge velocity is positive
        const double u1 = dir == 0 ? X.us[c] : X.vs[c];
        if (u1 <= 0.0) { d = 0.0; break; }
        const double cb = u1 + p.cstar, h = dir == 0 ? g.dx : g.dy;
        const int in = dir == 0 ? nn - 1 : i, jn = dir == 0 ? j : nn - 1;
        const double Un = L.step(in, jn, k, dir), dUn = L.dstep(D, in, jn, k, dir);
        d = src[c] - dtau * (src[c] * (u1 - Un) + cb * (src[c] - dUn)) / h;
      }
    }
    (dir == 0 ? D.u_dst : D.v_dst)[c] = d;
  }
}

template <class S>
__global__ void k_ad_uv(Grid g, Params p, Base<double> b, Metric<double> m, UVTraj<S> X, UVTan A,
                        double dtau) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i > g.nx || j > g.ny || k >= g.nz) return;
  const UVLin<S> L{g, p, b, m, X, dtau};
  const int c = g.id(i, j, k);
  for (int dir = 0; dir < 2; ++dir) {
    if (dir == 0 ? j >= g.ny : i >= g.nx) continue;
    const double ob = (dir == 0 ? A.u_dst : A.v_dst)[c];
    if (ob == 0.0) continue;
    const int n = dir == 0 ? i : j, nn = dir == 0 ? g.nx : g.ny;
    const int open = dir == 0 ? p.bcx : p.bcy;
    double* srcb = dir == 0 ? A.u_src : A.v_src;
    switch (face_kind(n, nn, open)) {
      case 0: L.ad_step(A, i, j, k, dir, ob); break;
      case 1:
        if (dir == 0) L.ad_step(A, 0, j, k, 0, ob);
        else L.ad_step(A, i, 0, k, 1, ob);
        break;
      case 2: {
        const double u0 = dir == 0 ? X.us[c] : X.vs[c];
        if (u0 >= 0.0) break;
        const double cb = u0 - p.cstar, h = dir == 0 ? g.dx : g.dy;
        const int i1 = dir == 0 ? 1 : i, j1 = dir == 0 ? j : 1;
        const double U1 = L.step(i1, j1, k, dir);
        // d = src (1 - dtau (U1 - u0) / h + dtau cb / h) - dtau cb / h dU1
        atomicAdd(&srcb[c], ob * (1.0 - dtau * (U1 - u0) / h + dtau * cb / h));
        L.ad_step(A, i1, j1, k, dir, -ob * dtau * cb / h);
        break;
      }
      default: {
        const double u1 = dir == 0 ? X.us[c] : X.vs[c];
        if (u1 <= 0.0) break;
        const double cb = u1 + p.cstar, h = dir == 0 ? g.dx : g.dy;
        const int in = dir == 0 ? nn - 1 : i, jn = dir == 0 ? j : nn - 1;
        const double Un = L.step(in, jn, k, dir);
        // d = src (1 - dtau (u1 - Un) / h - dtau cb / h) + dtau cb / h dUn
        atomicAdd(&srcb[c], ob * (1.0 - dtau * (u1 - Un) / h - dtau * cb / h));
        L.ad_step(A, in, jn, k, dir, ob * dtau * cb / h);
      }
    }
  }
}

// ------------------------------------------------ vertically implicit solve
// One thread per column, levels in local arrays (nz < MAXL). The forward
// column (kernels.cu:531-645 with flat terrain and no moisture) is
// recomputed from the trajectory, then linearised.
template <class S>
struct WPTraj {
  Traj<S> u, v;  // after this substep's horizontal step
  Traj<S> w;
  const double* pip;
  Traj<S> th;
  Traj<S> qn;  // time-n q_v (read when p.moisture)
  // relaxation pi' tendency and the rim weight (tpw.p null: no relaxation);
  // the tendency enters P in columns with a positive weight
  Traj<S> tpi, tpw;
};
struct WPTan {  // TL: perturbations; AD: adjoints
  double *u, *v, *w, *pip, *th, *tw, *tth;  // inputs
  double *w_out, *pip_out, *th_out;         // outputs
  double* qn;                               // time-n q_v (input)
  double* tpi;                              // relaxation pi' tendency (input)
};

// theta_v at the w level below cell c: its dry part and the moist factor.
template <class S>
__device__ __forceinline__ double thvw_dry(const Base<double>& b, const WPTraj<S>& X, int c,
                                           int PL) {
  return b.thbw[c] + 0.5 * (X.th[c - PL] + X.th[c]);
}
template <class S>
__device__ __forceinline__ double thvw_fq(const Params& p, const WPTraj<S>& X, int c, int PL) {
  return p.moisture ? 1.0 + kEps61 * 0.5 * (X.qn[c - PL] + X.qn[c]) : 1.0;
}

template <int MAXL>
struct WPCol {
  double P[MAXL], E[MAXL], G[MAXL], R[MAXL], Av[MAXL + 1], W[MAXL + 1];
  double A[MAXL], B[MAXL], C[MAXL], D[MAXL], cs[MAXL], ds[MAXL];
};

// Cross-metric vertical flux S_v = A_v (u^xz h_x + v^yz h_y)(1 - zeta_w / zt)
// at w level k of a column (zero at k = 0, nz and on flat ground), from the
// horizontal velocities after the substep's horizontal step; linear in
// (u, v), so the same function gives its TL on perturbations.
template <class V>
__device__ __forceinline__ double sv_at(const Grid& g, const Metric<double>& m,
                                        const Coef<double>& cf, const V& u, const V& v, int c0,
                                        int k, double hx, double hy) {
  if (k <= 0 || k >= g.nz || (hx == 0.0 && hy == 0.0)) return 0.0;
  const int PL = g.plane, NX = g.NX, c = c0 + k * PL;
  const double fac = 1.0 - m.zeta_w[k + g.ng] / m.zt;
  const double uw = 0.25 * (u[c - PL] + u[c - PL + 1] + u[c] + u[c + 1]);
  const double vw = 0.25 * (v[c - PL] + v[c - PL + NX] + v[c] + v[c + NX]);
  return cf.aw[c] * (uw * hx + vw * hy) * fac;
}

// Surface w(0) = u^x h_x + v^y h_y of a column (diagnostic; zero on flat
// ground).
template <class V>
__device__ __forceinline__ double w_surface(const Grid& g, const V& u, const V& v, int c0,
                                            double hx, double hy) {
  return 0.5 * (u[c0] + u[c0 + 1]) * hx + 0.5 * (v[c0] + v[c0 + g.NX]) * hy;
}

// Forward column: assembly, rows and the Thomas solve.
template <class S, int MAXL>
__device__ void wp_forward(const Grid& g, const Params& p, const Base<double>& b, const Metric<double>& m,
                           const Coef<double>& cf, const WPTraj<S>& X,
                           int c0, int c2d, double mc, double dtau, WPCol<MAXL>& s) {
  const int NX = g.NX, PL = g.plane, nz = g.nz;
  const double c1 = (1.0 + p.beta) * 0.5, c2 = (1.0 - p.beta) * 0.5;
  const double J = m.jac[c2d], hx = m.hx[c2d], hy = m.hy[c2d];
  for (int k = 0; k <= nz; ++k) s.Av[k] = (k <= 0 || k >= nz) ? 0.0 : (double)cf.aw[c0 + k * PL];
  for (int k = 0; k < nz; ++k) {
    const int c = c0 + k * PL, kk = k + g.ng;
    const double Kt = cf.kt3[c];
    const double Dh = mc * (((double)cf.rtjx[c + 1] * (double)X.u[c + 1] -
                             (double)cf.rtjx[c] * (double)X.u[c]) * g.rdx +
                            ((double)cf.rtjy[c + NX] * (double)X.v[c + NX] -
                             (double)cf.rtjy[c] * (double)X.v[c]) * g.rdy);
    const double rdzc = m.rdzc[kk];
    const double dsv = sv_at(g, m, cf, X.u, X.v, c0, k + 1, hx, hy) -
                       sv_at(g, m, cf, X.u, X.v, c0, k, hx, hy);
    s.P[k] = X.pip[c] - dtau * Kt *
                            (Dh - dsv * rdzc +
                             c2 * (s.Av[k + 1] * (double)X.w[c + PL] - s.Av[k] * (double)X.w[c]) *
                                 rdzc);
    if (X.tpw.p && X.tpw[c2d] > 0.0) s.P[k] += dtau * X.tpi[c];
    s.E[k] = dtau * Kt * c1 * rdzc;
    if (k >= 1) {
      const double rJdzw = m.rdzw[kk] / J;
      const double thvw = thvw_dry(b, X, c, PL) * thvw_fq(p, X, c, PL);
      s.G[k] = dtau * kCp * thvw * c1 * rJdzw;
      const double buoy = kGrav * 0.5 * ((double)X.th[c - PL] + (double)X.th[c]) / (double)b.thbw[c];
      s.R[k] = (double)X.w[c] + dtau * buoy - dtau * kCp * thvw * c2 * (X.pip[c] - X.pip[c - PL]) * rJdzw;
    }
  }
  for (int k = 1; k < nz; ++k) {
    s.A[k] = -s.G[k] * s.E[k - 1] * s.Av[k - 1];
    s.B[k] = 1.0 + s.G[k] * s.Av[k] * (s.E[k] + s.E[k - 1]);
    s.C[k] = -s.G[k] * s.E[k] * s.Av[k + 1];
    s.D[k] = s.R[k] - s.G[k] * (s.P[k] - s.P[k - 1]);
  }
  // the tendency terms enter R (tw) and nothing else of the matrix: the
  // caller adds dtau tw to R before the solve through D (see below)
}

template <int MAXL>
__device__ void thomas(int nz, const double* A, const double* B, const double* C, const double* D,
                       double* cs, double* ds, double* W) {
  for (int k = 1; k < nz; ++k) {
    const double mm = k == 1 ? B[k] : B[k] - A[k] * cs[k - 1];
    cs[k] = C[k] / mm;
    ds[k] = (k == 1 ? D[k] : D[k] - A[k] * ds[k - 1]) / mm;
  }
  W[nz] = 0.0;
  for (int k = nz - 1; k >= 1; --k) W[k] = ds[k] - cs[k] * W[k + 1];
  W[0] = 0.0;
}

template <class S, int MAXL>
__global__ void __launch_bounds__(64) k_tl_wpi(Grid g, Params p, Base<double> b, Metric<double> m,
                                               Coef<double> cf, WPTraj<S> X, const S* tw, WPTan D,
                                               double dtau) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int NX = g.NX, PL = g.plane, nz = g.nz;
  const int c0 = g.id(i, j, 0);
  const double c1 = (1.0 + p.beta) * 0.5, c2 = (1.0 - p.beta) * 0.5;
  WPCol<MAXL> s;
  const int c2d = g.id2(i, j);
  const double mc = m.mapf[c2d], J = m.jac[c2d], hx = m.hx[c2d], hy = m.hy[c2d];
  wp_forward<S, MAXL>(g, p, b, m, cf, X, c0, c2d, mc, dtau, s);
  for (int k = 1; k < nz; ++k) s.D[k] += dtau * (double)tw[c0 + k * PL];
  thomas<MAXL>(nz, s.A, s.B, s.C, s.D, s.cs, s.ds, s.W);
  s.W[0] = w_surface(g, X.u, X.v, c0, hx, hy);
  // perturbations of P, G, R and the rows
  double dP[MAXL], dG[MAXL], dR[MAXL], dD[MAXL], dW[MAXL + 1];
  for (int k = 0; k < nz; ++k) {
    const int c = c0 + k * PL, kk = k + g.ng;
    const double Kt = cf.kt3[c];
    const double dDh = mc *
                       ((cf.rtjx[c + 1] * D.u[c + 1] - cf.rtjx[c] * D.u[c]) * g.rdx +
                        (cf.rtjy[c + NX] * D.v[c + NX] - cf.rtjy[c] * D.v[c]) * g.rdy);
    const double rdzc = m.rdzc[kk];
    const double ddsv = sv_at(g, m, cf, D.u, D.v, c0, k + 1, hx, hy) -
                        sv_at(g, m, cf, D.u, D.v, c0, k, hx, hy);
    dP[k] = D.pip[c] - dtau * Kt *
                           (dDh - ddsv * rdzc +
                            c2 * (s.Av[k + 1] * D.w[c + PL] - s.Av[k] * D.w[c]) * rdzc);
    if (X.tpw.p && X.tpw[c2d] > 0.0) dP[k] += dtau * D.tpi[c];
    dG[k] = 0.0;
    dR[k] = 0.0;
    if (k >= 1) {
      const double rJdzw = m.rdzw[kk] / J;
      const double th_d = thvw_dry(b, X, c, PL), fq = thvw_fq(p, X, c, PL);
      const double thvw = th_d * fq;
      double dthvw = 0.5 * (D.th[c - PL] + D.th[c]) * fq;
      if (p.moisture) dthvw += th_d * kEps61 * 0.5 * (D.qn[c - PL] + D.qn[c]);
      dG[k] = dtau * kCp * dthvw * c1 * rJdzw;
      const double dbuoy = kGrav * 0.5 * (D.th[c - PL] + D.th[c]) / b.thbw[c];
      const double dp = X.pip[c] - X.pip[c - PL], ddp = D.pip[c] - D.pip[c - PL];
      dR[k] = D.w[c] + dtau * (dbuoy + D.tw[c]) -
              dtau * kCp * c2 * rJdzw * (dthvw * dp + thvw * ddp);
    }
  }
  // A dW = dD - dA W_{k-1} - dB W_k - dC W_{k+1}
  for (int k = 1; k < nz; ++k) {
    const double dA = -dG[k] * s.E[k - 1] * s.Av[k - 1];
    const double dB = dG[k] * s.Av[k] * (s.E[k] + s.E[k - 1]);
    const double dC = -dG[k] * s.E[k] * s.Av[k + 1];
    const double dDk = dR[k] - dG[k] * (s.P[k] - s.P[k - 1]) - s.G[k] * (dP[k] - dP[k - 1]);
    dD[k] = dDk - dA * s.W[k - 1] - dB * s.W[k] - dC * s.W[k + 1];
  }
  double cs2[MAXL], ds2[MAXL];
  thomas<MAXL>(nz, s.A, s.B, s.C, dD, cs2, ds2, dW);
  dW[0] = w_surface(g, D.u, D.v, c0, hx, hy);
  for (int k = 0; k < nz; ++k) {
    const int c = c0 + k * PL;
    D.pip_out[c] = dP[k] - s.E[k] * (s.Av[k + 1] * dW[k + 1] - s.Av[k] * dW[k]);
    D.th_out[c] = D.th[c] + dtau * (D.tth[c] - 0.5 * (dW[k] + dW[k + 1]) * b.dthbdz[c]);
    D.w_out[c] = dW[k];
  }
}

template <class S, int MAXL>
__global__ void __launch_bounds__(64) k_ad_wpi(Grid g, Params p, Base<double> b, Metric<double> m,
                                               Coef<double> cf, WPTraj<S> X, const S* tw, WPTan A,
                                               double dtau) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int NX = g.NX, PL = g.plane, nz = g.nz;
  const int c0 = g.id(i, j, 0);
  const double c1 = (1.0 + p.beta) * 0.5, c2 = (1.0 - p.beta) * 0.5;
  WPCol<MAXL> s;
  const int c2d = g.id2(i, j);
  const double mc = m.mapf[c2d], J = m.jac[c2d], hx = m.hx[c2d], hy = m.hy[c2d];
  wp_forward<S, MAXL>(g, p, b, m, cf, X, c0, c2d, mc, dtau, s);
  for (int k = 1; k < nz; ++k) s.D[k] += dtau * (double)tw[c0 + k * PL];
  thomas<MAXL>(nz, s.A, s.B, s.C, s.D, s.cs, s.ds, s.W);
  s.W[0] = w_surface(g, X.u, X.v, c0, hx, hy);
  // outputs -> W, P
  double Wb[MAXL + 1], Pb[MAXL], Gb[MAXL], Rb[MAXL];
  for (int k = 0; k <= nz; ++k) Wb[k] = 0.0;
  for (int k = 0; k < nz; ++k) {
    const int c = c0 + k * PL;
    const double pb = A.pip_out[c], tb = A.th_out[c], wb = A.w_out[c];
    A.pip_out[c] = 0.0;
    A.th_out[c] = 0.0;
    A.w_out[c] = 0.0;
    Pb[k] = pb;
    Wb[k + 1] += -pb * s.E[k] * s.Av[k + 1];
    Wb[k] += pb * s.E[k] * s.Av[k];
    A.th[c] += tb;
    A.tth[c] += dtau * tb;
    const double x = -0.5 * dtau * b.dthbdz[c] * tb;
    Wb[k] += x;
    Wb[k + 1] += x;
    Wb[k] += wb;
    Gb[k] = 0.0;
    Rb[k] = 0.0;
  }
  // W_k (k = 1..nz-1) solves the rows; W_0 and W_nz are fixed.
  // lambda = A^{-T} Wb over rows 1..nz-1: transposed system has sub-diagonal
  // C_{k-1}, diagonal B_k, super-diagonal A_{k+1}.
  double lam[MAXL], cs2[MAXL], ds2[MAXL];
  for (int k = 1; k < nz; ++k) {
    const double sub = k == 1 ? 0.0 : s.C[k - 1];
    const double mm = k == 1 ? s.B[k] : s.B[k] - sub * cs2[k - 1];
    const double sup = k + 1 < nz ? s.A[k + 1] : 0.0;
    cs2[k] = sup / mm;
    ds2[k] = (k == 1 ? Wb[k] : Wb[k] - sub * ds2[k - 1]) / mm;
  }
  for (int k = nz - 1; k >= 1; --k) lam[k] = ds2[k] - (k + 1 < nz ? cs2[k] * lam[k + 1] : 0.0);
  // the surface w(0) = u^x h_x + v^y h_y
  if (Wb[0] != 0.0 && (hx != 0.0 || hy != 0.0)) {
    atomicAdd(&A.u[c0], 0.5 * hx * Wb[0]);
    atomicAdd(&A.u[c0 + 1], 0.5 * hx * Wb[0]);
    atomicAdd(&A.v[c0], 0.5 * hy * Wb[0]);
    atomicAdd(&A.v[c0 + NX], 0.5 * hy * Wb[0]);
  }
  // rows -> G, R, P (Db = lam; Ab = -lam W_{k-1}, Bb = -lam W_k, Cb = -lam W_{k+1})
  for (int k = 1; k < nz; ++k) {
    const double Db = lam[k];
    const double Ab = -lam[k] * s.W[k - 1], Bb = -lam[k] * s.W[k], Cb = -lam[k] * s.W[k + 1];
    Gb[k] += Ab * (-s.E[k - 1] * s.Av[k - 1]) + Bb * s.Av[k] * (s.E[k] + s.E[k - 1]) +
             Cb * (-s.E[k] * s.Av[k + 1]) - Db * (s.P[k] - s.P[k - 1]);
    Rb[k] += Db;
    A.tw[c0 + k * PL] += dtau * Db;
    Pb[k] += -Db * s.G[k];
    Pb[k - 1] += Db * s.G[k];
  }
  // G, R, P -> inputs
  double Svb[MAXL + 1];
  for (int k = 0; k <= nz; ++k) Svb[k] = 0.0;
  for (int k = 0; k < nz; ++k) {
    const int c = c0 + k * PL, kk = k + g.ng;
    const double Kt = cf.kt3[c];
    const double rdzc = m.rdzc[kk];
    // P_k = pip - dtau Kt (Dh - (Sv_{k+1} - Sv_k) rdzc + c2 (Av_{k+1} w_{k+1} - Av_k w_k) rdzc),
    // Dh = m (...): the map factor multiplies Dh alone
    A.pip[c] += Pb[k];
    if (X.tpw.p && X.tpw[c2d] > 0.0) A.tpi[c] += dtau * Pb[k];
    const double Qb = -dtau * Kt * Pb[k];
    const double Dhb = Qb * mc;
    A.w[c + PL] += Qb * c2 * s.Av[k + 1] * rdzc;
    A.w[c] += -Qb * c2 * s.Av[k] * rdzc;
    Svb[k + 1] -= Qb * rdzc;
    Svb[k] += Qb * rdzc;
    atomicAdd(&A.u[c + 1], Dhb * cf.rtjx[c + 1] * g.rdx);
    atomicAdd(&A.u[c], -Dhb * cf.rtjx[c] * g.rdx);
    atomicAdd(&A.v[c + NX], Dhb * cf.rtjy[c + NX] * g.rdy);
    atomicAdd(&A.v[c], -Dhb * cf.rtjy[c] * g.rdy);
    if (k >= 1) {
      const double rJdzw = m.rdzw[kk] / J;
      const double th_d = thvw_dry(b, X, c, PL), fq = thvw_fq(p, X, c, PL);
      const double thvw = th_d * fq;
      const double dp = X.pip[c] - X.pip[c - PL];
      // G = dtau cp thvw c1 rJdzw; R = w + dtau g 0.5 (thm + thk)/thbw - dtau cp thvw c2 dp rJdzw;
      // thvw = (thbw + 0.5 (thm + thk)) (1 + 0.61 0.5 (qm + qk))
      const double thvwb = Gb[k] * dtau * kCp * c1 * rJdzw - Rb[k] * dtau * kCp * c2 * dp * rJdzw;
      if (p.moisture) {
        const double qb = th_d * kEps61 * 0.5 * thvwb;
        atomicAdd(&A.qn[c - PL], qb);
        atomicAdd(&A.qn[c], qb);
      }
      const double thsum_b = 0.5 * thvwb * fq + Rb[k] * dtau * kGrav * 0.5 / b.thbw[c];
      A.th[c - PL] += thsum_b;
      A.th[c] += thsum_b;
      A.w[c] += Rb[k];
      const double dpb = -Rb[k] * dtau * kCp * thvw * c2 * rJdzw;
      A.pip[c] += dpb;
      A.pip[c - PL] -= dpb;
    }
  }
  // S_v -> the u, v faces around each w level
  if (hx != 0.0 || hy != 0.0) {
    for (int k = 1; k < nz; ++k) {
      if (Svb[k] == 0.0) continue;
      const int c = c0 + k * PL;
      const double f = 0.25 * Svb[k] * cf.aw[c] * (1.0 - m.zeta_w[k + g.ng] / m.zt);
      atomicAdd(&A.u[c - PL], f * hx);
      atomicAdd(&A.u[c - PL + 1], f * hx);
      atomicAdd(&A.u[c], f * hx);
      atomicAdd(&A.u[c + 1], f * hx);
      atomicAdd(&A.v[c - PL], f * hy);
      atomicAdd(&A.v[c - PL + NX], f * hy);
      atomicAdd(&A.v[c], f * hy);
      atomicAdd(&A.v[c + NX], f * hy);
    }
  }
}

// ------------------------------------------------------------ objective
// y += a x over n points.
__global__ void k_axpy(int n, double a, const double* __restrict__ x, double* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

// J += scale sum_{i,j} wgt(i,j) 0.5 (u^2 + v^2) at the lowest level.
template <class S>
__global__ void k_ke_objective(Grid g, const S* __restrict__ u, const S* __restrict__ v,
                               const double* __restrict__ wgt, double scale, double* J) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const double w = wgt[g.id2(i, j)];
  if (w == 0.0) return;
  const int c = g.id(i, j, 0);
  const double uc = u[c], vc = v[c];
  atomicAdd(J, scale * w * 0.5 * (uc * uc + vc * vc));
}

// Its gradient: ub += scale wgt u, vb += scale wgt v at the lowest level.
template <class S>
__global__ void k_ke_gradient(Grid g, const S* __restrict__ u, const S* __restrict__ v,
                              const double* __restrict__ wgt, double scale, double* ub,
                              double* vb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const double w = wgt[g.id2(i, j)];
  if (w == 0.0) return;
  const int c = g.id(i, j, 0);
  ub[c] += scale * w * (double)u[c];
  vb[c] += scale * w * (double)v[c];
}

}  // namespace wfe::dev
