#pragma once
// Device-side views and helpers shared by the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <type_traits>

namespace wfe::dev {

// Device array view. Element i is addressed through a 32-bit unsigned byte
// offset from the base, so loads and stores take the scalar-base plus
// vector-offset form and need no 64-bit address arithmetic per access.
// Indices into the padded layout are never negative (ghost layers are part
// of it) and every array is below 4 GiB. Converts to and from T*.
template <class T>
struct Arr {
  T* p;
  Arr() = default;
  __host__ __device__ constexpr Arr(T* q) : p(q) {}
  __host__ __device__ constexpr Arr(std::nullptr_t) : p(nullptr) {}
  template <class U, class = std::enable_if_t<std::is_convertible_v<U*, T*>>>
  __host__ __device__ constexpr Arr(Arr<U> o) : p(o.p) {}
  __host__ __device__ constexpr operator T*() const { return p; }
  __device__ __forceinline__ T& operator[](int i) const {
    using B = std::conditional_t<std::is_const_v<T>, const char, char>;
    return *reinterpret_cast<T*>(reinterpret_cast<B*>(p) + (unsigned)i * (unsigned)sizeof(T));
  }
};
template <class T>
using CArr = Arr<const T>;

inline constexpr double kGrav = 9.81, kCp = 1004.5, kRd = 287.04, kCv = kCp - kRd;
inline constexpr double kP00 = 1.0e5, kLv = 2.5e6, kLf = 3.34e5, kLs = kLv + kLf;
inline constexpr double kRv = 461.5, kEps61 = 0.61, kTfrz = 273.15;

// Device layout of the padded grid. The reference's ordering (i fastest, ng
// ghosts on each side, one extra w level; src/core/grid.hpp:22-36) is kept,
// but each row is padded to a pitch NX that is a multiple of 32 elements and
// the interior i = 0 sits at column xo = 32, so every row's interior starts
// on a 128-byte line (256 bytes in FP64) and tiles of 32/64 columns map onto
// whole cache lines. Ghost columns i = -ng..-1 sit at xo - ng..xo - 1.
// Extents fit int.
struct Grid {
  int nx, ny, nz, ng;
  int NX;  // row pitch (elements), multiple of 32
  int NY, NZ;
  int plane;  // NX*NY
  int xo;     // column of i = 0 within a row
  double dx, dy;
  double rdx, rdy;  // 1/dx, 1/dy
  __host__ __device__ int id(int i, int j, int k) const {
    return ((k + ng) * NY + (j + ng)) * NX + (i + xo);
  }
  __host__ __device__ int id2(int i, int j) const { return (j + ng) * NX + (i + xo); }
};

inline constexpr int kRowAlign = 32;

// Workgroups are dispatched to the 8 XCDs round-robin by linear id; this
// bijection gives each XCD a contiguous range of logical tiles, so tiles
// that share cache lines (neighbours in x and y) share one L2.
__device__ __forceinline__ int xcd_remap(int id, int n) {
  const int x = id & 7, slot = id >> 3;
  const int q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + slot;
}

// This workgroup's logical (x, y, z) block after the XCD remap of its linear
// dispatch id (x fastest, as the dispatcher orders them). Logical ids run x
// fastest, then z, then y, so one XCD holds whole columns of a band of rows
// and the vertical stencil neighbours of a block share its L2.
struct BlockId {
  int x, y, z;
};
__device__ __forceinline__ BlockId xcd_block() {
  const int gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
  const int id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const int l = xcd_remap(id, gx * gy * gz);
  const int t = l / gx;
  return {l - t * gx, t / gz, t % gz};
}

struct Params {
  double diff_K;
  int coriolis_use_ub;
  double rayleigh_zd, rayleigh_alpha;
  double beta, smdiv;
  int bcx, bcy;  // 1 = open
  double cstar;
  int moisture, pd_moist, w_damping, microphysics;
  int terrain;  // 1 when any terrain slope h_x, h_y is non-zero
  // 1 when the base state and the acoustic coefficients hold the same
  // profile in every column (flat terrain, horizontally uniform base state):
  // kernels then read those profiles from column (0, 0).
  int hom;
};

// Base state (DevProf, src/core/base_state.hpp:16-27). pib stays FP64 so
// that pi = pibar + pi' is formed in FP64 in every precision mode.
template <class T>
struct Base {
  CArr<T> thb, rhob, dthbdz, thbw, rhobw, qvb, ub, vb;
  CArr<double> pib;
  // 1 when every column holds the same profile (flat ground, horizontally
  // uniform base state): at(g, c, k) then reads column (0, 0) of level k.
  int hom;
  __device__ __forceinline__ int at(const Grid& g, int c, int k) const {
    return hom ? g.id(0, 0, k) : c;
  }
};

// Metric (DevMetric, src/core/metric.hpp:16-30).
template <class T>
struct Metric {
  CArr<T> zeta_w, zeta_c, dzeta_c, dzeta_w;
  CArr<T> hx, hy, hxu, hyv, jac, fcor, mapf;
  T zt;
  // FP64 reciprocals for the column solve: 1/dzeta_c, 1/dzeta_w per level
  // (index k + ng) and 1/zt.
  CArr<double> rdzc, rdzw;
  double rzt;
  // sin^2 Rayleigh damping rate per level at cell centres and w levels
  // (ray_alpha, kernels.cu:42-48), formed on the host in FP64.
  CArr<T> ray_c, ray_w;
};

// Constant acoustic coefficients (AcousticCoef, src/dynamics/kernels.hpp:46-50).
template <class T>
struct Coef {
  CArr<T> rtjx, rtjy, kt3, aw;
};

// 1/x for well-scaled x (pivots of a diagonally dominant system, base-state
// theta, the Jacobian): v_rcp_f64 followed by two Newton-Raphson steps
// r <- r + r (1 - x r), within an ulp or so of the quotient.
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// FP32 1/x for well-scaled x: v_rcp_f32 then one Newton-Raphson step.
__device__ __forceinline__ float rcp_nr(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return fmaf(r, fmaf(-x, r, 1.0f), r);
}

// 5th-order upwind interface value between q0 and q1 (WS2002;
// src/dynamics/kernels.cu:51-57): the centred 6-point value minus/plus the
// 5th undivided difference / 60, upwinded on the sign of vel (>= 0 -> minus).
template <class T>
__device__ __forceinline__ T iface5(T qm2, T qm1, T q0, T q1, T q2, T q3, T vel) {
  const T c = (T)(1.0 / 60.0);
  const T cen = c * ((T)37 * (q0 + q1) - (T)8 * (qm1 + q2) + (qm2 + q3));
  const T upw = c * ((q3 - qm2) - (T)5 * (q2 - qm1) + (T)10 * (q1 - q0));
  return (vel >= (T)0) ? (cen - upw) : (cen + upw);
}

// Lateral ghost index maps, composed as the reference's fill order leaves
// them (vertical, then x over all rows, then y over all columns): periodic
// wrap or zero-gradient clamp to [0, last].
__device__ __forceinline__ int map_periodic(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}
__device__ __forceinline__ int map_lat(int i, int n, int last, int open) {
  if (open) return i < 0 ? 0 : (i > last ? last : i);
  return map_periodic(i, n);
}

}  // namespace wfe::dev
