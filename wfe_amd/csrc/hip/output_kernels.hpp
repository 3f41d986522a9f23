#pragma once
// Output-path kernels: the driver's interior diagnostics as a device
// reduction, and the interior packing of a field into the reference's
// output order (float32, i fastest, k slowest).
#include "hip/device_common.hpp"

namespace wfe::dev {

// Slots of one diagnostics record (partial or final), in doubles.
enum : int {
  kDgW = 0,     // max |w|, levels 0..nz
  kDgThMin,     // min(0, theta'), levels 0..nz-1
  kDgThMax,     // max(0, theta')
  kDgPip,       // max |pi'|
  kDgQc,        // max(0, qc)
  kDgQr,        // max(0, qr)
  kDgFinite,    // 1 when every w, theta', pi' and qr value is finite, else 0
  kDgSlots = 8  // record stride (one spare slot keeps records 64-byte sized)
};

// Combination rules of the host loops of src/main.cpp:202-240: a value
// replaces the running extreme only when it compares strictly beyond it, so
// NaN never enters (it only clears the finite flag), +-Inf does, and a -0.0
// never replaces the +0.0 start value.
__device__ __forceinline__ double dg_max(double m, double v) { return m < v ? v : m; }
__device__ __forceinline__ double dg_min(double m, double v) { return v < m ? v : m; }

struct DiagAcc {
  double w = 0, thmin = 0, thmax = 0, pip = 0, qc = 0, qr = 0;
  int finite = 1;
  __device__ void merge(const DiagAcc& o) {
    w = dg_max(w, o.w);
    thmin = dg_min(thmin, o.thmin);
    thmax = dg_max(thmax, o.thmax);
    pip = dg_max(pip, o.pip);
    qc = dg_max(qc, o.qc);
    qr = dg_max(qr, o.qr);
    finite &= o.finite;
  }
};

// Block-wide merge through LDS (blockDim.x * blockDim.y <= 256 threads);
// thread 0 returns the block's record.
__device__ __forceinline__ DiagAcc dg_block_merge(DiagAcc a) {
  __shared__ double s[7][256];
  const int t = threadIdx.x + blockDim.x * threadIdx.y;
  const int nt = blockDim.x * blockDim.y;
  s[0][t] = a.w;
  s[1][t] = a.thmin;
  s[2][t] = a.thmax;
  s[3][t] = a.pip;
  s[4][t] = a.qc;
  s[5][t] = a.qr;
  s[6][t] = a.finite;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h && t + h < nt) {
      s[0][t] = dg_max(s[0][t], s[0][t + h]);
      s[1][t] = dg_min(s[1][t], s[1][t + h]);
      s[2][t] = dg_max(s[2][t], s[2][t + h]);
      s[3][t] = dg_max(s[3][t], s[3][t + h]);
      s[4][t] = dg_max(s[4][t], s[4][t + h]);
      s[5][t] = dg_max(s[5][t], s[5][t + h]);
      s[6][t] = s[6][t] < s[6][t + h] ? s[6][t] : s[6][t + h];
    }
    __syncthreads();
  }
  DiagAcc r;
  r.w = s[0][0];
  r.thmin = s[1][0];
  r.thmax = s[2][0];
  r.pip = s[3][0];
  r.qc = s[4][0];
  r.qr = s[5][0];
  r.finite = s[6][0] != 0.0;
  return r;
}

__device__ __forceinline__ void dg_store(double* rec, const DiagAcc& a) {
  rec[kDgW] = a.w;
  rec[kDgThMin] = a.thmin;
  rec[kDgThMax] = a.thmax;
  rec[kDgPip] = a.pip;
  rec[kDgQc] = a.qc;
  rec[kDgQr] = a.qr;
  rec[kDgFinite] = a.finite ? 1.0 : 0.0;
  rec[kDgSlots - 1] = 0.0;
}

// Levels per workgroup of the partial pass.
inline constexpr int kDiagLevels = 8;

// Pass 1: one record per workgroup of 64 x 4 columns and kDiagLevels levels
// of 0..nz, written to part[block * kDgSlots]. qc and qr may be null (dry).
// The grid is (ceil(nx/64), ceil(ny/4), ceil((nz+1)/kDiagLevels)).
template <class T>
__global__ void __launch_bounds__(256)
k_diag_partial(Grid g, const T* __restrict__ w, const T* __restrict__ thp,
               const double* __restrict__ pip, const T* __restrict__ qc,
               const T* __restrict__ qr, double* __restrict__ part) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k0 = blockIdx.z * kDiagLevels;
  DiagAcc a;
  if (i < g.nx && j < g.ny) {
    const int k1 = min(k0 + kDiagLevels, g.nz + 1);
    for (int k = k0; k < k1; ++k) {
      const int c = g.id(i, j, k);
      const double wv = (double)w[c];
      if (!isfinite(wv)) a.finite = 0;
      a.w = dg_max(a.w, fabs(wv));
      if (k < g.nz) {
        const double tv = (double)thp[c];
        if (!isfinite(tv)) a.finite = 0;
        a.thmin = dg_min(a.thmin, tv);
        a.thmax = dg_max(a.thmax, tv);
        const double pv = fabs(pip[c]);
        if (!isfinite(pv)) a.finite = 0;
        a.pip = dg_max(a.pip, pv);
        if (qc) a.qc = dg_max(a.qc, (double)qc[c]);
        if (qr) {
          const double rv = (double)qr[c];
          if (!isfinite(rv)) a.finite = 0;
          a.qr = dg_max(a.qr, rv);
        }
      }
    }
  }
  const DiagAcc r = dg_block_merge(a);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    dg_store(part + (size_t)b * kDgSlots, r);
  }
}

// Pass 2: one workgroup of 256 threads merges n partial records into out.
// Internal linkage: every translation unit that launches it has its own copy.
static __global__ void __launch_bounds__(256)
k_diag_final(const double* __restrict__ part, int n, double* __restrict__ out) {
  DiagAcc a;
  for (int b = threadIdx.x; b < n; b += 256) {
    const double* r = part + (size_t)b * kDgSlots;
    DiagAcc o;
    o.w = r[kDgW];
    o.thmin = r[kDgThMin];
    o.thmax = r[kDgThMax];
    o.pip = r[kDgPip];
    o.qc = r[kDgQc];
    o.qr = r[kDgQr];
    o.finite = r[kDgFinite] != 0.0;
    a.merge(o);
  }
  const DiagAcc r = dg_block_merge(a);
  if (threadIdx.x == 0) dg_store(out, r);
}

// Interior of a field as float32 in the output order of
// src/io/writer.cpp:89-107: out[(k * ny + j) * nx + i] = f(i, j, k) for
// k < nlev. With plane2d the source is a 2-D array (index id2(i, j)) and
// nlev must be 1. Grid: (ceil(nx/64), ceil(ny/4), nlev), 64 x 4 threads.
template <class E>
__global__ void __launch_bounds__(256)
k_pack_interior(Grid g, const E* __restrict__ f, int plane2d, float* __restrict__ out) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const int c = plane2d ? g.id2(i, j) : g.id(i, j, k);
  out[((size_t)k * g.ny + j) * g.nx + i] = (float)f[c];
}

}  // namespace wfe::dev
