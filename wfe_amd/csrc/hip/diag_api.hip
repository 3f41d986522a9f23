// Stateless diagnostics over caller-owned device fields: the driver's
// diagnose (src/main.cpp:202-240) for fields that live outside a context,
// in the reference's padded layout (row length nx + 2 ng).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "wfe_hip.h"
#include "hip/output_kernels.hpp"

namespace wfe {
void set_last_error(const std::string& s);
}

// Workspace of wfe_diagnose_fields_noalloc (see below): 16385 records of 8
// doubles, about 1 MiB per GPU.
static __device__ __align__(256) double scratch[(16384 + 1) * 8];

namespace {

struct DiagError : std::runtime_error {
  int code;
  DiagError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw DiagError(WFE_E_HIP, std::string(what) + ": " + hipGetErrorName(e) + ": " +
                                   hipGetErrorString(e));
}

template <class F>
int guarded_diag(const char* fn, F&& f) {
  try {
    f();
    return WFE_OK;
  } catch (const DiagError& e) {
    wfe::set_last_error(std::string(fn) + ": " + e.what());
    return e.code;
  } catch (const std::exception& e) {
    wfe::set_last_error(std::string(fn) + ": " + e.what());
    return WFE_E_HIP;
  }
}

// Device view of the reference layout: pitch nx + 2 ng, interior i = 0 at
// column ng. Returns the launch grid of the partial pass.
wfe::dev::Grid ref_grid(const wfe_grid* g) {
  if (!g) throw DiagError(WFE_E_ARG, "null grid");
  if (g->ng != 3) throw DiagError(WFE_E_ARG, "ng must be 3");
  if (g->nx < 1 || g->ny < 1 || g->nz < 1) throw DiagError(WFE_E_ARG, "empty grid");
  const double np = (double)(g->nx + 6) * (g->ny + 6) * (g->nz + 7);
  if (np * sizeof(double) >= 4294967296.0)
    throw DiagError(WFE_E_ARG, "grid too large: an FP64 field must stay below 4 GiB");
  wfe::dev::Grid d{};
  d.nx = g->nx;
  d.ny = g->ny;
  d.nz = g->nz;
  d.ng = g->ng;
  d.NX = g->nx + 2 * g->ng;
  d.NY = g->ny + 2 * g->ng;
  d.NZ = g->nz + 1 + 2 * g->ng;
  d.plane = d.NX * d.NY;
  d.xo = g->ng;
  d.dx = g->dx;
  d.dy = g->dy;
  d.rdx = g->dx != 0 ? 1.0 / g->dx : 0.0;
  d.rdy = g->dy != 0 ? 1.0 / g->dy : 0.0;
  return d;
}

dim3 diag_grid(const wfe::dev::Grid& g) {
  return dim3((unsigned)((g.nx + 63) / 64), (unsigned)((g.ny + 3) / 4),
              (unsigned)((g.nz + 1 + wfe::dev::kDiagLevels - 1) / wfe::dev::kDiagLevels));
}

void to_diag(const double* r, wfe_diag* out) {
  using namespace wfe::dev;
  out->wmax = r[kDgW];
  out->thmin = r[kDgThMin];
  out->thmax = r[kDgThMax];
  out->pipmax = r[kDgPip];
  out->qcmax = r[kDgQc];
  out->qrmax = r[kDgQr];
  out->finite = r[kDgFinite] != 0.0;
}

// Both passes on `st` with the partial records at part and the final record
// after them; copies the final record into out and waits for it.
void run_diag(const wfe::dev::Grid& g, int precision, const void* w, const void* thp,
              const double* pip, const void* qc, const void* qr, hipStream_t st, double* part,
              int nb, wfe_diag* out) {
  using namespace wfe::dev;
  const dim3 grd = diag_grid(g);
  if (precision == WFE_PREC_FP64)
    k_diag_partial<double><<<grd, dim3(64, 4), 0, st>>>(
        g, static_cast<const double*>(w), static_cast<const double*>(thp), pip,
        static_cast<const double*>(qc), static_cast<const double*>(qr), part);
  else
    k_diag_partial<float><<<grd, dim3(64, 4), 0, st>>>(
        g, static_cast<const float*>(w), static_cast<const float*>(thp), pip,
        static_cast<const float*>(qc), static_cast<const float*>(qr), part);
  hip_ok(hipGetLastError(), "diag partial launch");
  double* fin = part + (size_t)nb * kDgSlots;
  k_diag_final<<<1, 256, 0, st>>>(part, nb, fin);
  hip_ok(hipGetLastError(), "diag final launch");
  double r[kDgSlots];
  hip_ok(hipMemcpyAsync(r, fin, sizeof r, hipMemcpyDeviceToHost, st), "diag copy");
  hip_ok(hipStreamSynchronize(st), "diag sync");
  to_diag(r, out);
}

void check_fields(int precision, const void* w, const void* thp, const double* pip,
                  const void* qc, const void* qr, wfe_diag* out) {
  if (precision != WFE_PREC_FP64 && precision != WFE_PREC_MIXED)
    throw DiagError(WFE_E_ARG, "unknown precision");
  if (!w || !thp || !pip || !out) throw DiagError(WFE_E_ARG, "null field or output");
  if ((qc == nullptr) != (qr == nullptr))
    throw DiagError(WFE_E_ARG, "qc and qr must both be given or both be null");
}

// Workspace of the allocation-free entry point: partial records of at most
// kNoAllocPartials workgroups plus the final record, in one file-scope device
// array. Every GPU holds its own copy at its own address; host code must use
// the address hipGetSymbolAddress returns for the current device, never the
// array's name (a host-side stand-in).
constexpr int kNoAllocPartials = 16384;
static_assert(sizeof(scratch) ==
                  (size_t)(kNoAllocPartials + 1) * wfe::dev::kDgSlots * sizeof(double),
              "workspace size");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
double* device_scratch() {
  int dev = 0;
  hip_ok(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw DiagError(WFE_E_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hip_ok(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) throw DiagError(WFE_E_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int wfe_diagnose_fields(const wfe_grid* g, int precision, const void* w, const void* thp,
                        const double* pip, const void* qc, const void* qr, void* stream,
                        wfe_diag* out) {
  return guarded_diag("wfe_diagnose_fields", [&] {
    const wfe::dev::Grid d = ref_grid(g);
    check_fields(precision, w, thp, pip, qc, qr, out);
    const dim3 grd = diag_grid(d);
    const int nb = (int)(grd.x * grd.y * grd.z);
    hipStream_t st = static_cast<hipStream_t>(stream);
    double* part = nullptr;
    hip_ok(hipMalloc(&part, (size_t)(nb + 1) * wfe::dev::kDgSlots * sizeof(double)),
           "partials allocation");
    struct Free {
      double* p;
      ~Free() { (void)hipFree(p); }
    } fr{part};
    run_diag(d, precision, w, thp, pip, qc, qr, st, part, nb, out);
  });
}

int wfe_diagnose_fields_noalloc(const wfe_grid* g, int precision, const void* w,
                                const void* thp, const double* pip, const void* qc,
                                const void* qr, void* stream, wfe_diag* out) {
  return guarded_diag("wfe_diagnose_fields_noalloc", [&] {
    const wfe::dev::Grid d = ref_grid(g);
    check_fields(precision, w, thp, pip, qc, qr, out);
    const dim3 grd = diag_grid(d);
    const long long nb = (long long)grd.x * grd.y * grd.z;
    if (nb > kNoAllocPartials)
      throw DiagError(WFE_E_ARG,
                      "grid too large for the fixed workspace: " + std::to_string(nb) +
                          " workgroups of 64 x 4 columns x " +
                          std::to_string(wfe::dev::kDiagLevels) + " levels, at most " +
                          std::to_string(kNoAllocPartials) +
                          " (use wfe_diagnose_fields for larger grids)");
    // Every record the kernels read is written earlier in the same call, so
    // nothing left by a previous call needs clearing.
    run_diag(d, precision, w, thp, pip, qc, qr, static_cast<hipStream_t>(stream),
             device_scratch(), (int)nb, out);
  });
}

}  // extern "C"
