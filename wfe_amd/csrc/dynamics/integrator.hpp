#pragma once

#include <cstddef>

#include "core/grid.hpp"
#include "dynamics/params.hpp"
#include "wfe_hip.h"

namespace wfe {

// A device field owned by the integrator context, reached through the C ABI
// (the Field3D upload/download of src/core/field3d.hpp:42-47). Host buffers
// are FP64 in the padded layout.
class Field3D {
 public:
  Field3D() = default;
  Field3D(wfe_ctx* ctx, int field, size_t n) : ctx_(ctx), field_(field), n(n) {}
  void upload(const double* h) const;
  void download(double* h) const;
  int field() const { return field_; }

 private:
  wfe_ctx* ctx_ = nullptr;
  int field_ = 0;

 public:
  size_t n = 0;
};

// Prognostic state at time n (src/dynamics/state.hpp:13-46).
struct State {
  Field3D u, v, w, thp, pip, qv, qc, qr;
};

// Time-dependent lateral boundary source (the BdyManager role of
// src/dynamics/boundary.hpp:16-38): advances the target pair and the time
// fraction for the context before each step.
class BoundarySource {
 public:
  virtual ~BoundarySource() = default;
  virtual bool active() const = 0;
  virtual void update(wfe_ctx* ctx, double t) = 0;
};

// WS2002 RK3 + Klemp-Wilhelmson split-explicit integrator, the interface of
// src/dynamics/integrator.hpp:18-42 over the gfx950 C ABI. Every C ABI
// failure prints the error and exits with status 1, as the reference's
// WFE_CUDA_CHECK does (src/core/cuda_check.hpp:7-16).
class Integrator {
 public:
  Integrator() = default;
  Integrator(const Integrator&) = delete;
  Integrator& operator=(const Integrator&) = delete;
  ~Integrator();

  void init(const GDims& g, const wfe_base& prof, const wfe_metric& m, const DynParams& dp,
            int precision = WFE_PREC_MIXED);
  // Adopts a context already created (e.g. by wfe_case_create_ctx).
  void adopt(const GDims& g, const DynParams& dp, wfe_ctx* ctx);
  void step(double dt, double t);
  void print_profile() const;

  State& state() { return state_; }
  const Field3D& rain() const { return rain_; }
  void set_boundary(BoundarySource* bdy) { bdy_ = bdy; }
  wfe_ctx* context() const { return ctx_; }

 private:
  void bind();
  GDims g_{};
  DynParams dp_{};
  wfe_ctx* ctx_ = nullptr;
  State state_{};
  Field3D rain_{};
  BoundarySource* bdy_ = nullptr;
  bool prof_on_ = false;
  long steps_ = 0;
};

// Aborts with the reference's behaviour on a non-zero C ABI status.
void wfe_check(int rc, const char* what);

}  // namespace wfe
