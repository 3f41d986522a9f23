#include "dynamics/integrator.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>

namespace wfe {

void wfe_check(int rc, const char* what) {
  if (rc != WFE_OK) {
    std::fprintf(stderr, "wfe error (%d) in %s: %s\n", rc, what, wfe_last_error());
    std::exit(1);
  }
}

void Field3D::upload(const double* h) const { wfe_check(wfe_upload(ctx_, field_, h, n), "upload"); }

void Field3D::download(double* h) const {
  wfe_check(wfe_download(ctx_, field_, h, n), "download");
}

Integrator::~Integrator() {
  if (ctx_) wfe_destroy(ctx_);
}

void Integrator::bind() {
  const size_t n = g_.npts();
  state_ = State{Field3D(ctx_, WFE_F_U, n),   Field3D(ctx_, WFE_F_V, n),
                 Field3D(ctx_, WFE_F_W, n),   Field3D(ctx_, WFE_F_THP, n),
                 Field3D(ctx_, WFE_F_PIP, n), Field3D(ctx_, WFE_F_QV, n),
                 Field3D(ctx_, WFE_F_QC, n),  Field3D(ctx_, WFE_F_QR, n)};
  rain_ = Field3D(ctx_, WFE_F_RAIN, g_.n2());
  const char* e = std::getenv("WFE_PROF");
  prof_on_ = e && e[0] == '1';
}

// Integrator::init (src/dynamics/integrator.cpp:43-69).
void Integrator::init(const GDims& g, const wfe_base& prof, const wfe_metric& m,
                      const DynParams& dp, int precision) {
  g_ = g;
  dp_ = dp;
  if (g.nz + 1 > kMaxColumnLevels) {
    std::fprintf(stderr, "nz=%d too large (column solver limit %d)\n", g.nz,
                 kMaxColumnLevels - 1);
    std::exit(1);
  }
  const wfe_grid gr{g.nx, g.ny, g.nz, g.ng, g.dx, g.dy, g.dz};
  const wfe_params p{dp.diff_K,         dp.coriolis_f,   dp.coriolis_use_ub, dp.rayleigh_zd,
                     dp.rayleigh_alpha, dp.acoustic_ns,  dp.acoustic_beta,   dp.acoustic_smdiv,
                     dp.bc_x_open,      dp.bc_y_open,    dp.cstar,           dp.moisture,
                     dp.microphysics,   dp.pd_moist,     dp.w_damping,       dp.w_abort};
  wfe_check(wfe_create(&gr, &p, &prof, &m, precision, &ctx_), "Integrator::init");
  bind();
}

void Integrator::adopt(const GDims& g, const DynParams& dp, wfe_ctx* ctx) {
  g_ = g;
  dp_ = dp;
  ctx_ = ctx;
  bind();
}

// Integrator::step (src/dynamics/integrator.cpp:71-111): boundary update,
// then the device step. With WFE_PROF=1 each step is timed synchronously.
void Integrator::step(double dt, double t) {
  if (bdy_ && bdy_->active()) bdy_->update(ctx_, t);
  std::chrono::steady_clock::time_point t0;
  if (prof_on_) {
    wfe_check(wfe_sync(ctx_), "sync");
    t0 = std::chrono::steady_clock::now();
  }
  wfe_check(wfe_step(ctx_, dt, t), "Integrator::step");
  if (prof_on_) {
    wfe_check(wfe_sync(ctx_), "sync");
    wall_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    ++steps_;
  }
}

void Integrator::print_profile() const {
  if (!prof_on_ || steps_ == 0) return;
  std::printf("\nprofile (WFE_PROF=1): %ld steps, %.3f s GPU-synchronous, %.3f ms/step\n",
              steps_, wall_, 1e3 * wall_ / (double)steps_);
}

}  // namespace wfe
