#include "dynamics/integrator.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

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
  // per-kernel HIP-event timing on the context's stream (steps then run
  // eagerly, one event pair around each launch)
  if (prof_on_) wfe_check(wfe_kernel_timing(ctx_, 2), "kernel timing");
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
// then the device step.
void Integrator::step(double dt, double t) {
  if (bdy_ && bdy_->active()) bdy_->update(ctx_, t);
  wfe_check(wfe_step(ctx_, dt, t), "Integrator::step");
  ++steps_;
}

// The WFE_PROF=1 section table of src/dynamics/integrator.cpp:17-41,113-122:
// kernel times summed into the reference's eight sections, in its format.
void Integrator::print_profile() const {
  if (!prof_on_ || steps_ == 0) return;
  static const char* kSection[8] = {"sinir kosullari", "kutle akisi+div", "yavas tendency",
                                    "sinir relaks",    "akustik dongu",   "nem asama",
                                    "kessler",         "yuzey/PBL"};
  struct Map {
    const char* kernel;
    int section;
  };
  static const Map kMap[] = {{"fill_ghosts", 0}, {"massflux", 1},      {"tendencies", 2},
                             {"pd_ratio", 2},    {"pd_apply", 2},      {"bdy_relax", 3},
                             {"acou_uv", 4},     {"acou_wpi", 4},      {"acoustic_fused", 4},
                             {"stage_q", 5},     {"kessler", 6},       {"sfc_scalar", 7},
                             {"sfc_u", 7},       {"sfc_v", 7},         {"set_bdy_tfrac", 3}};
  std::vector<char> buf(1 << 16);
  wfe_check(wfe_kernel_timing_report(ctx_, buf.data(), buf.size()), "kernel timing report");
  double sec[8] = {};
  std::istringstream in(buf.data());
  std::string name;
  long launches;
  double ms;
  while (in >> name >> launches >> ms)
    for (const Map& m : kMap)
      if (name == m.kernel) sec[m.section] += ms * 1e-3;
  double tot = 0;
  for (double v : sec) tot += v;
  std::printf("\nprofil (WFE_PROF=1, toplam %.2fs GPU kernel time, %ld adim):\n", tot, steps_);
  for (int q = 0; q < 8; ++q)
    if (sec[q] > 0)
      std::printf("  %-16s %7.2fs  %5.1f%%\n", kSection[q], sec[q], 100.0 * sec[q] / tot);
}

}  // namespace wfe
