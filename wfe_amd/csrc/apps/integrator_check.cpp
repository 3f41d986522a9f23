// wfe_integrator_check <case.ini> [steps=N] [key=value ...]
//
// Drives the gfx950 integrator through the reference's integrator interface
// (src/dynamics/integrator.hpp:18-42): Integrator::init from the grid, the
// base-state and metric views and DynParams, the initial state through
// state().<field>.upload, then step(dt, t). The same case is also created in
// one call through wfe_case_create_ctx; both are stepped and every state
// field is compared bit for bit. Prints one line per field and "identical"
// when all agree; exit status 0 when identical, 4 when not, 1 on setup
// errors. Idealised cases only (file-mode boundary forcing needs the case
// handle).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "core/case.hpp"
#include "dynamics/integrator.hpp"
#include "wfe_hip.h"

int main(int argc, char** argv) {
  using namespace wfe;
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <case.ini> [steps=N] [key=value ...]\n", argv[0]);
    return 1;
  }
  int steps = 3;
  std::vector<const char*> over;
  Case c;
  if (!c.cfg.load(argv[1])) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 1;
  }
  for (int a = 2; a < argc; ++a) {
    const std::string s = argv[a];
    const size_t eq = s.find('=');
    if (eq == std::string::npos) continue;
    if (s.substr(0, eq) == "steps") {
      steps = std::atoi(s.c_str() + eq + 1);
      continue;
    }
    c.cfg.set(s.substr(0, eq), s.substr(eq + 1));
    over.push_back(argv[a]);
  }
  try {
    c.build();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "case setup failed: %s\n", e.what());
    return 1;
  }
  if (c.file_mode) {
    std::fprintf(stderr, "file-mode cases are not supported by this check\n");
    return 1;
  }
  const int precision =
      c.cfg.get_str("precision", "mixed") == "fp64" ? WFE_PREC_FP64 : WFE_PREC_MIXED;

  // the reference's path: views of the host builders' FP64 arrays
  const BaseState& b = c.base;
  const Metric& m = c.metric;
  const wfe_base bv{b.thb.data(),  b.pib.data(),   b.rhob.data(), b.dthbdz.data(),
                    b.thbw.data(), b.rhobw.data(), b.thvb.data(), b.thvbw.data(),
                    b.qvb.data(),  b.ub.data(),    b.vb.data()};
  const wfe_metric mv{m.zeta_w.data(), m.zeta_c.data(), m.dzeta_c.data(), m.dzeta_w.data(),
                      m.h.data(),      m.hx.data(),     m.hy.data(),      m.hx_u.data(),
                      m.hy_v.data(),   m.jac.data(),    m.fcor.data(),    m.mapf.data(),
                      m.zt};
  Integrator integ;
  integ.init(c.g, bv, mv, c.dp, precision);
  State& s = integ.state();
  Field3D* fields[8] = {&s.u, &s.v, &s.w, &s.thp, &s.pip, &s.qv, &s.qc, &s.qr};
  for (int f = 0; f < 8; ++f) fields[f]->upload(c.ic[f].data());

  // the same case in one call
  wfe_case* wc = nullptr;
  wfe_check(wfe_case_load(argv[1], (int)over.size(), over.data(), &wc), "wfe_case_load");
  wfe_ctx* ctx = nullptr;
  wfe_check(wfe_case_create_ctx(wc, precision, &ctx), "wfe_case_create_ctx");

  for (int n = 0; n < steps; ++n) {
    integ.step(c.dt, n * c.dt);
    wfe_check(wfe_step(ctx, c.dt, n * c.dt), "wfe_step");
  }
  static const char* names[8] = {"u", "v", "w", "thp", "pip", "qv", "qc", "qr"};
  std::vector<double> a(c.g.npts()), r(c.g.npts());
  bool same = true;
  for (int f = 0; f < 8; ++f) {
    fields[f]->download(a.data());
    wfe_check(wfe_download(ctx, f, r.data(), r.size()), "wfe_download");
    double dmax = 0, amax = 0;
    for (size_t q = 0; q < a.size(); ++q) {
      if (std::memcmp(&a[q], &r[q], sizeof(double)) != 0) same = false;
      dmax = std::fmax(dmax, std::fabs(a[q] - r[q]));
      amax = std::fmax(amax, std::fabs(r[q]));
    }
    std::printf("%-4s max|init - case| = %.3e  (max|field| = %.3e)\n", names[f], dmax, amax);
  }
  std::printf("%d steps, precision %s: %s\n", steps,
              precision == WFE_PREC_FP64 ? "fp64" : "mixed", same ? "identical" : "DIFFERENT");
  wfe_destroy(ctx);
  wfe_case_destroy(wc);
  return same ? 0 : 4;
}
