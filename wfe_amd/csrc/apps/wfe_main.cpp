// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.
// wfe <case.ini> [key=value ...] -- the model driver of src/main.cpp:245-477
// over the gfx950 integrator: same case keys, stdout diagnostic lines, output
// files and exit codes (0 ok, 1 config/setup/runtime error, 2 NaN in the
// diagnostics, 3 |w| guard). Extra keys: precision = mixed | fp64.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "core/case.hpp"
#include "dynamics/integrator.hpp"
#include "io/writer.hpp"
#include "wfe_hip.h"

namespace {

struct Diag {
  double wmax = 0, thmin = 0, thmax = 0, pipmax = 0, qcmax = 0, qrmax = 0;
  bool finite = true;
};

// Interior diagnostics (src/main.cpp:197-240) on the host from whole-field
// downloads: the reference's path, kept for synchronous output runs.
Diag diagnose_host(const wfe::GDims& g, wfe::State& s, std::vector<double>& buf, bool moist) {
  Diag d;
  s.w.download(buf.data());
  for (int k = 0; k <= g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double v = buf[g.idx(i, j, k)];
        if (!std::isfinite(v)) d.finite = false;
        if (std::fabs(v) > d.wmax) d.wmax = std::fabs(v);
      }
  s.thp.download(buf.data());
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double v = buf[g.idx(i, j, k)];
        if (!std::isfinite(v)) d.finite = false;
        d.thmin = std::min(d.thmin, v);
        d.thmax = std::max(d.thmax, v);
      }
  s.pip.download(buf.data());
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double v = std::fabs(buf[g.idx(i, j, k)]);
        if (!std::isfinite(v)) d.finite = false;
        if (v > d.pipmax) d.pipmax = v;
      }
  if (moist) {
    s.qc.download(buf.data());
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) d.qcmax = std::max(d.qcmax, buf[g.idx(i, j, k)]);
    s.qr.download(buf.data());
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const double v = buf[g.idx(i, j, k)];
          if (!std::isfinite(v)) d.finite = false;
          d.qrmax = std::max(d.qrmax, v);
        }
  }
  return d;
}

// The same record as a device reduction: only the record is copied back.
Diag diagnose_device(wfe_ctx* ctx) {
  wfe_diag r{};
  wfe::wfe_check(wfe_diagnose(ctx, &r), "diagnose");
  Diag d;
  d.wmax = r.wmax;
  d.thmin = r.thmin;
  d.thmax = r.thmax;
  d.pipmax = r.pipmax;
  d.qcmax = r.qcmax;
  d.qrmax = r.qrmax;
  d.finite = r.finite != 0;
  return d;
}

class CaseBoundary : public wfe::BoundarySource {
 public:
  explicit CaseBoundary(wfe_case* c) : c_(c) {}
  bool active() const override { return true; }
  void update(wfe_ctx* ctx, double t) override {
    wfe::wfe_check(wfe_case_update_bdy(c_, ctx, t), "boundary update");
  }

 private:
  wfe_case* c_;
};

}  // namespace

int main(int argc, char** argv) {
  using namespace wfe;
  const std::string cfg_path = argc > 1 ? argv[1] : "cases/warm_bubble.ini";
  std::vector<const char*> over;
  for (int a = 2; a < argc; ++a) over.push_back(argv[a]);
  wfe_case* cs = nullptr;
  int rc = wfe_case_load(cfg_path.c_str(), (int)over.size(), over.data(), &cs);
  if (rc != WFE_OK) {
    std::fprintf(stderr, "case setup failed: %s\n", wfe_last_error());
    return 1;
  }
  // The case is also rebuilt here as a C++ object for the driver's own keys.
  Case c;
  c.cfg.load(cfg_path);
  for (const char* o : over) {
    const std::string s = o;
    const size_t eq = s.find('=');
    if (eq != std::string::npos) c.cfg.set(s.substr(0, eq), s.substr(eq + 1));
  }
  try {
    c.build();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "case setup failed: %s\n", e.what());
    return 1;
  }
  const GDims& g = c.g;
  const std::string prec = c.cfg.get_str("precision", "mixed");
  const int precision = prec == "fp64" ? WFE_PREC_FP64 : WFE_PREC_MIXED;

  char devname[256] = "unknown";
  wfe_device_name(devname, sizeof(devname));
  std::printf("WFE | GPU: %s | grid %dx%dx%d (dx=%gm) | dt=%gs | t_end=%gs\n", devname, g.nx,
              g.ny, g.nz, g.dx, c.dt, c.t_end);
  std::printf("precision: %s\n", precision == WFE_PREC_FP64 ? "fp64" : "mixed (fp32 + fp64 pi')");
  for (const std::string& w : c.warnings) std::fprintf(stderr, "%s\n", w.c_str());
  if (c.sfc.on)
    std::printf("fizik: yuzey + %s PBL + toprak + radyasyon (doy=%d, %02d UTC)\n",
                c.sfc.nonlocal ? "nonlocal" : "yerel-K", c.sfc.doy, (int)c.sfc.start_hour);

  wfe_ctx* ctx = nullptr;
  wfe_check(wfe_case_create_ctx(cs, precision, &ctx), "context");
  Integrator integ;
  integ.adopt(g, c.dp, ctx);
  CaseBoundary bdy(cs);
  if (c.file_mode) integ.set_boundary(&bdy);

  Writer writer;
  try {
    writer.init(g, c.out_dir, c.dt, c.dp.moisture, c.sfc.on);
    std::vector<float> zc((size_t)g.nx * g.ny * g.nz);
    size_t o = 0;
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i)
          zc[o++] = (float)c.metric.z_at(g, i, j, c.metric.zeta_c[k + g.ng]);
    writer.write_static("zc", zc);
    writer.write(ctx, 0, 0);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  {
    const std::string p = c.out_dir + "/run_info.txt";
    if (std::FILE* f = std::fopen(p.c_str(), "w")) {
      std::fprintf(f, "wfe_version = 0.5.0-mi355x\ngit = unknown\nprecision = %s\ngpu = %s\n\n",
                   precision == WFE_PREC_FP64 ? "fp64" : "mixed", devname);
      for (const auto& kv : c.cfg.raw())
        std::fprintf(f, "%s = %s\n", kv.first.c_str(), kv.second.c_str());
      std::fclose(f);
    }
  }
  for (const std::string& k : c.cfg.unused())
    if (k != "precision")
      std::fprintf(stderr, "WARNING: config key never read (typo?): %s\n", k.c_str());

  std::vector<double> diag_buf(writer.asynchronous() ? 0 : g.npts());
  const int nsteps = (int)std::ceil(c.t_end / c.dt);
  const int diag_steps = std::max(1, (int)std::round(c.diag_interval / c.dt));
  const int out_steps = std::max(1, (int)std::round(c.out_interval / c.dt));
  try {
    const auto t0 = std::chrono::steady_clock::now();
    for (int step = 1; step <= nsteps; ++step) {
      integ.step(c.dt, (step - 1) * c.dt);
      const double t = step * c.dt;
      if (step % 10 == 0) {
        float wm = 0;
        wfe_check(wfe_absmax(ctx, WFE_F_W, &wm), "absmax");
        if (!std::isfinite(wm) || wm > (float)c.dp.w_abort) {
          std::fprintf(stderr, "ERROR: |w|max=%.1f m/s exceeded the limit (step %d, t=%.0fs) -- "
                               "writing emergency output and stopping\n",
                       wm, step, t);
          writer.write(ctx, step, t);
          writer.finish();
          wfe_case_destroy(cs);
          return 3;
        }
      }
      if (step % diag_steps == 0 || step == nsteps) {
        const Diag d = writer.asynchronous()
                           ? diagnose_device(ctx)
                           : diagnose_host(g, integ.state(), diag_buf, c.dp.moisture);
        if (c.dp.moisture)
          std::printf("adim %6d  t=%7.1fs  |w|max=%7.3f  th'=[%+.2f,%+.2f]K  "
                      "qc=%.2fg/kg qr=%.2fg/kg\n",
                      step, t, d.wmax, d.thmin, d.thmax, d.qcmax * 1000, d.qrmax * 1000);
        else
          std::printf("adim %6d  t=%7.1fs  |w|max=%7.3f  th'=[%+.3f,%+.3f]K  |pi'|max=%.2e\n",
                      step, t, d.wmax, d.thmin, d.thmax, d.pipmax);
        std::fflush(stdout);
        if (!d.finite) {
          std::fprintf(stderr, "ERROR: NaN/Inf detected, simulation stopped.\n");
          writer.finish();
          wfe_case_destroy(cs);
          return 2;
        }
      }
      if (step % out_steps == 0 || step == nsteps) {
        writer.write(ctx, step, t, true);
      }
    }
    wfe_check(wfe_sync(ctx), "sync");
    writer.finish();
    const double wall =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("bitti: %d adim, %.1f s duvar zamani, %.1f adim/s, gercek-zaman orani %.0fx\n",
                nsteps, wall, nsteps / wall, c.t_end / wall);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  integ.print_profile();
  wfe_case_destroy(cs);
  return 0;
}
