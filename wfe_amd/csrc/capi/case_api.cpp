// C ABI of the host case builder (include/wfe_hip.h, wfe_case_*): the
// Config + Metric + BaseState + initial-condition assembly of
// src/main.cpp:248-383, exposed for the driver, the Python layer and tests.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "core/case.hpp"
#include "wfe_hip.h"

namespace wfe {
void set_last_error(const std::string& s);
}

// A case and, per context it drives, that context's boundary bracket (the
// BdyManager state of src/dynamics/boundary.cpp:46-48,83-91), so one case
// can drive several contexts that stand at different times. Converted
// boundary files are kept for the two most recent file indices.
struct wfe_case {
  wfe::Case c;
  struct Bracket {
    int idx_hi = 1;
    double t_lo = 0, t_hi = 0;
  };
  std::map<const wfe_ctx*, Bracket> brackets;
  std::map<int, std::array<std::vector<double>, 5>> files;
};

namespace {

template <class F>
int guarded_case(F&& f) {
  try {
    f();
    return WFE_OK;
  } catch (const std::bad_alloc&) {
    wfe::set_last_error("host allocation failed");
    return WFE_E_NOMEM;
  } catch (const std::exception& e) {
    wfe::set_last_error(e.what());
    return WFE_E_CONFIG;
  }
}

const std::vector<double>* case_array(const wfe::Case& c, const std::string& n) {
  const wfe::BaseState& b = c.base;
  const wfe::Metric& m = c.metric;
  struct NV {
    const char* n;
    const std::vector<double>* v;
  };
  const NV t[] = {{"thb", &b.thb},       {"pib", &b.pib},         {"rhob", &b.rhob},
                  {"dthbdz", &b.dthbdz}, {"thbw", &b.thbw},       {"rhobw", &b.rhobw},
                  {"thvb", &b.thvb},     {"thvbw", &b.thvbw},     {"qvb", &b.qvb},
                  {"ub", &b.ub},         {"vb", &b.vb},           {"zeta_w", &m.zeta_w},
                  {"zeta_c", &m.zeta_c}, {"dzeta_c", &m.dzeta_c}, {"dzeta_w", &m.dzeta_w},
                  {"h", &m.h},           {"hx", &m.hx},           {"hy", &m.hy},
                  {"hx_u", &m.hx_u},     {"hy_v", &m.hy_v},       {"jac", &m.jac},
                  {"fcor", &m.fcor},     {"mapf", &m.mapf},       {"u", &c.ic[0]},
                  {"v", &c.ic[1]},       {"w", &c.ic[2]},         {"thp", &c.ic[3]},
                  {"pip", &c.ic[4]},     {"qv", &c.ic[5]},        {"qc", &c.ic[6]},
                  {"qr", &c.ic[7]},      {"bdy_wgt", &c.bdy.wgt},
                  {"sfc_tsk", &c.sfc.tsk}, {"sfc_land", &c.sfc.land}, {"sfc_lat", &c.sfc.lat},
                  {"sfc_lon", &c.sfc.lon}, {"sfc_soilw", &c.sfc.soilw}};
  for (const NV& e : t)
    if (n == e.n) return e.v;
  static const char* vn[5] = {"u", "v", "thp", "pip", "qv"};
  for (int v = 0; v < 5; ++v) {
    if (n == std::string("bdy_lo_") + vn[v]) return &c.bdy.lo[v];
    if (n == std::string("bdy_hi_") + vn[v]) return &c.bdy.hi[v];
  }
  return nullptr;
}

void fill_views(const wfe::Case& c, wfe_grid& g, wfe_params& p, wfe_base& b, wfe_metric& m) {
  g = wfe_grid{c.g.nx, c.g.ny, c.g.nz, c.g.ng, c.g.dx, c.g.dy, c.g.dz};
  const wfe::DynParams& d = c.dp;
  p = wfe_params{d.diff_K,         d.coriolis_f,     d.coriolis_use_ub, d.rayleigh_zd,
                 d.rayleigh_alpha, d.acoustic_ns,    d.acoustic_beta,   d.acoustic_smdiv,
                 d.bc_x_open,      d.bc_y_open,      d.cstar,           d.moisture,
                 d.microphysics,   d.pd_moist,       d.w_damping,       d.w_abort};
  const wfe::BaseState& s = c.base;
  b = wfe_base{s.thb.data(),  s.pib.data(),   s.rhob.data(), s.dthbdz.data(),
               s.thbw.data(), s.rhobw.data(), s.thvb.data(), s.thvbw.data(),
               s.qvb.data(),  s.ub.data(),    s.vb.data()};
  const wfe::Metric& t = c.metric;
  m = wfe_metric{t.zeta_w.data(), t.zeta_c.data(), t.dzeta_c.data(), t.dzeta_w.data(),
                 t.h.data(),      t.hx.data(),     t.hy.data(),      t.hx_u.data(),
                 t.hy_v.data(),   t.jac.data(),    t.fcor.data(),    t.mapf.data(),
                 t.zt};
}

}  // namespace

extern "C" {

int wfe_case_load(const char* ini_path, int n_over, const char* const* over, wfe_case** out) {
  if (!ini_path || !out || (n_over > 0 && !over)) {
    wfe::set_last_error("null argument");
    return WFE_E_ARG;
  }
  *out = nullptr;
  auto cs = std::make_unique<wfe_case>();
  if (!cs->c.cfg.load(ini_path)) {
    wfe::set_last_error(std::string("cannot read config: ") + ini_path);
    return WFE_E_IO;
  }
  for (int a = 0; a < n_over; ++a) {
    const std::string s = over[a] ? over[a] : "";
    const size_t eq = s.find('=');
    if (eq != std::string::npos) cs->c.cfg.set(s.substr(0, eq), s.substr(eq + 1));
  }
  const int rc = guarded_case([&] { cs->c.build(); });
  if (rc == WFE_OK) *out = cs.release();
  return rc;
}

int wfe_case_destroy(wfe_case* c) {
  delete c;
  return WFE_OK;
}

int wfe_case_grid(const wfe_case* c, wfe_grid* g) {
  if (!c || !g) return WFE_E_ARG;
  *g = wfe_grid{c->c.g.nx, c->c.g.ny, c->c.g.nz, c->c.g.ng, c->c.g.dx, c->c.g.dy, c->c.g.dz};
  return WFE_OK;
}

int wfe_case_params(const wfe_case* c, wfe_params* p) {
  if (!c || !p) return WFE_E_ARG;
  wfe_grid g;
  wfe_base b;
  wfe_metric m;
  fill_views(c->c, g, *p, b, m);
  return WFE_OK;
}

int wfe_case_scalar(const wfe_case* c, const char* name, double* out) {
  if (!c || !name || !out) return WFE_E_ARG;
  const std::string n(name);
  if (n == "dt") *out = c->c.dt;
  else if (n == "t_end") *out = c->c.t_end;
  else if (n == "out_interval") *out = c->c.out_interval;
  else if (n == "diag_interval") *out = c->c.diag_interval;
  else if (n == "zt") *out = c->c.metric.zt;
  else if (n == "sfc_on") *out = c->c.sfc.on;
  else if (n == "sfc_start_hour") *out = c->c.sfc.start_hour;
  else if (n == "sfc_doy") *out = c->c.sfc.doy;
  else if (n == "sfc_nonlocal") *out = c->c.sfc.nonlocal;
  else {
    wfe::set_last_error("unknown scalar: " + n);
    return WFE_E_ARG;
  }
  return WFE_OK;
}

size_t wfe_case_array_size(const wfe_case* c, const char* name) {
  if (!c || !name) return 0;
  const std::vector<double>* v = case_array(c->c, name);
  return v ? v->size() : 0;
}

int wfe_case_array(const wfe_case* c, const char* name, double* out, size_t n) {
  if (!c || !name || !out) return WFE_E_ARG;
  const std::vector<double>* v = case_array(c->c, name);
  if (!v || v->size() != n) {
    wfe::set_last_error(std::string("unknown array or size mismatch: ") + name);
    return WFE_E_ARG;
  }
  std::memcpy(out, v->data(), n * sizeof(double));
  return WFE_OK;
}

int wfe_case_create_ctx(const wfe_case* cs, int precision, wfe_ctx** out) {
  if (!cs || !out) return WFE_E_ARG;
  const wfe::Case& c = cs->c;
  wfe_grid g;
  wfe_params p;
  wfe_base b;
  wfe_metric m;
  fill_views(c, g, p, b, m);
  int rc = wfe_create(&g, &p, &b, &m, precision, out);
  if (rc != WFE_OK) return rc;
  static const int fields[8] = {WFE_F_U,   WFE_F_V,  WFE_F_W,  WFE_F_THP,
                                WFE_F_PIP, WFE_F_QV, WFE_F_QC, WFE_F_QR};
  for (int q = 0; q < 8 && rc == WFE_OK; ++q)
    rc = wfe_upload(*out, fields[q], c.ic[q].data(), c.ic[q].size());
  if (rc == WFE_OK && c.bdy.active) {
    rc = wfe_set_bdy_weights(*out, c.bdy.wgt.data(), c.bdy.wgt.size());
    // A new context starts at the first two boundary times. The host pair
    // follows the context advanced last, so once any context of this case has
    // crossed a boundary time the first pair is read again from its files.
    std::array<std::vector<double>, 5> lo0, hi0;
    const std::array<std::vector<double>, 5>* lo = &c.bdy.lo;
    const std::array<std::vector<double>, 5>* hi = &c.bdy.hi;
    if (c.bdy.idx_hi != 1) {
      bool ok = false;
      rc = guarded_case([&] {
        ok = c.bdy.load_into(c.g, c.input, c.base.thb, c.base.pib, 0, lo0) &&
             c.bdy.load_into(c.g, c.input, c.base.thb, c.base.pib, std::min(1, c.input.n_bdy - 1),
                             hi0);
      });
      if (rc == WFE_OK && !ok) {
        wfe::set_last_error("cannot read the first boundary files");
        rc = WFE_E_IO;
      }
      lo = &lo0;
      hi = &hi0;
    }
    for (int v = 0; v < 5 && rc == WFE_OK; ++v) {
      rc = wfe_set_bdy_target(*out, 0, v, (*lo)[v].data(), (*lo)[v].size());
      if (rc == WFE_OK) rc = wfe_set_bdy_target(*out, 1, v, (*hi)[v].data(), (*hi)[v].size());
    }
  }
  if (rc == WFE_OK && c.sfc.on) {
    const wfe_sfc sf{c.sfc.tsk.data(), c.sfc.land.data(), c.sfc.lat.data(), c.sfc.lon.data(),
                     c.sfc.soilw.data(), c.sfc.start_hour, c.sfc.doy, c.sfc.nonlocal ? 1 : 0};
    rc = wfe_set_physics(*out, &sf, c.sfc.tsk.size());
  }
  if (rc != WFE_OK) {
    wfe_destroy(*out);
    *out = nullptr;
    return rc;
  }
  if (c.bdy.active) {
    wfe_case::Bracket b;
    b.t_hi = c.input.bdy_interval;
    const_cast<wfe_case*>(cs)->brackets[*out] = b;
  }
  return rc;
}

int wfe_case_update_bdy(wfe_case* cs, wfe_ctx* ctx, double t) {
  if (!cs || !ctx) return WFE_E_ARG;
  wfe::Case& c = cs->c;
  if (!c.bdy.active) return WFE_OK;
  // this context's bracket; a context first seen here holds the targets of
  // wfe_case_create_ctx (the first two boundary times)
  auto ins = cs->brackets.emplace(ctx, wfe_case::Bracket{});
  wfe_case::Bracket& b = ins.first->second;
  if (ins.second) b.t_hi = c.input.bdy_interval;
  int rc = WFE_OK;
  // BdyManager::update (boundary.cpp:83-91): while t has passed the upper
  // time and a later file exists, the upper target becomes the lower one
  // (a buffer swap on the device) and the next file is uploaded as the upper
  while (t >= b.t_hi && b.idx_hi + 1 < c.input.n_bdy) {
    const int next = b.idx_hi + 1;
    auto f = cs->files.find(next);
    if (f == cs->files.end()) {
      std::array<std::vector<double>, 5> tg;
      bool ok = false;
      rc = guarded_case([&] { ok = c.bdy.load_into(c.g, c.input, c.base.thb, c.base.pib, next, tg); });
      if (rc != WFE_OK) return rc;
      if (!ok) {
        wfe::set_last_error("cannot read boundary file " + std::to_string(next));
        return WFE_E_IO;
      }
      while (cs->files.size() >= 2) cs->files.erase(cs->files.begin());
      f = cs->files.emplace(next, std::move(tg)).first;
    }
    rc = wfe_swap_bdy_targets(ctx);
    for (int v = 0; v < 5 && rc == WFE_OK; ++v)
      rc = wfe_set_bdy_target(ctx, 1, v, f->second[v].data(), f->second[v].size());
    if (rc != WFE_OK) return rc;
    // the host copy shows the pair of the most recently advanced context
    c.bdy.lo.swap(c.bdy.hi);
    c.bdy.hi = f->second;
    c.bdy.idx_hi = next;
    b.idx_hi = next;
    b.t_lo = b.t_hi;
    b.t_hi += c.input.bdy_interval;
    c.bdy.t_lo = b.t_lo;
    c.bdy.t_hi = b.t_hi;
  }
  // BdyManager::tfrac (boundary.cpp:93-97)
  double tf = 0;
  if (b.t_hi > b.t_lo) tf = std::clamp((t - b.t_lo) / (b.t_hi - b.t_lo), 0.0, 1.0);
  return wfe_set_bdy_tfrac(ctx, tf);
}

int wfe_case_unused_keys(const wfe_case* c, char* buf, size_t n) {
  if (!c || !buf || n == 0) return WFE_E_ARG;
  std::string s;
  for (const std::string& k : c->c.cfg.unused()) s += k + "\n";
  std::snprintf(buf, n, "%s", s.c_str());
  return WFE_OK;
}

}  // extern "C"
