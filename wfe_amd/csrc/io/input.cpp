// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.
#include "io/input.hpp"

#include <cstdio>
#include <memory>

#include "core/config.hpp"

namespace wfe {

namespace {
struct FileCloser {
  void operator()(std::FILE* f) const {
    if (f) std::fclose(f);
  }
};
using File = std::unique_ptr<std::FILE, FileCloser>;

bool read_f32(std::FILE* f, std::vector<double>& dst, size_t n) {
  std::vector<float> tmp(n);
  if (n && std::fread(tmp.data(), sizeof(float), n, f) != n) return false;
  dst.assign(tmp.begin(), tmp.end());
  return true;
}
}  // namespace

bool InputData::load(const GDims& g, const std::string& dir) {
  dir_ = dir;
  Config meta;
  if (!meta.load(dir + "/wfe_input.ini")) {
    std::fprintf(stderr, "cannot read input metadata: %s/wfe_input.ini\n", dir.c_str());
    return false;
  }
  if (meta.get_int("nx", -1) != g.nx || meta.get_int("ny", -1) != g.ny ||
      meta.get_int("nz", -1) != g.nz) {
    std::fprintf(stderr, "input grid does not match the case grid\n");
    return false;
  }
  if (meta.get_int("version", 1) < 3) {
    std::fprintf(stderr, "input format older than version 3\n");
    return false;
  }
  const int npf = meta.get_int("np_prof", 0);
  bdy_interval = meta.get_real("bdy_interval", 10800);
  n_bdy = meta.get_int("n_bdy", 0);
  start = meta.get_str("start", "");

  File f(std::fopen((dir + "/wfe_init.bin").c_str(), "rb"));
  if (!f) {
    std::fprintf(stderr, "cannot open %s/wfe_init.bin\n", dir.c_str());
    return false;
  }
  const size_t n2 = (size_t)g.nx * g.ny, n3 = n2 * g.nz;
  std::FILE* fp = f.get();
  const bool ok = read_f32(fp, prof_z, npf) && read_f32(fp, prof_th, npf) &&
                  read_f32(fp, prof_qv, npf) && read_f32(fp, prof_u, npf) &&
                  read_f32(fp, h, n2) && read_f32(fp, fcor, n2) && read_f32(fp, tsk, n2) &&
                  read_f32(fp, land, n2) && read_f32(fp, lat, n2) && read_f32(fp, lon, n2) &&
                  read_f32(fp, soilw, n2) && read_f32(fp, u, n3) && read_f32(fp, v, n3) &&
                  read_f32(fp, th, n3) && read_f32(fp, pi, n3) && read_f32(fp, qv, n3);
  if (!ok) std::fprintf(stderr, "wfe_init.bin is truncated or corrupt\n");
  return ok;
}

bool InputData::load_bdy(const GDims& g, int idx, std::vector<double> fields[5]) const {
  char path[1024];
  std::snprintf(path, sizeof(path), "%s/wfe_bdy_%03d.bin", dir_.c_str(),
                (int)(idx * bdy_interval / 3600));
  File f(std::fopen(path, "rb"));
  if (!f) {
    std::fprintf(stderr, "cannot open boundary file %s\n", path);
    return false;
  }
  const size_t n3 = (size_t)g.nx * g.ny * g.nz;
  for (int v = 0; v < 5; ++v)
    if (!read_f32(f.get(), fields[v], n3)) {
      std::fprintf(stderr, "boundary file truncated or corrupt: %s\n", path);
      return false;
    }
  return true;
}

}  // namespace wfe
