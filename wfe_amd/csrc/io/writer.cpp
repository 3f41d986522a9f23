#include "io/writer.hpp"

#include <cstdio>
#include <filesystem>
#include <stdexcept>

namespace wfe {

void Writer::init(const GDims& g, const std::string& out_dir, double dt, bool moisture,
                  bool has_tsk) {
  g_ = g;
  dir_ = out_dir;
  moist_ = moisture;
  std::filesystem::create_directories(dir_);
  full_.resize(g.npts());
  out_.resize((size_t)g.nx * g.ny * (g.nz + 1));
  const std::string meta = dir_ + "/meta.json";
  std::FILE* f = std::fopen(meta.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + meta);
  std::fprintf(f,
               "{\n"
               "  \"nx\": %d, \"ny\": %d, \"nz\": %d,\n"
               "  \"dx\": %g, \"dy\": %g, \"dz\": %g,\n"
               "  \"dt\": %g,\n"
               "  \"vars\": {\"u\": %d, \"v\": %d, \"w\": %d, \"thp\": %d, \"pip\": %d",
               g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, dt, g.nz, g.nz, g.nz + 1, g.nz, g.nz);
  if (moist_)
    std::fprintf(f, ", \"qv\": %d, \"qc\": %d, \"qr\": %d, \"rain\": 1", g.nz, g.nz, g.nz);
  if (has_tsk) std::fprintf(f, ", \"tsk\": 1, \"pblh\": 1, \"t2m\": 1, \"u10\": 1");
  std::fprintf(f, "}\n}\n");
  std::fclose(f);
}

void Writer::write_raw(const std::string& path, const float* data, size_t n) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot open output file " + path);
  const size_t wrote = std::fwrite(data, sizeof(float), n, f);
  std::fclose(f);
  if (wrote != n) throw std::runtime_error("short write to " + path);
}

void Writer::write_field(wfe_ctx* ctx, int field, const char* name, int step, int nzlev) {
  if (wfe_download(ctx, field, full_.data(), full_.size()) != WFE_OK)
    throw std::runtime_error(std::string("download failed: ") + wfe_last_error());
  size_t o = 0;
  for (int k = 0; k < nzlev; ++k)
    for (int j = 0; j < g_.ny; ++j)
      for (int i = 0; i < g_.nx; ++i) out_[o++] = (float)full_[g_.idx(i, j, k)];
  char path[1024];
  std::snprintf(path, sizeof(path), "%s/%s_%06d.bin", dir_.c_str(), name, step);
  write_raw(path, out_.data(), o);
}

void Writer::write(wfe_ctx* ctx, int step, double t) {
  write_field(ctx, WFE_F_U, "u", step, g_.nz);
  write_field(ctx, WFE_F_V, "v", step, g_.nz);
  write_field(ctx, WFE_F_W, "w", step, g_.nz + 1);
  write_field(ctx, WFE_F_THP, "thp", step, g_.nz);
  write_field(ctx, WFE_F_PIP, "pip", step, g_.nz);
  if (moist_) {
    write_field(ctx, WFE_F_QV, "qv", step, g_.nz);
    write_field(ctx, WFE_F_QC, "qc", step, g_.nz);
    write_field(ctx, WFE_F_QR, "qr", step, g_.nz);
  }
  std::printf("  cikti yazildi: step %d (t = %.1f s)\n", step, t);
}

void Writer::write_field2d(wfe_ctx* ctx, int field, const char* name, int step) {
  std::vector<double> h2(g_.n2());
  if (wfe_download(ctx, field, h2.data(), h2.size()) != WFE_OK)
    throw std::runtime_error(std::string("download failed: ") + wfe_last_error());
  std::vector<float> o((size_t)g_.nx * g_.ny);
  size_t idx = 0;
  for (int j = 0; j < g_.ny; ++j)
    for (int i = 0; i < g_.nx; ++i) o[idx++] = (float)h2[g_.idx2(i, j)];
  char path[1024];
  std::snprintf(path, sizeof(path), "%s/%s_%06d.bin", dir_.c_str(), name, step);
  write_raw(path, o.data(), o.size());
}

void Writer::write_static(const char* name, const std::vector<float>& data) {
  char path[1024];
  std::snprintf(path, sizeof(path), "%s/%s.bin", dir_.c_str(), name);
  write_raw(path, data.data(), data.size());
}

}  // namespace wfe
