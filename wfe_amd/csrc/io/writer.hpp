#pragma once

#include <string>
#include <vector>

#include "core/grid.hpp"
#include "wfe_hip.h"

namespace wfe {

// Snapshot writer with the reference's on-disk format (src/io/writer.hpp:14-33,
// src/io/writer.cpp:12-107): out_dir/meta.json plus one raw float32 file per
// field and output step, <var>_%06d.bin, interior only, i fastest, k slowest,
// nz levels (nz+1 for w); 2-D fields (rain) likewise; static fields as
// <name>.bin.
class Writer {
 public:
  // Throws std::runtime_error when meta.json cannot be written.
  void init(const GDims& g, const std::string& out_dir, double dt, bool moisture, bool has_tsk);
  void write(wfe_ctx* ctx, int step, double t);
  void write_field2d(wfe_ctx* ctx, int field, const char* name, int step);
  void write_static(const char* name, const std::vector<float>& data);
  const std::string& dir() const { return dir_; }

 private:
  void write_field(wfe_ctx* ctx, int field, const char* name, int step, int nzlev);
  void write_raw(const std::string& path, const float* data, size_t n);

  GDims g_{};
  std::string dir_;
  bool moist_ = false;
  std::vector<double> full_;
  std::vector<float> out_;
};

}  // namespace wfe
