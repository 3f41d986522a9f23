#pragma once
// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.

#include <string>
#include <vector>

#include "core/grid.hpp"

namespace wfe {

// File-mode input (src/io/input.hpp:14-31, src/io/input.cpp:19-73 of the
// reference): <dir>/wfe_input.ini (version >= 3) with nx, ny, nz, np_prof,
// bdy_interval [s], n_bdy, start (YYYYMMDDHH); <dir>/wfe_init.bin holding,
// as consecutive float32 arrays,
//   prof_z, prof_th, prof_qv, prof_u              [np_prof]
//   h, fcor, tsk, land, lat, lon, soilw           [ny*nx]
//   u, v, th, pi, qv                              [nz*ny*nx]   (i fastest)
// and <dir>/wfe_bdy_%03d.bin (file number = hours since start) with
//   u, v, th, pi, qv                              [nz*ny*nx].
class InputData {
 public:
  // Returns false (and writes the reason to stderr) on a missing, old or
  // mismatched input.
  bool load(const GDims& g, const std::string& dir);
  bool load_bdy(const GDims& g, int idx, std::vector<double> fields[5]) const;

  std::vector<double> prof_z, prof_th, prof_qv, prof_u;
  std::vector<double> h, fcor, tsk, land, lat, lon, soilw;
  std::vector<double> u, v, th, pi, qv;
  double bdy_interval = 10800;
  int n_bdy = 0;
  std::string start;

 private:
  std::string dir_;
};

}  // namespace wfe
