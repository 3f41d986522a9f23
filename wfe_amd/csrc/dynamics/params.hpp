#pragma once
// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.

namespace wfe {

// Run parameters of the dynamical core (src/dynamics/params.hpp:8-25 of the
// reference), keyed from the case file by src/main.cpp:284-298.
struct DynParams {
  double diff_K = 0;             // constant diffusivity [m2 s-1] (0 = off)
  double coriolis_f = 0;         // f-plane parameter [s-1] (f(lat) in file mode)
  bool coriolis_use_ub = true;   // idealised: f (u - u_b); file mode: full wind
  double rayleigh_zd = -1;       // Rayleigh layer base [m] (< 0 = off)
  double rayleigh_alpha = 0;     // maximum damping rate [s-1]
  int acoustic_ns = 6;           // acoustic substeps per large step
  double acoustic_beta = 0.2;    // vertical implicit off-centring
  double acoustic_smdiv = 0.1;   // divergence damping (pi' forward weighting)
  bool bc_x_open = false;        // x boundary: periodic or open/radiation
  bool bc_y_open = false;
  double cstar = 30;             // radiation phase speed [m s-1]
  bool moisture = false;         // qv/qc/qr prognostics
  bool microphysics = true;      // Kessler when moisture is on
  bool pd_moist = false;         // positive-definite moisture advection
  bool w_damping = false;        // local w damping where the vertical Courant > 1
  double w_abort = 150;          // |w| guard threshold [m s-1]
};

}  // namespace wfe
