#pragma once
// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.

// Physical constants of the dynamics/microphysics path. The values are fixed
// by the model definition (src/core/constants.hpp:7-17 of the reference);
// they must agree bit-for-bit with the reference for parity.

namespace wfe::phys {

inline constexpr double grav = 9.81;     // gravity [m s-2]
inline constexpr double cp = 1004.5;     // dry-air cp [J kg-1 K-1]
inline constexpr double Rd = 287.04;     // dry-air gas constant [J kg-1 K-1]
inline constexpr double cv = cp - Rd;    // dry-air cv
inline constexpr double p00 = 1.0e5;     // reference pressure [Pa]
inline constexpr double Lv = 2.5e6;      // latent heat of vaporisation [J kg-1]
inline constexpr double Lf = 3.34e5;     // latent heat of fusion [J kg-1]
inline constexpr double Ls = Lv + Lf;    // latent heat of sublimation [J kg-1]
inline constexpr double Rv = 461.5;      // water-vapour gas constant [J kg-1 K-1]
inline constexpr double eps61 = 0.61;    // virtual-temperature factor
inline constexpr double Tfrz = 273.15;   // freezing point [K]

}  // namespace wfe::phys
