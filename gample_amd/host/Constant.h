// Constant.h -- CGS-Gaussian constants and unit conversions, the values of
// the reference's src/Constant.h:1-24.
#pragma once

namespace gample {
constexpr double me = 9.10938188e-28;      // electron mass, g
constexpr double ec = 4.803204197e-10;     // elementary charge, statC
constexpr double c = 2.99792458e10;        // speed of light, cm/s
constexpr double kb = 1.3806503e-16;       // Boltzmann constant, erg/K
constexpr double evtoerg = 1.6021764262e-12;
constexpr double ergtoev = 1 / 1.6021764262e-12;
constexpr double eVtoK = 11604;
constexpr double VtoG = 1e6 / c;            // V/m -> statV/cm
constexpr double VtosV = 1e8 / c;           // V -> statV
constexpr double sVtoV = 1 / (1e8 / c);
constexpr double JtosJ = 1 / (10 / c * 1e4);
constexpr double GtoV = 1 / (1e6 / c);       // statV/cm -> V/m
constexpr double sJtoJ = 10 / c * 1e4;       // statA/cm^2 -> A/m^2
constexpr double GtoT = 1e-4;                // G -> T
constexpr double TtoG = 1e4;
}  // namespace gample
