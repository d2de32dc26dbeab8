// Aqueous sulfur chemistry of one cell: mo_setsox's setsox_single_level
// (mo_setsox.hpp:943-1259) -- in-cloud pH from the electro-neutrality
// equation, S(IV) oxidation by H2O2 and O3, H2SO4 uptake by cloud water, and
// the cloud-borne sulfate gain split over the modes by their activated
// number.  FP64; one lane per cell.  It produces the cloud-chemistry change
// of the cloud-borne sulfate that the microphysics chain reads as its
// pre-cloud-chemistry state.
#pragma once

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#else
#include <cmath>
// plain C++ compilers see the device annotations as no-ops (the test suite
// builds this file for the CPU: tests/cpp/sox_host.cpp)
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#ifndef __forceinline__
#define __forceinline__ inline
#endif
#endif

namespace mam4x {
namespace sox {

__host__ __device__ __forceinline__ double dmax(double x, double y) { return x < y ? y : x; }
__host__ __device__ __forceinline__ double dmin(double x, double y) { return y < x ? y : x; }

// the constants of mo_setsox::Config (mo_setsox.hpp:14-77)
constexpr double kRa = 8314.0 / 101325.0;       // (atm)/(M K)
constexpr double kKw = 1.0e-14;                 // water acidity
constexpr double kConst0 = 1.0e3 / 6.023e23;    // cm3/L / Avogadro
constexpr double kCo2g = 330.0e-6;              // atm
constexpr double kP0 = 101300.0;                // Pa
constexpr double kSmallLwc = 1.0e-8;            // kg/L

struct Met {
  double press;   // [Pa]
  double tfld;    // [K]
  double lwc;     // cloud liquid water [kg/kg]
  double cldfrc;  // cloud fraction
  double cldnum;  // droplet number [#/kg]
  double xhnm;    // air number density [#/cm3]
};

// the tracers the routine reads or writes (gas_pcnst slots in brackets)
struct Tracers {
  double o3, h2o2, h2so4, so2;  // gases [0], [1], [2], [3], mol/mol
  double so4c[3];               // cloud-borne sulfate, acc/ait/coarse [6], [14], [21]
  double numc[4];               // cloud-borne number [13], [18], [26], [30]; the
                                // last one has no sulfate slot and is never used
};

struct Rates {
  double aqso4[3];    // so4_c tendency from aqueous oxidation [mol/mol/s]
  double aqh2so4[3];  // so4_c tendency from H2SO4 uptake [mol/mol/s]
};

// net positive charge at pH yph (calc_ynetpos, :189-223); xph = [H+]
__host__ __device__ __forceinline__ double ynetpos(double yph, double f1, double f2, double f3,
                                                   double f4, double eco2, double eso4,
                                                   double &xph) {
  xph = pow(10.0, -yph);
  const double eso2 = f1 / (1.0 + f2 * (1.0 + (f3 / xph) * (1.0 + f4 / xph)));
  const double hso3 = eso2 / xph;
  const double so3 = hso3 * 2.0 * f4 / xph;
  return xph - (kKw / xph + eco2 / xph + hso3 + so3 + 1.0 * eso4);
}

// [H+] by bisection of the electro-neutrality equation over pH 2..7
// (calc_ph_values, :225-413).  The bracket halves from width 5 until it is
// <= 0.005, so every cell runs the same ten bisection steps.
__host__ __device__ __forceinline__ double ph_solve(double tfld, double patm, double xlwc,
                                                    double tf, double xso2, double xso4,
                                                    double xhnm) {
  const double xk = 1.230 * exp(3120.0 * tf), xe = 1.7e-2 * exp(2090.0 * tf),
               x2 = 6.0e-8 * exp(1120.0 * tf);
  const double f1 = xk * xe * patm * xso2, f2 = xk * kRa * tfld * xlwc;
  const double eco2 = (3.1e-2 * exp(2423.0 * tf)) * (4.3e-7 * exp(-913.0 * tf)) * kCo2g * patm;
  const double eso4 = xso4 * xhnm * kConst0 / xlwc;
  double xph;
  double lo = 2.0, hi = 7.0;
  (void)ynetpos(lo, f1, f2, xe, x2, eco2, eso4, xph);
  (void)ynetpos(hi, f1, f2, xe, x2, eco2, eso4, xph);
  for (int it = 3; it < 20; ++it) {
    const double y = 0.5 * (lo + hi);
    if (ynetpos(y, f1, f2, xe, x2, eco2, eso4, xph) >= 0.0)
      lo = y;
    else
      hi = y;
    if (fabs(hi - lo) <= 0.005) return pow(10.0, -(0.5 * (hi + lo)));
  }
  return xph;
}

// H2SO4 uptake rate by cloud drops [1/s] (cldaero_uptakerate, :597-695)
__host__ __device__ __forceinline__ double uptake_rate(double xl, double cldnum, double cfact,
                                                       double cldfrc, double tfld,
                                                       double press) {
  const double num_cd = dmax(1.0e-3 * cldnum * cfact / cldfrc, 0.0);
  const double vol = (xl * 0.75) / 3.14159265358979323846;
  double rxn = pow(vol * (num_cd * num_cd), 0.3333333);
  double rad;
  if (rxn <= vol * 4.0e4) {
    rxn = vol * 4.0e4;
    rad = 50.0e-4;
  } else if (rxn >= vol * 4.0e8) {
    rxn = vol * 4.0e8;
    rad = 0.5e-4;
  } else {
    rad = rxn / num_cd;
  }
  const double diffus = 0.557 * pow(tfld, 1.75) / press;
  const double speed = 1.455e4 * sqrt(tfld / 98.0);
  const double kn = 3.0 * diffus / (speed * rad);
  const double fs = (0.4875 * (1.0 + kn)) / (kn * (1.184 + kn) + 0.4875);
  return 12.56637 * rxn * diffus * fs;
}

// One cell.  q is updated in place; r receives the so4_c tendencies.
__host__ __device__ __forceinline__ void single_level(double dt, const Met &m, Tracers &q,
                                                      Rates &r) {
  const double cfact = m.xhnm * 1.0e6 * 1.38e-23 / 287.0 * 1.0e-3;  // [kg/L]
  // sox_cldaero_create_obj (:88-131): in-cloud water, cloud-borne sulfate
  const double xlwc = m.cldfrc > 0.0 ? (m.lwc * cfact) / m.cldfrc : 0.0;
  const double so4c = q.so4c[0] + q.so4c[1] + q.so4c[2];
  const double tf = (1.0 / m.tfld) - (1.0 / 298.0);
  const double patm = m.press / kP0;
  const bool wet = xlwc >= kSmallLwc;
  double xso2 = q.so2, xh2o2 = q.h2o2, xso4 = 0.0;
  double xph = 1.0e-7;
  if (wet) {
    if (m.cldfrc > 0.0) xso4 = so4c / m.cldfrc;
    xph = ph_solve(m.tfld, patm, xlwc, tf, xso2, xso4, m.xhnm);
  }
  // Henry's-law constants at the solved pH and the gas-phase fractions
  const double heh2o2 = (7.4e4 * exp(6621.0 * tf)) * (1.0 + (2.2e-12 * exp(-3730.0 * tf)) / xph);
  const double xk = 1.230 * exp(3120.0 * tf), xe = 1.7e-2 * exp(2090.0 * tf),
               x2 = 6.0e-8 * exp(1120.0 * tf);
  const double heso2 = xk * (1.0 + xe / xph * (1.0 + x2 / xph));
  const double heo3 = 1.15e-2 * exp(2560.0 * tf);
  const double h2o2g = xh2o2 / (1.0 + heh2o2 * kRa * m.tfld * xlwc);
  const double so2g = xso2 / (1.0 + heso2 * kRa * m.tfld * xlwc);
  const double o3g = q.o3 / (1.0 + heo3 * kRa * m.tfld * xlwc);
  const double rah2o2 = 8.e4 * exp(-3650.0 * tf) / (0.1 + xph);
  const double rao3 = 4.39e11 * exp(-4131.0 / m.tfld) + 2.56e3 * exp(-996.0 / m.tfld) / xph;
  // calc_sox_aqueous (:415-522): S(IV) + H2O2, then S(IV) + O3
  double del_hp = 0.0, xso4_init = 0.0;
  if (wet) {
    double pso4 = rah2o2 * 7.4e4 * exp(6621.0 * tf) * h2o2g * patm * 1.23 * exp(3120.0 * tf) *
                  so2g * patm;
    pso4 = pso4 * xlwc / kConst0 / m.xhnm;
    double ds = dmax(pso4 * dt, 1.0e-30);
    xso4_init = xso4;
    if ((ds <= xso2) && (ds <= xh2o2)) {
      xso4 = xso4 + ds;
      xh2o2 = xh2o2 - ds;
      xso2 = xso2 - ds;
    } else if (xh2o2 > xso2) {
      xso4 = xso4 + xso2;
      xh2o2 = xh2o2 - xso2;
      xso2 = 1.0e-20;
    } else {
      xso4 = xso4 + xh2o2;
      xso2 = xso2 - xh2o2;
      xh2o2 = 1.0e-20;
    }
    del_hp = xso4 - xso4_init;
    pso4 = rao3 * heo3 * o3g * patm * heso2 * so2g * patm;
    pso4 = pso4 * xlwc / kConst0 / m.xhnm;
    ds = dmax(pso4 * dt, 1.0e-30);
    xso4_init = xso4;
    if (ds > xso2) {
      xso4 = xso4 + xso2;
      xso2 = 1.0e-20;
    } else {
      xso4 = xso4 + ds;
      xso2 = xso2 - ds;
    }
  }
  // sox_cldaero_update (:732-941)
#pragma unroll
  for (int k = 0; k < 3; ++k) r.aqso4[k] = r.aqh2so4[k] = 0.0;
  if ((m.cldfrc >= 1.0e-5) && wet) {
    // compute_aer_factor (:524-595): shares of the activated number
    double faq[3], sumf = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      faq[k] = dmax(0.0, q.numc[k]);
      if (k == 0) faq[k] = dmax(1.0e-10, faq[k]);
      sumf += faq[k];
    }
    if (sumf > 0.0)
#pragma unroll
      for (int k = 0; k < 3; ++k) faq[k] /= sumf;
    double up = uptake_rate(xlwc, m.cldnum, cfact, m.cldfrc, m.tfld, m.press);
    up = (1.0 - exp(-1.0 * dmin(100.0, dt * up))) / dt;
    const double d_gas = q.h2so4 * up;
    const double d_aq = ((xso4 - xso4_init) + del_hp) / dt;
    const double d_hp = del_hp / dt;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      r.aqso4[k] = faq[k] * d_aq * m.cldfrc;
      r.aqh2so4[k] = faq[k] * d_gas * m.cldfrc;
      q.so4c[k] += ((r.aqso4[k] + r.aqh2so4[k]) + 0.0) * dt;
    }
    q.h2so4 = q.h2so4 - d_gas * dt * m.cldfrc;
    q.so2 += (-d_aq * m.cldfrc + 0.0) * dt;
    q.h2o2 += (-d_hp * m.cldfrc + 0.0) * dt;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) q.so4c[k] = dmax(q.so4c[k], 1.0e-20);
  q.so2 = dmax(q.so2, 1.0e-20);
}

}  // namespace sox
}  // namespace mam4x
