// Linearized stratospheric ozone chemistry (LINOZ, Hsu and Prather 2008) of
// one column and the invariant densities of one cell -- the feeders of the
// gas-phase chemistry carried by mo_gas_phase_chemdr.hpp:
//   * the overhead ozone column, a running sum from the model top
//     (compute_o3_column_density, compute_o3_column_density.hpp:59-83);
//   * the relaxation of ozone towards its climatological steady state above
//     the tropopause, with polar-stratospheric-cloud loss (lin_strat_chem.hpp:
//     31-222);
//   * the surface sink of the linoz ozone in the lowest levels
//     (lin_strat_chem.hpp:263-316);
//   * setinv_single_level (mo_setinv.hpp:45-86): M, N2, O2, H2O and the four
//     offline species as number densities.
// FP64.  One lane per column for the linoz walk (the column sum couples the
// levels), one lane per cell for the invariants.
#pragma once

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#else
#include <cmath>
// plain C++ compilers see the device annotations as no-ops (the test suite
// builds this file for the CPU: tests/cpp/linoz_host.cpp)
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
namespace lnz {

constexpr double kPi = 3.14159265358979323846264;
constexpr double kRadToDeg = 180. / kPi;
// [mol/cm2] per (Pa x vmr): 2.8704e21 / (g k) with the reference's constants
constexpr double kColumnFactor = 2.8704e21 / (9.80616 * 1.38044);
constexpr double kToDobson = 1.0 / 2.687e16;  // [mol/cm2] -> [DU]
constexpr double kChlorine1987 = 2.5977;      // EESC [ppbv]

// the eight climatology planes, in LinozData order (mo_gas_phase_chemdr.hpp:
// 35-52)
enum { kO3Clim, kTClim, kO3colClim, kPmLClim, kdPmLdO3, kdPmLdT, kdPmLdO3col, kCariolle, kNClim };

struct Conf {
  double chlorine_loading;  // EESC [ppbv]
  double psc_T;             // PSC ozone-loss temperature threshold [K]
  double o3_sfc;            // surface-sink target [vmr]
  double o3_tau;            // surface-sink time scale [s]
  int o3_lbl;               // levels of the surface sink (0: none)
};

struct LevelOut {
  double do3, do3_psc, ss_o3;
};

// lin_strat_chem_solve_kk with psc_activation: the new ozone of one level.
// The PSC loss, where it applies, replaces the relaxation (it starts again
// from the old ozone).
__host__ __device__ __forceinline__ double solve_level(double o3col, double temp, double sza,
                                                       double pmid, double dt, double lats,
                                                       const double clim[kNClim],
                                                       const Conf &cf, double o3_old,
                                                       LevelOut &out) {
  const bool excess_chlorine = (cf.chlorine_loading - 0.0) > 0.0;
  const double o3col_du = o3col * kToDobson;
  const double d_temp = temp - clim[kTClim];
  const double d_o3col = o3col_du - clim[kO3colClim];
  out.ss_o3 = clim[kO3Clim] - (clim[kPmLClim] + d_o3col * clim[kdPmLdO3col] +
                               d_temp * clim[kdPmLdT]) /
                                  clim[kdPmLdO3];
  const double delta_o3 = (out.ss_o3 - o3_old) * (1.0 - exp(clim[kdPmLdO3] * dt));
  double o3_new = o3_old + delta_o3;
  out.do3 = delta_o3 / dt;
  out.do3_psc = 0.0;
  if (fabs(lats) > 40.0 && excess_chlorine && temp <= cf.psc_T) {
    // the largest zenith angle with sunlight at this height (tangent height
    // at sunset)
    const double l = 16.0 * log10(1e5 / pmid);
    const double max_sza = 90.0 + sqrt(l > 0.0 ? l : 0.0);
    if (sza * kRadToDeg <= max_sza) {
      const double r = cf.chlorine_loading / kChlorine1987;
      const double psc_loss = exp(-clim[kCariolle] * (r * r) * dt);
      o3_new = o3_old * psc_loss;
      out.do3_psc = (o3_new - o3_old) / dt;
    }
  }
  return o3_new;
}

// lin_strat_sfcsink: relax the lowest o3_lbl levels of a column towards
// o3_sfc, bottom up; returns the ozone mass loss [Tg/yr/m2].
__host__ __device__ __forceinline__ double surface_sink(int nlev, long ks, double dt,
                                                        const Conf &cf, const double *pdel,
                                                        double *o3) {
  double do3mass = 0.0;
  if (cf.o3_lbl > 0) {
    const double efactor = 1.0 - exp(-dt / cf.o3_tau);
    for (int k = nlev - 1; k > nlev - cf.o3_lbl - 1 && k >= 0; --k) {
      const long i = k * ks;
      const double o3l_old = o3[i];
      const double mass = pdel[i] * (1.0 / 9.80616);  // air mass [kg/m2]
      const double do3 = (cf.o3_sfc - o3l_old) * efactor;
      o3[i] = o3l_old + do3;
      do3mass += do3 * mass * 48. / (0.028966 * 1e3);
    }
  }
  return do3mass / dt * 1.0e-9 * (86400. * 365.0);
}

// One column.  Level k of a plane sits at p[k * ks].  o3 [vmr] is advanced
// in place; the optional outputs (NULL: not written) are the overhead column
// [mol/cm2] at every level and, above the tropopause (k < ltrop), the linoz
// and PSC tendencies [vmr/s] and the steady-state ozone [vmr] (0 below);
// sfcsink: the surface sink's ozone mass loss [Tg/yr/m2].  The outputs and
// o3 must not overlap the inputs (restrict): the loads of later levels are
// then free to run ahead of the stores of earlier ones.
__host__ __device__ __forceinline__ void column(
    int nlev, long ks, double dt, double sza, double rlats, int ltrop, double col0,
    const Conf &cf, const double *__restrict__ temp, const double *__restrict__ pmid,
    const double *__restrict__ pdel, const double *const clim[kNClim], double *__restrict__ o3,
    double *__restrict__ o3col_out, double *__restrict__ do3_out,
    double *__restrict__ do3_psc_out, double *__restrict__ ss_out, double *sfcsink) {
  const double lats = rlats * kRadToDeg;
  const double *__restrict__ c0 = clim[0];
  const double *__restrict__ c1 = clim[1];
  const double *__restrict__ c2 = clim[2];
  const double *__restrict__ c3 = clim[3];
  const double *__restrict__ c4 = clim[4];
  const double *__restrict__ c5 = clim[5];
  const double *__restrict__ c6 = clim[6];
  const double *__restrict__ c7 = clim[7];
  double above = 0.0;  // the column above level k (exclusive sum)
  // above the tropopause: column, solve, diagnostics
  for (int k = 0; k < ltrop; ++k) {
    const long i = k * ks;
    const double o3_old = o3[i];
    const double delta = kColumnFactor * pdel[i] * o3_old;
    const double o3col = col0 + above + 0.5 * delta;
    above += delta;
    const double c[kNClim] = {c0[i], c1[i], c2[i], c3[i], c4[i], c5[i], c6[i], c7[i]};
    LevelOut lo;
    o3[i] = solve_level(o3col, temp[i], sza, pmid[i], dt, lats, c, cf, o3_old, lo);
    if (o3col_out) o3col_out[i] = o3col;
    if (do3_out) do3_out[i] = lo.do3;
    if (do3_psc_out) do3_psc_out[i] = lo.do3_psc;
    if (ss_out) ss_out[i] = lo.ss_o3;
  }
  // below: the column only
  for (int k = ltrop; k < nlev; ++k) {
    const long i = k * ks;
    const double delta = kColumnFactor * pdel[i] * o3[i];
    if (o3col_out) o3col_out[i] = col0 + above + 0.5 * delta;
    above += delta;
    if (do3_out) do3_out[i] = 0.0;
    if (do3_psc_out) do3_psc_out[i] = 0.0;
    if (ss_out) ss_out[i] = 0.0;
  }
  const double sink = surface_sink(nlev, ks, dt, cf, pdel, o3);
  if (sfcsink) *sfcsink = sink;
}

// setinv_single_level: number densities [#/cm3] from T [K], p [Pa], the
// water vapour mass mixing ratio [kg/kg] and the offline species' volume
// mixing ratios (their invariant slots 7, 4, 5, 6).
__host__ __device__ __forceinline__ void invariants(double tfld, double pmid, double qv,
                                                    const double cnst_offline[4],
                                                    double inv[8]) {
  const double h2ovmr = qv * 0.028966 / 0.018016;
  const double m = 10.0 * pmid / (0.13806500000000001E-015 * tfld);
  inv[0] = m;
  inv[1] = 0.79 * m;
  inv[2] = 0.21 * m;
  inv[3] = h2ovmr * m;
  inv[7] = cnst_offline[0] * m;
  inv[4] = cnst_offline[1] * m;
  inv[5] = cnst_offline[2] * m;
  inv[6] = cnst_offline[3] * m;
}

}  // namespace lnz
}  // namespace mam4x
