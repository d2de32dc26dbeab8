// Washout rates of one column: mo_sethet::sethet_detail (mo_sethet.hpp:217-
// 500) with the wrapper's species (:549-556): H2O2 (gas_pcnst 1), H2SO4 (2)
// and SO2 (3).
//   * the column's precipitation, rescaled to its net (positive and
//     negative) total, which removes point storms (calc_precip_rescale);
//   * the levels below the tropopause proxy ktop (100 hPa, 300 hPa poleward
//     of 60 degrees; find_ktop);
//   * the washout of H2O2 by the rain of each raining level onto every level
//     below it, where the column is not saturated (gas_washout), and the
//     washout lifetime that gives the below-cloud rate;
//   * the in-cloud rates from the effective Henry's-law constants.
// SO2 takes the H2O2 rate, as in the reference; its own washout there is
// computed and never used, so it is not computed here.  FP64; one lane per
// column.  The H2O2, H2SO4 and SO2 output planes serve as the column's
// scratch (washed gas, layer depths, saturation terms) until their final
// values are written.
#pragma once

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#else
#include <cmath>
// plain C++ compilers see the device annotations as no-ops (the test suite
// builds this file for the CPU: tests/cpp/sethet_host.cpp)
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
namespace sh {

constexpr double kMassH2O = 0.018016 * 1000;  // [g/mol]
constexpr double kAvo = 6.022214076e23;
constexpr double kConst0 = 1.380649e-23 * 1.e7 * 1.0e-6;  // [atm/K/cm3]
constexpr double kAvo2 = kAvo * 1.0e-3 * 1.0e-6;
constexpr double kXrm = .189;  // mean rain-drop diameter [cm]
constexpr double kXum = 748.0;  // mean rain-drop terminal velocity [cm/s]
constexpr double kMassAir = 29.0;
constexpr double kSatf = .016;  // in-cloud saturation factor (H2O2, HNO3)
constexpr double kRga = 1.0 / 9.80616;

// calc_het_rates (mo_sethet.hpp:36-52)
__host__ __device__ __forceinline__ double het_rate(double rain, double xhen, double tmp,
                                                    double work1, double work2) {
  const double h2o_mol = 1.0e3 / kMassH2O;
  const double r = rain / (h2o_mol * (work1 + 1.0 / (xhen * work2)));
  return kSatf * (r > 0.0 ? r : 0.0) + tmp;
}

// One column; level k of an input or output plane at [k * ks], level 0 the
// model top.  xkgm: the drops' mass-transfer coefficient (host-computed,
// mo_sethet.hpp:330-331).  zmid is the reference's argument of that name
// (its formulas treat it as km).  het1/het2/het3: the H2O2, H2SO4 and SO2
// rate planes [1/s], written at every level.
__host__ __device__ __forceinline__ void column(int nlev, long ks, double rlat, double phis,
                                                double delt, double xkgm, const double *press,
                                                const double *zmid, const double *tfld,
                                                const double *cmfdqr, const double *nrain,
                                                const double *nevapr, const double *inv,
                                                const double *qh2o2, double *het1,
                                                double *het2, double *het3) {
  // find_ktop: the lowest level above the pressure limit
  const double p_limit = fabs(rlat) > 60.0 * (3.14159265358979323846264 / 180.0) ? 300.0e2
                                                                                 : 100.0e2;
  int ktop = 0;
  for (int k = nlev - 1; k >= 0; --k)
    if (press[k * ks] < p_limit) {
      ktop = k;
      break;
    }
  // calc_precip_rescale's column totals
  double total_rain = 0.0, total_pos = 0.0;
  for (int k = 0; k < nlev; ++k) {
    const long i = k * ks;
    const double p = cmfdqr[i] + nrain[i] - nevapr[i];
    total_rain += p;
    total_pos += p < 0.0 ? 0.0 : p;
  }
  auto precip = [&](long i) {
    double p = cmfdqr[i] + nrain[i] - nevapr[i];
    if (p < 0.0) p = 0.0;
    return total_rain <= 0.0 ? 0.0 : p * total_rain / total_pos;
  };
  const double zsurf = 1.0e-3 * phis * kRga;  // [km]
  // scratch: het1 = H2O2 gas being washed, het2 = its saturation term,
  // het3 = layer depth [cm]
  for (int k = ktop; k < nlev; ++k) {
    const long i = k * ks;
    const double t_factor = (298.0 - tfld[i]) / (298.0 * tfld[i]);
    const double xhen_h2o2 = 7.45e4 * exp(6620.0 * t_factor);
    het1[i] = qh2o2[i] * inv[i];
    het2[i] = 1.0 / (xhen_h2o2 * kConst0 * tfld[i]);
    het3[i] = k < nlev - 1 ? fabs((zmid[i] - zmid[i + ks]) * 1.0e5)
                           : fabs((zmid[i] - zsurf) * 1.0e5);
  }
  // gas_washout of H2O2: the rain of level kk onto the levels k >= kk
  for (int kk = ktop; kk < nlev; ++kk) {
    const long ii = kk * ks;
    const double pr = precip(ii);
    const double rain = kMassAir * pr * inv[ii] / kMassH2O;
    if (rain == 0.0) continue;
    const double xliq = pr * delt * inv[ii] / kAvo * kMassAir * 1.0e6;  // [g/m3]
    double accumulator = 0.0;
    for (int k = kk; k < nlev; ++k) {
      const long i = k * ks;
      const double xgas = het1[i];
      const double xca = 6.0 * xkgm * xgas / (kXrm * kXum) * het3[i] * xliq * 1.0e-6;
      accumulator += xca;
      const double xeqca = xgas / (xliq * kAvo2 + het2[i]) * xliq * kAvo2;
      if (accumulator < xeqca) het1[i] = xgas - xca > 0.0 ? xgas - xca : 0.0;
    }
  }
  // the below-cloud lifetime and the in-cloud rates
  for (int kk = ktop; kk < nlev; ++kk) {
    const long i = kk * ks;
    const double pr = precip(i);
    const double rain = kMassAir * pr * inv[i] / kMassH2O;
    double stay = 1.0;
    if (rain != 0.0) {
      stay = ((zmid[i] - zsurf) * 1.0e5) / (kXum * delt);
      stay = stay < 1.0 ? stay : 1.0;
    }
    const double xdtm = het3[i] / kXum;
    const double xh2o2 = qh2o2[i] * inv[i];
    const double xxx2 = (xh2o2 - het1[i]);
    const double yh2o2 = xxx2 != 0.0 ? xh2o2 / xxx2 * xdtm : 1.0e29;
    const double lt = 1.0 / yh2o2;
    const double tmp = (lt > 0.0 ? lt : 0.0) * stay;
    if (rain <= 0.0) {
      het1[i] = het2[i] = het3[i] = 0.0;
      continue;
    }
    const double xliq = pr * delt * inv[i] / kAvo * kMassAir * 1.0e6;
    const double work1 = kAvo2 * xliq, work2 = kConst0 * tfld[i];
    const double t_factor = (298.0 - tfld[i]) / (298.0 * tfld[i]);
    const double xhen_h2o2 = 7.45e4 * exp(6620.0 * t_factor);
    const double xhen_hno3 = 2.1e5 * exp(8700.0 * t_factor) * (1.0 + 15.4 / 1.0e-5);
    het1[i] = het_rate(rain, xhen_h2o2, tmp, work1, work2);
    het3[i] = het1[i];
    het2[i] = het_rate(rain, xhen_hno3, 0.0, work1, work2);
  }
  // no washout at or above ktop
  for (int k = 0; k <= ktop && k < nlev; ++k) het1[k * ks] = het2[k * ks] = het3[k * ks] = 0.0;
}

}  // namespace sh
}  // namespace mam4x
