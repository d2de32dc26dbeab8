// Table photolysis rates (table_photo, mo_photo.hpp:798-893) -- the
// photolysis_rate input of the gas-phase chemistry -- split in two passes:
//   * the column pass (one lane per column): cloud_mod (mo_photo.hpp:169-427),
//     the cloud optical depths integrated from the top and from the bottom,
//     giving each level's effective albedo and cloud multiplier.  Two level
//     walks: the top-down walk forms the cloud above a level and its
//     multiplier, the bottom-up walk the cloud below it and its albedo, with
//     the reference's serial running sums;
//   * the level pass (one lane per cell): interpolate_rsf (mo_photo.hpp:
//     498-645, find_index :428-450, calc_sum_wght :452-496) and jlong
//     (:647-796) for the one reaction the rates feed (lng_indexer), then
//     photo = cld_mult esfact (photo + pht_alias_mult j).
// Tables row-major: rsf_tab [nw][nump][numsza][numcolo3][numalb], xsqy
// [numj][nw][nt][np_xs].  FP64; every sum in the reference's order.
#pragma once

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#else
#include <cmath>
// plain C++ compilers see the device annotations as no-ops (the test suite
// builds this file for the CPU: tests/cpp/photo_host.cpp)
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
namespace pht {

constexpr double kPi = 3.14159265358979323846264;
constexpr double kRadToDeg = 180.0 / kPi;
constexpr double kMaxZenith = 88.85;  // [deg]: night at and beyond
constexpr double kPaToMb = 1.e-2;
constexpr int kNt = 201;  // temperature index range of the cross sections

struct Table {
  int nw, nt, np_xs, numj, nump, numsza, numcolo3, numalb;
  const double *xsqy, *sza, *del_sza, *alb, *del_alb, *press, *del_p, *colo3, *o3rat,
      *del_o3rat, *etfphot, *prs, *dprs, *rsf_tab;
  double alias;  // pht_alias_mult_1 of the reaction
  int ind;       // its lng_indexer entry (< 0: the rates are left as they are)
};

__host__ __device__ __forceinline__ double kmax(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ double kmin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ double bound01(double x) { return kmax(0.0, kmin(1.0, x)); }

// the solar zenith angle [deg]; daylight in [0, 88.85)
__host__ __device__ __forceinline__ double sza_deg(double zen) { return zen * kRadToDeg; }
__host__ __device__ __forceinline__ bool daylight(double sza) {
  return sza >= 0.0 && sza < kMaxZenith;
}

// mo_photo.hpp:428-450
__host__ __device__ __forceinline__ int find_index(const double *v, int len, double vmin,
                                                   int idx0) {
  for (int ii = 0; ii < len; ++ii)
    if (v[ii] > vmin) {
      const int m = ii < len - 1 ? ii : len - 1;
      return m - 1 > 0 ? m - 1 : 0;
    }
  return idx0;
}

// ---------------------------------------------------------------- column
constexpr double kRgrav = 1.0 / 9.80616;
constexpr double kLwpToTau = .155;
constexpr double kTauMin = 5.0;
constexpr double kC1 = 11.905, kC2 = 9.524;

// a layer's liquid water path [g/m2] and cloud optical depth
__host__ __device__ __forceinline__ double layer_lwp(double cld, double lwc, double delp) {
  return cld != 0.0 ? kRgrav * lwc * delp * 1000.0 / cld : 0.0;
}
__host__ __device__ __forceinline__ double layer_tau(double cld, double lwc, double delp) {
  if (cld == 0.0) return 0.0;
  const double del_lwp = kRgrav * lwc * delp * 1000.0 / cld;
#if defined(__HIP_DEVICE_COMPILE__)
  return del_lwp * kLwpToTau * ::pow(cld, 1.5);
#else
  return del_lwp * kLwpToTau * std::pow(cld, 1.5);
#endif
}

__host__ __device__ __forceinline__ double coschi_of(double zen) {
#if defined(__HIP_DEVICE_COMPILE__)
  return kmax(::cos(zen), 0.5);
#else
  return kmax(std::cos(zen), 0.5);
#endif
}

// cloud_mod for one column whose level k sits at offset k * ks: the
// multiplier (top-down walk) into cld_mult, the effective albedo (bottom-up
// walk) into eff_alb, both at the same offsets.
__host__ __device__ __forceinline__ void cloud_column(int nlev, long ks, double zen,
                                                      const double *clouds, const double *lwc,
                                                      const double *delp, double srf_alb,
                                                      double *eff_alb, double *cld_mult) {
  const double coschi = coschi_of(zen);
  // top down: the cloud above level k (above_tau / above_cld of the
  // reference, after its normalisation and the optical-depth threshold)
  double sum_tau = 0.0, sum_ctau = 0.0, prev_cld = 0.0;
  for (int k = 0; k < nlev; ++k) {
    const long i = k * ks;
    double a_tau = 0.0, a_cld = 0.0;
    if (k > 0) {
      a_tau = sum_tau;
      a_cld = sum_ctau;
      a_cld = a_tau != 0.0 ? a_cld / a_tau : prev_cld;
      prev_cld = a_cld;
      if (a_cld != 0.0) a_tau /= a_cld;
      if (a_tau < kTauMin) a_cld = 0.0;
    }
    const double cld = clouds[i], w = lwc[i], dp = delp[i];
    const double d_tau = layer_tau(cld, w, dp);
    const double del_lwp = layer_lwp(cld, w, dp);
    double fac1 = 0.0;
    if (del_lwp * kLwpToTau >= kTauMin) fac1 = 1.4 * coschi - 1.0;
    const double above_tra = kC1 / (kC2 + a_tau);
    const double fac2 = kmin(0.0, 1.6 * coschi * above_tra - 1.0);
    cld_mult[i] = kmax(.05, 1.0 + fac1 * cld + fac2 * a_cld);
    sum_tau = d_tau + sum_tau;
    sum_ctau = cld * d_tau + sum_ctau;
  }
  // bottom up: the cloud below level k
  sum_tau = 0.0;
  sum_ctau = 0.0;
  prev_cld = 0.0;
  for (int k = nlev - 1; k >= 0; --k) {
    const long i = k * ks;
    double b_tau = 0.0, b_cld = 0.0;
    if (k < nlev - 1) {
      b_tau = sum_tau;
      b_cld = sum_ctau;
      b_cld = b_tau != 0.0 ? b_cld / b_tau : prev_cld;
      prev_cld = b_cld;
      if (b_cld != 0.0) b_tau /= b_cld;
      if (b_tau < kTauMin) b_cld = 0.0;
    }
    const double below_tra = kC1 / (kC2 + b_tau);
    eff_alb[i] = srf_alb + b_cld * (1.0 - below_tra) * (1.0 - srf_alb);
    const double cld = clouds[i];
    const double d_tau = layer_tau(cld, lwc[i], delp[i]);
    sum_tau = d_tau + sum_tau;
    sum_ctau = cld * d_tau + sum_ctau;
  }
}

// ---------------------------------------------------------------- level
// the zenith-angle interpolation, the same for every level of a column
struct SzaPos {
  int is;
  double dels0, wrk0;
};
__host__ __device__ __forceinline__ SzaPos sza_pos(const Table &t, double sza_in) {
  SzaPos s;
  s.is = find_index(t.sza, t.numsza, sza_in, 0);
  s.dels0 = bound01((sza_in - t.sza[s.is]) * t.del_sza[s.is]);
  s.wrk0 = 1.0 - s.dels0;
  return s;
}

// calc_sum_wght's eight corner weights
struct Corners {
  double w000, w001, w010, w011, w100, w101, w110, w111;
};
__host__ __device__ __forceinline__ Corners corners(double d0, double wrk0, double d1,
                                                    double d2) {
  Corners c;
  double wrk1 = (1.0 - d1) * (1.0 - d2);
  c.w000 = wrk0 * wrk1;
  c.w100 = d0 * wrk1;
  wrk1 = (1.0 - d1) * d2;
  c.w001 = wrk0 * wrk1;
  c.w101 = d0 * wrk1;
  wrk1 = d1 * (1.0 - d2);
  c.w010 = wrk0 * wrk1;
  c.w110 = d0 * wrk1;
  wrk1 = d1 * d2;
  c.w011 = wrk0 * wrk1;
  c.w111 = d0 * wrk1;
  return c;
}

// r: the rsf_tab entry (wn, iz, is, iv, ial); sv / ss: the o3-ratio and
// zenith-angle strides
__host__ __device__ __forceinline__ double corner_sum(const Corners &c, const double *r, long sv,
                                                      long ss) {
  return c.w000 * r[0] + c.w001 * r[1] + c.w010 * r[sv] + c.w011 * r[sv + 1] +
         c.w100 * r[ss] + c.w101 * r[ss + 1] + c.w110 * r[ss + sv] +
         c.w111 * r[ss + sv + 1];
}

// photolysis rate of reaction t.ind at one level: sza position s, effective
// albedo alb_in, pressure p_in [hPa], temperature t_in [K], o3 column
// colo3_in (interpolate_rsf + jlong)
__host__ __device__ __forceinline__ double level_rate(const Table &t, const SzaPos &s,
                                                      double alb_in, double p_in, double t_in,
                                                      double colo3_in) {
  const int ial = find_index(t.alb, t.numalb, alb_in, 0);
  int pind;
  double wght1;
  if (p_in > t.press[0]) {
    pind = 1;
    wght1 = 1.0;
  } else if (p_in <= t.press[t.nump - 1]) {
    pind = t.nump - 1;
    wght1 = 0.0;
  } else {
    int iz;
    for (iz = 1; iz < t.nump; ++iz)
      if (t.press[iz] < p_in) break;
    iz = iz < t.nump - 1 ? iz : t.nump - 1;
    pind = iz > 1 ? iz : 1;
    wght1 = bound01((p_in - t.press[pind]) * t.del_p[pind - 1]);
  }
  const double v3ratu = colo3_in / t.colo3[pind - 1];
  const int ratindu = find_index(t.o3rat, t.numcolo3, v3ratu, 0);
  double v3ratl;
  int ratindl;
  if (t.colo3[pind] != 0.0) {
    v3ratl = colo3_in / t.colo3[pind];
    ratindl = find_index(t.o3rat, t.numcolo3, v3ratl, 0);
  } else {
    ratindl = ratindu;
    v3ratl = t.o3rat[ratindu];
  }
  const double d2 = bound01((alb_in - t.alb[ial]) * t.del_alb[ial]);
  const double d1l = bound01((v3ratl - t.o3rat[ratindl]) * t.del_o3rat[ratindl]);
  const double d1u = bound01((v3ratu - t.o3rat[ratindu]) * t.del_o3rat[ratindu]);
  const Corners cl = corners(s.dels0, s.wrk0, d1l, d2);
  const Corners cu = corners(s.dels0, s.wrk0, d1u, d2);
  // the cross-section interpolation in pressure and its temperature row
  const int t_index = (int)(kmin(201.0, kmax(t_in - 148.5, 1.0)) - 1.0);
  int pndx = 0, pcase;
  double delp = 0.0;
  if (p_in >= t.prs[0]) {
    pcase = 0;
  } else if (p_in <= t.prs[t.np_xs - 1]) {
    pcase = 1;
  } else {
    pcase = 2;
    for (int km = 1; km < t.np_xs; ++km)
      if (p_in >= t.prs[km]) {
        pndx = km - 1;
        delp = (t.prs[pndx] - p_in) * t.dprs[pndx];
        break;
      }
  }
  const long sv = t.numalb, ss = (long)t.numcolo3 * sv, sz = (long)t.numsza * ss,
             sw = (long)t.nump * sz;
  const double *rl = t.rsf_tab + pind * sz + s.is * ss + ratindl * sv + ial;
  const double *ru = t.rsf_tab + (pind - 1) * sz + s.is * ss + ratindu * sv + ial;
  const long xw = (long)t.nt * t.np_xs;
  const double *x = t.xsqy + (long)t.ind * t.nw * xw + (long)t_index * t.np_xs;
  double suma = 0.0;
  for (int wn = 0; wn < t.nw; ++wn) {
    const double psum_l = corner_sum(cl, rl + wn * sw, sv, ss);
    const double psum_u = corner_sum(cu, ru + wn * sw, sv, ss);
    const double rsf = (psum_l + wght1 * (psum_u - psum_l)) * t.etfphot[wn];
    const double *xr = x + wn * xw;
    double xswk;
    if (pcase == 0)
      xswk = xr[0];
    else if (pcase == 1)
      xswk = xr[t.np_xs - 1];
    else
      xswk = xr[pndx] + delp * (xr[pndx + 1] - xr[pndx]);
    suma += xswk * rsf;
  }
  return suma;
}

// table_photo's update of one level's rate
__host__ __device__ __forceinline__ double update(const Table &t, double photo, double cld_mult,
                                                  double esfact, double j) {
  return cld_mult * esfact * (photo + t.alias * j);
}

}  // namespace pht
}  // namespace mam4x
