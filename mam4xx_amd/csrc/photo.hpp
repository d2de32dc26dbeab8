// Table photolysis rates (table_photo, mo_photo.hpp:798-893) -- the
// photolysis_rate input of the gas-phase chemistry -- split in two passes:
//   * the column pass (one lane per column): cloud_mod (mo_photo.hpp:169-427),
//     the cloud optical depths integrated from the top and from the bottom,
//     giving each level's effective albedo and cloud multiplier.  Two level
//     walks: the top-down walk forms the cloud above a level and its
//     multiplier, the bottom-up walk the cloud below it and its albedo, with
//     the reference's serial running sums;
//   * the level pass: interpolate_rsf (mo_photo.hpp:498-645, find_index
//     :428-450, calc_sum_wght :452-496) and jlong (:647-796) for the one
//     reaction the rates feed (lng_indexer), then photo = cld_mult esfact
//     (photo + pht_alias_mult j): a cell's position in the tables
//     (level_pos), then the sum over wavelengths of the terms wn_term, in
//     wavelength order.
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

// One layer's values the walks read: its cloud optical depth, the depth
// weighted by its cloud fraction, and fac1 times the fraction, fac1 the
// thick-cloud term of the layer's own multiplier (mo_photo.hpp:196-206,
// :398-411).
struct Layer {
  double tau, ctau, f1c;
};
__host__ __device__ __forceinline__ Layer layer(double coschi, double cld, double lwc,
                                                double delp) {
  Layer l;
  l.tau = layer_tau(cld, lwc, delp);
  l.ctau = cld * l.tau;
  double fac1 = 0.0;
  if (layer_lwp(cld, lwc, delp) * kLwpToTau >= kTauMin) fac1 = 1.4 * coschi - 1.0;
  l.f1c = fac1 * cld;
  return l;
}

// Running sums over the layers on one side of a level (above from the top,
// below from the bottom) and the last normalised cloud fraction.
struct Side {
  double tau, ctau, prev;
};
// the cloud on that side of an interior level: above_tau / above_cld (or
// below_*) after the normalisation, the copy of the previous fraction where
// the depth is 0, the cloudy part's depth and the 5.0 threshold
__host__ __device__ __forceinline__ void side_cloud(Side &s, double &tau, double &cld) {
  tau = s.tau;
  cld = s.ctau;
  cld = tau != 0.0 ? cld / tau : s.prev;
  s.prev = cld;
  if (cld != 0.0) tau /= cld;
  if (tau < kTauMin) cld = 0.0;
}
__host__ __device__ __forceinline__ void side_add(Side &s, double tau, double ctau) {
  s.tau = tau + s.tau;
  s.ctau = ctau + s.ctau;
}
// the cloud multiplier from the cloud above, the effective albedo from the
// cloud below
__host__ __device__ __forceinline__ double multiplier(double coschi, double f1c, double a_tau,
                                                      double a_cld) {
  const double above_tra = kC1 / (kC2 + a_tau);
  const double fac2 = kmin(0.0, 1.6 * coschi * above_tra - 1.0);
  return kmax(.05, 1.0 + f1c + fac2 * a_cld);
}
__host__ __device__ __forceinline__ double albedo(double srf_alb, double b_tau, double b_cld) {
  const double below_tra = kC1 / (kC2 + b_tau);
  return srf_alb + b_cld * (1.0 - below_tra) * (1.0 - srf_alb);
}

// cloud_mod for one column whose level k sits at offset k * ks: the
// multiplier (top-down walk) into cld_mult, the effective albedo (bottom-up
// walk) into eff_alb, both at the same offsets.
__host__ __device__ __forceinline__ void cloud_column(int nlev, long ks, double zen,
                                                      const double *clouds, const double *lwc,
                                                      const double *delp, double srf_alb,
                                                      double *eff_alb, double *cld_mult) {
  const double coschi = coschi_of(zen);
  Side s{0.0, 0.0, 0.0};
  for (int k = 0; k < nlev; ++k) {
    const long i = k * ks;
    double a_tau = 0.0, a_cld = 0.0;
    if (k > 0) side_cloud(s, a_tau, a_cld);
    const Layer l = layer(coschi, clouds[i], lwc[i], delp[i]);
    cld_mult[i] = multiplier(coschi, l.f1c, a_tau, a_cld);
    side_add(s, l.tau, l.ctau);
  }
  s = Side{0.0, 0.0, 0.0};
  for (int k = nlev - 1; k >= 0; --k) {
    const long i = k * ks;
    double b_tau = 0.0, b_cld = 0.0;
    if (k < nlev - 1) side_cloud(s, b_tau, b_cld);
    eff_alb[i] = albedo(srf_alb, b_tau, b_cld);
    const Layer l = layer(coschi, clouds[i], lwc[i], delp[i]);
    side_add(s, l.tau, l.ctau);
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

// Where one level sits in the tables: everything of interpolate_rsf and
// jlong but the sum over wavelengths.
struct LevelPos {
  int pind, is, ial, ratl, ratu;  // rsf_tab indices: layers pind, pind - 1
  int t_index, pcase, pndx;       // xsqy row; pressure case 0 / 1 / 2 and layer
  double dels0, wrk0, d1l, d1u, d2, wght1, delp;
};
__host__ __device__ __forceinline__ LevelPos level_pos(const Table &t, const SzaPos &s,
                                                       double alb_in, double p_in, double t_in,
                                                       double colo3_in) {
  LevelPos q;
  q.is = s.is;
  q.dels0 = s.dels0;
  q.wrk0 = s.wrk0;
  q.ial = find_index(t.alb, t.numalb, alb_in, 0);
  if (p_in > t.press[0]) {
    q.pind = 1;
    q.wght1 = 1.0;
  } else if (p_in <= t.press[t.nump - 1]) {
    q.pind = t.nump - 1;
    q.wght1 = 0.0;
  } else {
    int iz;
    for (iz = 1; iz < t.nump; ++iz)
      if (t.press[iz] < p_in) break;
    iz = iz < t.nump - 1 ? iz : t.nump - 1;
    q.pind = iz > 1 ? iz : 1;
    q.wght1 = bound01((p_in - t.press[q.pind]) * t.del_p[q.pind - 1]);
  }
  const double v3ratu = colo3_in / t.colo3[q.pind - 1];
  q.ratu = find_index(t.o3rat, t.numcolo3, v3ratu, 0);
  double v3ratl;
  if (t.colo3[q.pind] != 0.0) {
    v3ratl = colo3_in / t.colo3[q.pind];
    q.ratl = find_index(t.o3rat, t.numcolo3, v3ratl, 0);
  } else {
    q.ratl = q.ratu;
    v3ratl = t.o3rat[q.ratu];
  }
  q.d2 = bound01((alb_in - t.alb[q.ial]) * t.del_alb[q.ial]);
  q.d1l = bound01((v3ratl - t.o3rat[q.ratl]) * t.del_o3rat[q.ratl]);
  q.d1u = bound01((v3ratu - t.o3rat[q.ratu]) * t.del_o3rat[q.ratu]);
  // the cross sections' temperature row and pressure interpolation
  q.t_index = (int)(kmin(201.0, kmax(t_in - 148.5, 1.0)) - 1.0);
  q.pndx = 0;
  q.delp = 0.0;
  if (p_in >= t.prs[0]) {
    q.pcase = 0;
  } else if (p_in <= t.prs[t.np_xs - 1]) {
    q.pcase = 1;
  } else {
    q.pcase = 2;
    for (int km = 1; km < t.np_xs; ++km)
      if (p_in >= t.prs[km]) {
        q.pndx = km - 1;
        q.delp = (t.prs[q.pndx] - p_in) * t.dprs[q.pndx];
        break;
      }
  }
  return q;
}

// Table memory order (mam4x_photo_table): rsf_tab [nw][nump][numsza]
// [numcolo3][numalb], xsqy [numj][nw][nt][np_xs].
struct RowMajor {
  const double *rsf, *xs;  // xs: xsqy of the reaction
  long s_al, s_v, s_s, s_z, s_w, x_p, x_t, x_w;
  __host__ __device__ RowMajor(const Table &t) {
    rsf = t.rsf_tab;
    s_al = 1;
    s_v = t.numalb;
    s_s = (long)t.numcolo3 * s_v;
    s_z = (long)t.numsza * s_s;
    s_w = (long)t.nump * s_z;
    x_p = 1;
    x_t = t.np_xs;
    x_w = (long)t.nt * x_t;
    xs = t.xsqy + (long)t.ind * t.nw * x_w;
  }
};

template <class Ix>
__host__ __device__ __forceinline__ double corner_sum(const Corners &c, const Ix &x,
                                                      const double *r) {
  return c.w000 * r[0] + c.w001 * r[x.s_al] + c.w010 * r[x.s_v] + c.w011 * r[x.s_v + x.s_al] +
         c.w100 * r[x.s_s] + c.w101 * r[x.s_s + x.s_al] + c.w110 * r[x.s_s + x.s_v] +
         c.w111 * r[x.s_s + x.s_v + x.s_al];
}

// one wavelength's term of jlong's sum: the cross section times the source
// function interpolated in zenith angle, ozone ratio, albedo and pressure
template <class Ix>
__host__ __device__ __forceinline__ double wn_term(const Table &t, const Ix &x,
                                                   const LevelPos &q, const Corners &cl,
                                                   const Corners &cu, int wn) {
  const double *rl = x.rsf + q.pind * x.s_z + q.is * x.s_s + q.ratl * x.s_v + q.ial * x.s_al +
                     wn * x.s_w;
  const double *ru = x.rsf + (q.pind - 1) * x.s_z + q.is * x.s_s + q.ratu * x.s_v +
                     q.ial * x.s_al + wn * x.s_w;
  const double psum_l = corner_sum(cl, x, rl);
  const double psum_u = corner_sum(cu, x, ru);
  const double rsf = (psum_l + q.wght1 * (psum_u - psum_l)) * t.etfphot[wn];
  const double *xr = x.xs + (long)q.t_index * x.x_t + wn * x.x_w;
  double xswk;
  if (q.pcase == 0) {
    xswk = xr[0];
  } else if (q.pcase == 1) {
    xswk = xr[(t.np_xs - 1) * x.x_p];
  } else {
    const double x0 = xr[q.pndx * x.x_p], x1 = xr[(q.pndx + 1) * x.x_p];
    xswk = x0 + q.delp * (x1 - x0);
  }
  return xswk * rsf;
}

// photolysis rate of reaction t.ind at one level, the wavelengths summed in
// order (interpolate_rsf + jlong), from the tables as given
__host__ __device__ __forceinline__ double level_rate(const Table &t, const SzaPos &s,
                                                      double alb_in, double p_in, double t_in,
                                                      double colo3_in) {
  const LevelPos q = level_pos(t, s, alb_in, p_in, t_in, colo3_in);
  const Corners cl = corners(q.dels0, q.wrk0, q.d1l, q.d2);
  const Corners cu = corners(q.dels0, q.wrk0, q.d1u, q.d2);
  const RowMajor x(t);
  double suma = 0.0;
  for (int wn = 0; wn < t.nw; ++wn) suma += wn_term(t, x, q, cl, cu, wn);
  return suma;
}

// table_photo's update of one level's rate
__host__ __device__ __forceinline__ double update(const Table &t, double photo, double cld_mult,
                                                  double esfact, double j) {
  return cld_mult * esfact * (photo + t.alias * j);
}

}  // namespace pht
}  // namespace mam4x
