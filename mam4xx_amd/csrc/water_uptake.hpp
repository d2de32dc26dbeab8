// Aerosol water uptake of one cell: modal_aero_water_uptake_dr with the wet
// density (water_uptake.hpp:362-505) -- the producer of the chain's
// dgncur_awet / wetdens inputs.  FP64; one lane per cell.
//
// The reference solves the Kohler quartic in closed form with Kokkos::complex
// arithmetic; the complex helpers below follow the published Kokkos
// algorithms (product; quotient scaled by |re| + |im| of the divisor; the
// libstdc++-style square root; pow through the polar form).
#pragma once

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#else
#include <cmath>
#endif

namespace mam4x {
namespace wu {

struct Cx {
  double re, im;
};
__host__ __device__ __forceinline__ Cx cx(double re, double im = 0.0) { return Cx{re, im}; }
__host__ __device__ __forceinline__ Cx operator+(Cx a, Cx b) { return cx(a.re + b.re, a.im + b.im); }
__host__ __device__ __forceinline__ Cx operator-(Cx a, Cx b) { return cx(a.re - b.re, a.im - b.im); }
__host__ __device__ __forceinline__ Cx operator-(Cx a) { return cx(-a.re, -a.im); }
__host__ __device__ __forceinline__ Cx operator*(Cx a, Cx b) {
  return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__host__ __device__ __forceinline__ Cx operator/(Cx a, double r) { return cx(a.re / r, a.im / r); }
__host__ __device__ __forceinline__ Cx operator/(Cx x, Cx y) {
  const double s = fabs(y.re) + fabs(y.im);
  const Cx xs = cx(x.re / s, x.im / s), ys = cx(y.re / s, -y.im / s);
  return (xs * ys) / (ys.re * ys.re + ys.im * ys.im);
}
__host__ __device__ __forceinline__ double cabs(Cx z) { return hypot(z.re, z.im); }
__host__ __device__ __forceinline__ Cx csqrt(Cx z) {
  if (z.re == 0.0) {
    const double t = sqrt(fabs(z.im) / 2);
    return cx(t, z.im < 0.0 ? -t : t);
  }
  const double t = sqrt(2 * (cabs(z) + fabs(z.re))), u = t / 2;
  return z.re > 0.0 ? cx(u, z.im / t) : cx(fabs(z.im) / t, z.im < 0.0 ? -u : u);
}
__host__ __device__ __forceinline__ Cx cpow(Cx z, double y) {
  const double rho = pow(cabs(z), y), th = y * atan2(z.im, z.re);
  return cx(rho * cos(th), rho * sin(th));
}

__host__ __device__ __forceinline__ double dmax(double x, double y) { return x < y ? y : x; }
__host__ __device__ __forceinline__ double dmin(double x, double y) { return y < x ? y : x; }
__host__ __device__ __forceinline__ double dcube(double x) { return x * x * x; }

constexpr double kEps = 1e-4;  // Water_Uptake::eps (water_uptake.hpp:38)
constexpr double kPi = 3.14159265358979323846264;

// makoh_quartic (water_uptake.hpp:124-169): roots of x^4 + p3 x^3 + p2 x^2 +
// p1 x + p0, then find_real_solution (:93-117): the smallest real root in
// [rdry (1 - eps), 1000 rdry]
__host__ __device__ __forceinline__ double quartic_wet_radius(double p3, double p2, double p1,
                                                              double p0, double rdry) {
  Cx r[4];
  const double qq = -p2 * p2 / 36.0 + (p3 * p1 - 4.0 * p0) / 12.0;
  const double rr = -dcube(p2 / 6.0) + p2 * (p3 * p1 - 4.0 * p0) / 48.0 +
                    (4.0 * p0 * p2 - p0 * p3 * p3 - p1 * p1) / 16.0;
  Cx crad = csqrt(cx(rr * rr + qq * qq * qq));
  Cx cb = cx(rr) - crad;
  if (cb.re == 0.0 && cb.im == 0.0) {  // insoluble particle
    const double c = cbrt(-p1);
    r[0] = r[1] = r[2] = r[3] = cx(c);
  } else {
    cb = cpow(cb, 1.0 / 3.0);
    const Cx cy = -cb + cx(qq) / cb + cx(p2 / 6.0);
    const Cx cb0 = csqrt(cy * cy - cx(p0));
    const Cx cb1 = (cx(p3) * cy - cx(p1)) / (cx(2.0) * cb0);
    cb = cx(p3 / 2.0) + cb1;
    crad = csqrt(cb * cb - cx(4.0) * (cy + cb0));
    r[0] = (-cb + crad) / 2.0;
    r[1] = (-cb - crad) / 2.0;
    cb = cx(p3 / 2.0) - cb1;
    crad = csqrt(cb * cb - cx(4.0) * (cy - cb0));
    r[2] = (-cb + crad) / 2.0;
    r[3] = (-cb - crad) / 2.0;
  }
  double rwet = 1000.0 * rdry;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const double xr = r[n].re, xi = r[n].im;
    if (fabs(xi) > fabs(xr) * kEps) continue;
    if (xr > rwet) continue;
    if (xr < rdry * (1.0 - kEps)) continue;
    if (xr != xr) continue;  // NaN
    rwet = xr;
  }
  return rwet;
}

// modal_aero_kohler (water_uptake.hpp:178-229): dry radius [m] -> wet [m]
__host__ __device__ __forceinline__ double kohler(double rdry_in, double hygro, double rh) {
  const double rhow = 1.0, surften = 76.0, mw = 18.0, tair = 273.0, ugascon = 8.3e7;
  const double aa = 2.0e4 * mw * surften / (ugascon * tair * rhow);
  const double rdry = rdry_in * 1.e6;  // microns
  const double vol = dcube(rdry);
  const double bb = vol * hygro;
  const double ss = dmax(1.e-10, dmin(1.0 - kEps, rh));
  const double slog = log(ss);
  const double p43 = -aa / slog, p42 = 0.0, p41 = bb / slog - vol, p40 = aa * vol / slog;
  const double pp = fabs(-bb / aa) / (rdry * rdry);
  double rwet;
  if (pp < kEps)  // approximate solution for small particles
    rwet = rdry * (1.0 + pp * (1.0 / 3.0) / (1.0 - slog * rdry / aa));
  else
    rwet = quartic_wet_radius(p43, p42, p41, p40, rdry);
  rwet = dmin(rwet, 30.0);  // upper bound based on 1 day lifetime
  return rwet * 1.e-6;
}

// E3SM species type (1-based) of the s-th species of each mode in
// mode_aero_species order, and density / hygroscopicity by type
// (get_e3sm_parameters, water_uptake.hpp:312-360).  Types 2 and 3 are unused
// by the four modes.
__host__ __device__ __forceinline__ int spectype(int m, int s) {
  constexpr int t[4][7] = {{1, 4, 5, 6, 8, 7, 9},
                           {1, 5, 7, 9, 0, 0, 0},
                           {8, 7, 1, 6, 4, 5, 9},
                           {4, 6, 9, 0, 0, 0, 0}};
  return t[m][s];
}
__host__ __device__ __forceinline__ double specdens(int t) {
  constexpr double d[9] = {1770.0, 0.1797693135 + 309, 0.1797693135 + 309, 1000.0, 1000.0,
                           1700.0, 1900.0,             2600.0,             1601.0};
  return d[t - 1];
}
__host__ __device__ __forceinline__ double spechygro(int t) {
  constexpr double h[9] = {0.507,           0.1797693135 + 309, 0.1797693135 + 309,
                           0.1000000083e-09, 0.14,              0.1000000013e-09,
                           1.16,            0.068,             0.1000000015};
  return h[t - 1];
}

// Goff & Gratch saturation vapour pressure over water and the saturation
// specific humidity (wv_sat_methods.hpp:13-133)
__host__ __device__ __forceinline__ double qsat_water(double t, double p) {
  const double tboil = 373.16, ten = 10, one = 1;
  const double es = pow(ten, -7.90298 * (tboil / t - one) + 5.02808 * log10(tboil / t) -
                                 1.3816e-7 * (pow(ten, 11.344 * (one - t / tboil)) - one) +
                                 8.1328e-3 * (pow(ten, -3.49149 * (tboil / t - one)) - one) +
                                 log10(1013.246)) *
                    ten * ten;
  const double epsilo = 0.018016 / 0.028966;
  return (p - es) <= 0.0 ? one : epsilo * es / (p - (one - epsilo) * es);
}

struct WaterOut {
  double dgncur_awet[4], qaerwat[4], wetdens[4];
};

// One cell.  qaer(m, s): interstitial mass mixing ratio [kg/kg] of the s-th
// species of mode m; vol_from_dgn[m] = exp(4.5 ln^2 sigma_m), the modes'
// lognormal volume factor (a host-computed constant).
template <class Q>
__host__ __device__ __forceinline__ void water_uptake_cell(Q qaer, const int nspec[4],
                                                           const double vol_from_dgn[4],
                                                           double temp,
                                                           double pmid, double qv, double cldn,
                                                           const double dgncur_a[4],
                                                           WaterOut &o) {
  constexpr double rhcrystal = 0.35, rhdeliques = 0.8;  // aero_modes.hpp:76-77
  // modal_aero_water_uptake_rh_clearair (:259-285)
  const double qs = qsat_water(temp, pmid);
  double rh = qs > qv ? qv / qs : 0.98;
  rh = dmax(0.0, dmin(0.98, rh));
  if (cldn < 1.0) rh = (rh - cldn) / (1.0 - cldn);
  rh = dmax(rh, 0.0);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    // modal_aero_water_uptake_dryaer (:362-437)
    double dryvolmr = 0.0, maer = 0.0, hygro = 0.0;
    const double specdens_1 = specdens(spectype(m, 0));
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      if (s >= nspec[m]) continue;
      const int t = spectype(m, s);
      const double raer = qaer(m, s);
      const double vol_tmp = raer / specdens(t);
      maer = maer + raer;
      dryvolmr += vol_tmp;
      hygro += vol_tmp * spechygro(t);
    }
    hygro = dryvolmr > 1e-30 ? hygro / dryvolmr : spechygro(spectype(m, 0));
    const double v2ncur_a = 1.0 / ((kPi / 6.0) * dcube(dgncur_a[m]) * vol_from_dgn[m]);
    const double naer = dryvolmr * v2ncur_a;
    const double drydens = maer > 1e-31 ? maer / dryvolmr : 1.0;
    const double dryvol = 1.0 / v2ncur_a;
    const double drymass = drydens * dryvol;
    const double dryrad = cbrt(dryvol / (kPi * 4.0 / 3.0));
    // modal_aero_water_uptake_wetaer (:241-257)
    const double hystfac = 1.0 / dmax(1.0e-5, rhdeliques - rhcrystal);
    double wetrad = dmax(kohler(dryrad, hygro, rh), dryrad);
    double wetvol = dmax((kPi * 4.0 / 3.0) * dcube(wetrad), dryvol);
    double wtrvol = dmax(wetvol - dryvol, 0.0);
    if (rh < rhcrystal) {
      wetrad = dryrad;
      wetvol = dryvol;
      wtrvol = 0.0;
    } else if (rh < rhdeliques) {
      wtrvol = dmax(wtrvol * hystfac * (rh - rhcrystal), 0.0);
      wetvol = dryvol + wtrvol;
      wetrad = cbrt(wetvol / (4.0 / 3.0 * kPi));
    }
    o.dgncur_awet[m] = dgncur_a[m] * (wetrad / dryrad);
    o.qaerwat[m] = 1.0e3 * naer * wtrvol;
    // modal_aero_water_uptake_wetdens (:67-90)
    o.wetdens[m] = wetvol > 1.0e-30 ? (drymass + 1.0e3 * wtrvol) / wetvol : specdens_1;
  }
}

}  // namespace wu
}  // namespace mam4x
