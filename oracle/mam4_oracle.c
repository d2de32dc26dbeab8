/*
 * CPU FP64 oracle for the MAM4 aerosol-microphysics chain -- TEST
 * INFRASTRUCTURE ONLY (see mam4_oracle.h).
 *
 * Every function names the reference file:line it restates.  The reference
 * is eagles-project/mam4xx; paths are relative to its src/mam4xx/.  Quirks of
 * the reference are reproduced on purpose and flagged "QUIRK".
 *
 * min/max follow Kokkos::min/max exactly (ternaries, Kokkos_MinMax.hpp):
 *   max(a,b) = (a < b) ? b : a      min(a,b) = (b < a) ? b : a
 * which matters for NaN operands.
 */
#include "mam4_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "whitby_coag_tables.inc"

static inline double kmax(double a, double b) { return (a < b) ? b : a; }
static inline double kmin(double a, double b) { return (b < a) ? b : a; }
static inline double sq(double x) { return x * x; }
static inline double cube(double x) { return x * x * x; }
/* utils.hpp:23-26 */
static inline double min_max_bound(double lo, double hi, double x) {
  return kmax(lo, kmin(hi, x));
}


/* ---------------- libm-rounding noise (conditioning probe) ----------------
 * When enabled (orc_set_libm_noise(seed != 0)), every exp/log/pow/cbrt/erfc/
 * sqrt result is moved by -1, 0 or +1 ulp at random (per-cell seeded xorshift).
 * Running the oracle under a few seeds measures, cell by cell, how far the
 * chain's outputs move when only the last bit of its transcendental
 * functions changes -- the spread two correct libm implementations (glibc on
 * the host, the GPU's device libm) can legitimately produce. */
static unsigned long long g_noise_seed = 0;
static _Thread_local unsigned long long t_rng = 0;
void orc_set_libm_noise(unsigned long long seed) { g_noise_seed = seed; }
static void noise_cell(long cell) {
  unsigned long long z = g_noise_seed + 0x9E3779B97F4A7C15ULL * (unsigned long long)(cell + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  t_rng = (z ^ (z >> 31)) | 1ULL;
}
static inline double nz(double v) {
  if (!g_noise_seed) return v;
  t_rng ^= t_rng << 13;
  t_rng ^= t_rng >> 7;
  t_rng ^= t_rng << 17;
  const unsigned r = (unsigned)(t_rng % 3ULL);
  if (r == 0 || !isfinite(v)) return v;
  return nextafter(v, r == 1 ? INFINITY : -INFINITY);
}
#define M_exp(x) nz(exp(x))
#define M_log(x) nz(log(x))
#define M_pow(x, y) nz(pow((x), (y)))
#define M_cbrt(x) nz(cbrt(x))
#define M_erfc(x) nz(erfc(x))
#define M_sqrt(x) nz(sqrt(x))
#define M_log10(x) nz(log10(x))
#define M_cos(x) nz(cos(x))

/* ---------------- constants (mam4_constants.hpp, aero_modes.hpp) -------- */
static const double PI = 3.14159265358979323846264;
#define PI_SIXTH (PI / 6)
static const double MW_H2O = 0.018016;
static const double MW_DRY_AIR = 0.028966;
static const double FREEZING_PT_H2O = 273.15;
static const double PRESSURE_STP = 101325.0;

enum { NACC = 0, NAIT = 1, NCOA = 2, NPCA = 3 };
enum { A_SOA = 0, A_SO4 = 1, A_POM = 2, A_BC = 3, A_NCL = 4, A_DST = 5, A_MOM = 6 };

/* aero_modes.hpp:78-93 */
static const double MODE_DMIN[4] = {5.35e-8, 8.7e-9, 1e-6, 1e-8};
static const double MODE_DNOM[4] = {1.1e-7, 2.6e-8, 2e-6, 5e-8};
static const double MODE_DMAX[4] = {4.4e-7, 5.2e-8, 4e-6, 1e-7};
static const double MODE_SIGMA[4] = {1.8, 1.6, 1.8, 1.60000002384186};
/* aero_modes.hpp:154-161: densities by AeroId */
static const double AER_DENS[7] = {1000.0, 1770.0, 1000.0, 1700.0,
                                   1900.0, 2600.0, 1601.0};
/* aero_modes.hpp:214-243: species in each mode (AeroId) */
static const int NSPEC_MODE[4] = {7, 4, 7, 3};
static const int MODE_SPEC[4][7] = {
    {A_SO4, A_POM, A_SOA, A_BC, A_DST, A_NCL, A_MOM},
    {A_SO4, A_SOA, A_NCL, A_MOM, -1, -1, -1},
    {A_DST, A_NCL, A_SO4, A_BC, A_POM, A_SOA, A_MOM},
    {A_POM, A_BC, A_MOM, -1, -1, -1, -1}};

/* gas_chem_mechanism.hpp:31-37 */
static const double ADV_MASS[31] = {
    47.998200,  34.013600, 98.078400,     64.064800, 62.132400, 12.011000,
    115.107340, 12.011000, 12.011000,     12.011000, 135.064039, 58.442468,
    250092.672000, 1.007400, 115.107340,  12.011000, 58.442468, 250092.672000,
    1.007400,   135.064039, 58.442468,    115.107340, 12.011000, 12.011000,
    12.011000,  250092.672000, 1.007400,  12.011000, 12.011000, 250092.672000,
    1.007400};

/* mam4_amicphys.hpp:35-63 */
static const int NUMPTR[4] = {13, 18, 26, 30};
static const int LMASSPTR[7][4] = {{6, 14, 19, 27},  {7, 15, 20, 28},
                                   {8, 16, 21, 29},  {9, 17, 22, -6},
                                   {10, -6, 23, -6}, {11, -6, 24, -6},
                                   {12, -6, 25, -6}};
enum { LMAPCC_NUL = 0, LMAPCC_GAS = 1, LMAPCC_AER = 2, LMAPCC_NUM = 3 };
static const int LMAPCC_ALL[31] = {0, 0, 1, 0, 0, 1, 2, 2, 2, 2, 2,
                                   2, 2, 3, 2, 2, 2, 2, 3, 2, 2, 2,
                                   2, 2, 2, 2, 3, 2, 2, 2, 3};
/* mam4_amicphys.hpp:66-69, 117-123 */
static const int LMAP_GAS[2] = {5, 2};
static const int LMAP_AER[7][4] = {{8, 15, 24, -1},  {6, 14, 21, -1},
                                   {7, -1, 23, 27},  {9, -1, 22, 28},
                                   {11, 16, 20, -1}, {10, -1, 19, -1},
                                   {12, 17, 25, 29}};
/* mam4_amicphys.hpp:131-135 */
static const double FCVT_AER[7] = {8.000000000000000E-002, 1, 8.000000000000000E-002,
                                   1, 1, 1, 1};
#define MAXSUB 3
#define IGAS_H2SO4 1
#define IGAS_NH3 (-1)

/* mam4_amicphys.hpp:141-154 */
static double fcvt_gas(int igas) {
  const double mwuse_soa = 150;
  const double mw_gas = ADV_MASS[LMAP_GAS[igas]];
  double denom = mw_gas;
  if (igas < 1) denom = mwuse_soa;
  return mw_gas / denom;
}
static const double FCVT_NUM = 1.0;
static double fcvt_wtr(void) { return MW_DRY_AIR / MW_H2O; }

/* aging.hpp:130-137 / mam4_amicphys.hpp:1512-1518 (AeroId order) */
static const double MASS_2_VOL[7] = {0.15,
                                     6.4971751412429377e-002,
                                     0.15,
                                     7.0588235294117650e-003,
                                     3.0789473684210526e-002,
                                     5.1923076923076926e-002,
                                     156.20986883198000};

/* ---------------- conversions.hpp ---------------------------------------- */
/* conversions.hpp:299-305 */
double orc_mean_particle_diameter_from_volume(double vol, double sigma) {
  const double pio6 = PI_SIXTH;
  return M_cbrt(vol / pio6) * M_exp(-1.5 * sq(M_log(sigma)));
}
/* conversions.hpp:312-317 */
double orc_mean_particle_volume_from_diameter(double dgn, double sigma) {
  const double pio6 = PI_SIXTH;
  return cube(dgn) * M_exp(4.5 * sq(M_log(sigma))) * pio6;
}
/* conversions.hpp:148-180 */
static double vapor_saturation_pressure_magnus(double T, double P) {
  const double e0 = 610.94, exp_num = 17.625, exp_den = 234.04;
  const double celsius = T - FREEZING_PT_H2O;
  const double ew = e0 * M_exp(exp_num * celsius / (exp_den + celsius));
  return 1.00071 * M_exp(4.5e-8 * P) * ew;
}

double orc_magnus_es(double T, double P) { return vapor_saturation_pressure_magnus(T, P); }

/* ======================= subarea set-up (mam4_amicphys.hpp) ============== */

/* mam4_amicphys.hpp:241-317 */
static void setup_subareas(double cld, int *nsubarea, int *ncldy_subarea,
                           int *jclea, int *jcldy, int iscldy[MAXSUB],
                           double afracsub[MAXSUB], double *fclea,
                           double *fcldy) {
  const double fcld_locutoff = 1.0e-5;
  const double fcld_hicutoff = 0.999;
  if (cld < fcld_locutoff) {
    *fcldy = 0;
    *nsubarea = 1;
    *ncldy_subarea = 0;
    *jclea = 1;
    *jcldy = 0;
  } else if (cld > fcld_hicutoff) {
    *fcldy = 1;
    *nsubarea = 1;
    *ncldy_subarea = 1;
    *jclea = 0;
    *jcldy = 1;
  } else {
    *fcldy = cld;
    *nsubarea = 2;
    *ncldy_subarea = 1;
    *jclea = 1;
    *jcldy = 2;
  }
  *fclea = 1.0 - *fcldy;
  for (int j = 0; j < MAXSUB; ++j) {
    iscldy[j] = 0;
    afracsub[j] = 0;
  }
  if (*jcldy > 0) iscldy[*jcldy] = 1;
  if (*jclea > 0) afracsub[*jclea] = *fclea;
  if (*jcldy > 0) afracsub[*jcldy] = *fcldy;
}

/* mam4_amicphys.hpp:322-361 */
static void set_subarea_rh(int ncldy_subarea, int jclea, int jcldy,
                           const double afracsub[MAXSUB], double relhumgcm,
                           double relhumsub[MAXSUB]) {
  if (ncldy_subarea <= 0) {
    for (int j = 0; j < MAXSUB; ++j) relhumsub[j] = relhumgcm;
  } else {
    relhumsub[jcldy] = 1;
    if (jclea > 0) {
      const double t = (relhumgcm - afracsub[jcldy]) / afracsub[jclea];
      relhumsub[jclea] = min_max_bound(0, 1, t);
    }
  }
}

/* mam4_amicphys.hpp:366-408 (qsub_a and qsub_b alias in the only call) */
static void compute_qsub_from_gcm_and_other(const int lcompute[31], double f_a,
                                            double f_b, const double qgcm[31],
                                            int jclea, int jcldy,
                                            double qsub[31][MAXSUB]) {
  for (int i = 0; i < 31; ++i) {
    if (lcompute[i]) {
      qsub[i][jcldy] = (qgcm[i] - f_a * qsub[i][jclea]) / f_b;
      if (qsub[i][jcldy] < 0) {
        qsub[i][jcldy] = 0;
        qsub[i][jclea] = qgcm[i] / f_a;
      }
    }
  }
}

/* mam4_amicphys.hpp:414-445 */
static void set_subarea_qnumb_for_cldbrn(int jclea, int jcldy, double fcldy,
                                         const double qqcwgcm[31],
                                         double qqcwsub[31][MAXSUB]) {
  for (int m = 0; m < 4; ++m) {
    const int i = NUMPTR[m];
    qqcwsub[i][jclea] = 0;
    qqcwsub[i][jcldy] = qqcwgcm[i] / fcldy;
  }
}

/* mam4_amicphys.hpp:451-487 */
static void set_subarea_qmass_for_cldbrn(int jclea, int jcldy, double fcldy,
                                         const double qqcwgcm[31],
                                         double qqcwsub[31][MAXSUB]) {
  for (int m = 0; m < 4; ++m)
    for (int s = 0; s < NSPEC_MODE[m]; ++s) {
      const int i = LMASSPTR[s][m];
      qqcwsub[i][jclea] = 0;
      qqcwsub[i][jcldy] = qqcwgcm[i] / fcldy;
    }
}

/* mam4_amicphys.hpp:492-546 */
static void get_partition_factors(double qgcm_intrst, double qgcm_cldbrn,
                                  double fcldy, double fclea,
                                  double *factor_clea, double *factor_cldy) {
  const double tmp_q_cldbrn_cldy = qgcm_cldbrn / fcldy;
  const double tmp_q_intrst_cldy =
      kmax(0, ((qgcm_intrst + qgcm_cldbrn) - tmp_q_cldbrn_cldy));
  const double tmp_q_intrst_clea =
      (qgcm_intrst - fcldy * tmp_q_intrst_cldy) / fclea;
  const double eps = 1.e-35;
  double clea2gcm = kmax(eps, tmp_q_intrst_clea * fclea) / kmax(eps, qgcm_intrst);
  clea2gcm = kmax(0, kmin(1, clea2gcm));
  *factor_clea = clea2gcm / fclea;
  *factor_cldy = (1 - clea2gcm) / fcldy;
}

/* mam4_amicphys.hpp:552-604 */
static void set_subarea_qnumb_for_intrst(int jclea, int jcldy, double fclea,
                                         double fcldy, const double qgcm[31],
                                         const double qqcwgcm[31],
                                         const double qgcmx[31],
                                         double qsubx[31][MAXSUB]) {
  for (int m = 0; m < 4; ++m) {
    double fa, fc;
    get_partition_factors(qgcm[NUMPTR[m]], qqcwgcm[NUMPTR[m]], fcldy, fclea,
                          &fa, &fc);
    const int i = NUMPTR[m];
    qsubx[i][jclea] = qgcmx[i] * fa;
    qsubx[i][jcldy] = qgcmx[i] * fc;
  }
}

/* mam4_amicphys.hpp:610-673 */
static void set_subarea_qmass_for_intrst(int jclea, int jcldy, double fclea,
                                         double fcldy, const double qgcm[31],
                                         const double qqcwgcm[31],
                                         const double qgcmx[31],
                                         double qsubx[31][MAXSUB]) {
  for (int m = 0; m < 4; ++m) {
    double qi = 0, qc = 0;
    for (int s = 0; s < NSPEC_MODE[m]; ++s) {
      qi = qi + qgcm[LMASSPTR[s][m]];
      qc = qc + qqcwgcm[LMASSPTR[s][m]];
    }
    double fa, fc;
    get_partition_factors(qi, qc, fcldy, fclea, &fa, &fc);
    for (int s = 0; s < NSPEC_MODE[m]; ++s) {
      const int i = LMASSPTR[s][m];
      qsubx[i][jclea] = qgcmx[i] * fa;
      qsubx[i][jcldy] = qgcmx[i] * fc;
    }
  }
}

/* mam4_amicphys.hpp:678-911 */
static void set_subarea_gases_and_aerosols(
    int nsubarea, int jclea, int jcldy, double fclea, double fcldy,
    const double qgcm1[31], const double qgcm2[31], const double qqcwgcm2[31],
    const double qgcm3[31], const double qqcwgcm3[31], double qsub1[31][MAXSUB],
    double qsub2[31][MAXSUB], double qqcwsub2[31][MAXSUB],
    double qsub3[31][MAXSUB], double qqcwsub3[31][MAXSUB]) {
  for (int i = 0; i < 31; ++i)
    for (int j = 0; j < MAXSUB; ++j) {
      qsub1[i][j] = 0;
      qsub2[i][j] = 0;
      qsub3[i][j] = 0;
      qqcwsub2[i][j] = 0;
      qqcwsub3[i][j] = 0;
    }
  const int only_clea = (jclea == 1) && (jcldy == 0) && (nsubarea == 1);
  const int only_cldy = (jclea == 0) && (jcldy == 1) && (nsubarea == 1);
  const int partly = (jclea > 0) && (jcldy > 0) && (jclea + jcldy == 3) &&
                     (nsubarea == 2);
  if (only_clea || only_cldy) {
    const int j = 1;
    for (int i = 0; i < 31; ++i)
      if (LMAPCC_ALL[i] > 0) {
        qsub1[i][j] = qgcm1[i];
        qsub2[i][j] = qgcm2[i];
        qsub3[i][j] = qgcm3[i];
        qqcwsub2[i][j] = qqcwgcm2[i];
        qqcwsub3[i][j] = qqcwgcm3[i];
      }
  } else if (partly) {
    int is_gas[31];
    for (int i = 0; i < 31; ++i) is_gas[i] = (LMAPCC_ALL[i] == LMAPCC_GAS);
    for (int i = 0; i < 31; ++i)
      if (is_gas[i])
        for (int j = 1; j <= nsubarea; ++j) qsub1[i][j] = qgcm1[i];
    for (int i = 0; i < 31; ++i)
      if (is_gas[i])
        for (int j = 1; j <= nsubarea; ++j) qsub2[i][j] = qgcm2[i];
    for (int i = 0; i < 31; ++i)
      if (is_gas[i]) qsub3[i][jclea] = qsub2[i][jclea];
    compute_qsub_from_gcm_and_other(is_gas, fclea, fcldy, qgcm3, jclea, jcldy,
                                    qsub3);
    set_subarea_qnumb_for_cldbrn(jclea, jcldy, fcldy, qqcwgcm2, qqcwsub2);
    set_subarea_qmass_for_cldbrn(jclea, jcldy, fcldy, qqcwgcm2, qqcwsub2);
    set_subarea_qnumb_for_cldbrn(jclea, jcldy, fcldy, qqcwgcm3, qqcwsub3);
    set_subarea_qmass_for_cldbrn(jclea, jcldy, fcldy, qqcwgcm3, qqcwsub3);
    set_subarea_qnumb_for_intrst(jclea, jcldy, fclea, fcldy, qgcm2, qqcwgcm2,
                                 qgcm2, qsub2);
    set_subarea_qmass_for_intrst(jclea, jcldy, fclea, fcldy, qgcm2, qqcwgcm2,
                                 qgcm2, qsub2);
    set_subarea_qnumb_for_intrst(jclea, jcldy, fclea, fcldy, qgcm2, qqcwgcm2,
                                 qgcm3, qsub3);
    set_subarea_qmass_for_intrst(jclea, jcldy, fclea, fcldy, qgcm2, qqcwgcm2,
                                 qgcm3, qsub3);
  }
}

/* ======================= gas-aerosol exchange ============================ */

/* gasaerexch.hpp:222-236 */
/* Test switch: 0 (default) keeps the reference's constants; 1 replaces the
 * air terms of Fuller's diffusivity -- 0.035 for 1/M_air, 2.7189 for
 * V_air^(1/3), and 1.013e5 Pa per atmosphere -- by 1/28.966, 20.1^(1/3) and
 * 1.01325e5, the values under which the Fortran MAM4 known answers of
 * src/tests/mam4_amicphys_1gridcell.cpp:1788-2386 were produced. */
static int g_fortran_air = 0;
void orc_set_fortran_air_constants(int on) { g_fortran_air = on; }

static double gas_diffusivity(double T, double p_atm, double mw, double vd) {
  const double onethird = 1.0 / 3.0;
  const double inv_mw_air = g_fortran_air ? 1.0 / 28.966 : 0.035;
  const double vd_air_cbrt = g_fortran_air ? pow(20.1, onethird) : 2.7189;
  const double dgas = (1.0e-3 * M_pow(T, 1.75) * M_sqrt(1. / mw + inv_mw_air)) /
                      (p_atm * M_pow((M_pow(vd, onethird) + vd_air_cbrt), 2.0));
  return dgas * 1.0e-4;
}
/* gasaerexch.hpp:240-250 */
static double mean_molecular_speed(double temp, double rmw) {
  return 145.5 * M_sqrt(temp / rmw);
}
/* gasaerexch.hpp:253-265 */
static double fuchs_sutugin(double D_p, double gasfreepath, double a283,
                            double a75) {
  const double kn = 2.0 * gasfreepath / D_p;
  return (a75 * (1.0 + kn)) / (kn * (kn + 1.0 + a283) + a75);
}
/* gasaerexch.hpp:267-350 */
static void gas_aer_uptkrates_1box1gas(double accom, double gasdiffus,
                                       double gasfreepath, double beta_inp,
                                       const double dgncur_awet[4],
                                       const double lnsg[4], double uptkaer[4]) {
  const double tworootpi = 2 * M_sqrt(PI);
  const double root2 = M_sqrt(2.0);
  const double xghq[2] = {7.071067690849304E-01, -7.071067690849304E-01};
  const double wghq[2] = {8.862269520759583E-01, 8.862269520759583E-01};
  const double a283 = accom * 0.283;
  const double a75 = accom * 0.75;
  for (int n = 0; n < 4; ++n) {
    const double lndpgn = M_log(dgncur_awet[n]);
    double beta = 0;
    if (fabs(beta_inp - 1.5) > 0.5) {
      const double D_p = dgncur_awet[n];
      const double kn = 2.0 * gasfreepath / D_p;
      const double tmpa = 1.0 / (1.0 + kn) -
                          (2.0 * kn + 1.0 + a283) / (kn * (kn + 1.0 + a283) + a75);
      beta = 1.0 - kn * tmpa;
      beta = kmax(1.0, kmin(2.0, beta));
    } else {
      beta = beta_inp;
    }
    const double constant =
        tworootpi * M_exp(beta * lndpgn + 0.5 * M_pow(beta * lnsg[n], 2.0));
    double sumghq = 0.0;
    for (int iq = 0; iq < 2; ++iq) {
      const double lndp = lndpgn + beta * lnsg[n] * lnsg[n] + root2 * lnsg[n] * xghq[iq];
      const double D_p = M_exp(lndp);
      const double hh = fuchs_sutugin(D_p, gasfreepath, a283, a75);
      sumghq += wghq[iq] * D_p * hh / M_pow(D_p, beta);
    }
    uptkaer[n] = constant * gasdiffus * sumghq;
  }
}

/* gasaerexch_soaexch.hpp:79-341 (9-argument overload). Returns the number of
 * sub-steps taken. */
static int mam_soaexch_1subarea(double dtsubstep, double temp, double pmid,
                                double qgas_cur[2], double qgas_avg[2],
                                double qaer_cur[7][4], const double uptkaer[2][4]) {
  const int niter_max = 1000;
  int niter = 0;
  const double dtfull = dtsubstep;
  double tcur = 0.0, dtcur = 0.0, dtsum_qgas_avg = 0.0;
  double qgas_prv[2] = {0.0, 0.0};
  double qaer_prv[7][4];
  double uptk_tmp[4] = {0.0, 0.0, 0.0, 0.0};
  double a_ooa_sum_tmp[4] = {0.0, 0.0, 0.0, 0.0};
  double a_opoa[4] = {0.0, 0.0, 0.0, 0.0};
  double a_soa[4] = {0.0, 0.0, 0.0, 0.0};
  double a_soa_tmp[4] = {0.0, 0.0, 0.0, 0.0};
  double beta[4] = {0.0, 0.0, 0.0, 0.0};
  double del_g_soa_tmp = 0.0;
  double g_soa = 0.0;
  double g_star[4] = {0.0, 0.0, 0.0, 0.0};
  double phi[4] = {0.0, 0.0, 0.0, 0.0};
  double sat[4] = {0.0, 0.0, 0.0, 0.0};
  const double delh_vap_soa = 156.0e3, p0_soa_298 = 1.0e-10;
  /* opoa_frac: 0.1 except 0 for the primary-carbon mode (:123-138) */
  const double opoa_frac[4] = {0.1, 0.1, 0.1, 0.0};
  int skip[4] = {1, 1, 1, 1};
  const double alpha_astem = 0.05, rgas = 8.3144, a_min1 = 1.0e-20,
               g_min1 = 1.0e-20;
  double tot_soa = 0.0;
  double tmpa, tmpb, tmpc;
  memset(qaer_prv, 0, sizeof(qaer_prv));

  const double p0_soa =
      p0_soa_298 * M_exp(-(delh_vap_soa / rgas) * ((1.0 / temp) - (1.0 / 298.0)));
  const double g0_soa = 1.01325e5 * p0_soa / pmid;
  qgas_avg[0] = 0.0;
  dtsum_qgas_avg = 0.0;
  while (tcur < dtfull - 1.0e-3) {
    niter++;
    if (niter > niter_max) break;
    qgas_prv[0] = qgas_cur[0];
    for (int a = 0; a < 7; ++a)
      for (int n = 0; n < 4; ++n) qaer_prv[a][n] = qaer_cur[a][n];
    for (int n = 0; n < 4; ++n) {
      skip[n] = 1;
      if (uptkaer[0][n] > 1.0e-15) {
        uptk_tmp[n] = uptkaer[0][n];
        skip[n] = 0;
      } else {
        uptk_tmp[n] = 0.0;
      }
    }
    g_soa = kmax(qgas_prv[0], 0.0);
    tot_soa = g_soa;
    for (int n = 0; n < 4; ++n) {
      if (skip[n]) continue;
      a_soa[n] = kmax(qaer_prv[0][n], 0.0);
      tot_soa += a_soa[n];
    }
    for (int n = 0; n < 4; ++n) {
      if (skip[n]) continue;
      a_opoa[n] = 0.0;
      a_opoa[n] += opoa_frac[n] * kmax(qaer_prv[2][n], 0.0); /* iaer_pom = 2 */
    }
    tmpa = 0.0;
    for (int n = 0; n < 4; ++n) {
      if (skip[n]) continue;
      a_ooa_sum_tmp[n] = a_opoa[n];
      a_ooa_sum_tmp[n] += a_soa[n];
    }
    tmpb = 0.0;
    for (int n = 0; n < 4; ++n) {
      if (skip[n]) continue;
      sat[n] = g0_soa / kmax(a_ooa_sum_tmp[n], a_min1);
      g_star[n] = sat[n] * a_soa[n];
      phi[n] = (g_soa - g_star[n]) / kmax(g_soa, kmax(g_star[n], g_min1));
      tmpb += uptk_tmp[n] * fabs(phi[n]);
    }
    tmpa = kmax(tmpa, tmpb);
    {
      const double dtmax = dtfull - tcur;
      if (dtmax * tmpa <= alpha_astem) {
        dtcur = dtmax;
        tcur = dtfull;
      } else {
        dtcur = alpha_astem / tmpa;
        tcur += dtcur;
      }
    }
    /* Step 1 (:277-304) */
    for (int n = 0; n < 4; ++n) {
      if (skip[n]) continue;
      a_soa_tmp[n] = a_soa[n];
      beta[n] = dtcur * uptk_tmp[n];
      del_g_soa_tmp = g_soa - g_star[n];
      if (del_g_soa_tmp > 0.0) a_soa_tmp[n] += beta[n] * del_g_soa_tmp;
      a_ooa_sum_tmp[n] = a_opoa[n];
      a_ooa_sum_tmp[n] += a_soa_tmp[n];
      if (del_g_soa_tmp > 0.0) {
        sat[n] = g0_soa / kmax(a_ooa_sum_tmp[n], a_min1);
        g_star[n] = sat[n] * a_soa_tmp[n];
      }
    }
    /* Step 2 (:306-320) */
    tmpa = 0.0;
    tmpb = 0.0;
    for (int n = 0; n < 4; ++n) {
      if (skip[n]) continue;
      tmpa += a_soa[n] / (1.0 + beta[n] * sat[n]);
      tmpb += beta[n] / (1.0 + beta[n] * sat[n]);
    }
    g_soa = (tot_soa - tmpa) / (1.0 + tmpb);
    g_soa = kmax(0.0, g_soa);
    for (int n = 0; n < 4; ++n) {
      if (skip[n]) continue;
      a_soa[n] = (a_soa[n] + beta[n] * g_soa) / (1.0 + beta[n] * sat[n]);
    }
    /* update (:322-333).  QUIRK: every mode's SOA is overwritten, skipped
     * modes with the a_soa they never set (0). */
    qgas_cur[0] = g_soa;
    tmpc = qgas_cur[0] - qgas_prv[0];
    qgas_avg[0] += dtcur * (qgas_prv[0] + 0.5 * tmpc);
    for (int n = 0; n < 4; ++n) qaer_cur[0][n] = a_soa[n];
    dtsum_qgas_avg += dtcur;
  }
  qgas_avg[0] = kmax(0.0, qgas_avg[0] / dtsum_qgas_avg);
  (void)beta;
  return niter;
}

/* SOA exchange of one subarea (gasaerexch_soaexch.hpp:79-341) on its own,
 * for the closed-form and brute-force checks: q_soag [mol/mol, fcvt applied],
 * soa[4] and pom[4] per mode, uptk_soag[4] the SOAG uptake rates [1/s].
 * Returns the sub-step count; *q_soag_avg receives the time average. */
int orc_soaexch(double dt, double temp, double pmid, double *q_soag, double *q_soag_avg,
                double soa[4], const double pom[4], const double uptk_soag[4]) {
  double qgas[2] = {*q_soag, 0.0}, qavg[2] = {0.0, 0.0};
  double qaer[7][4], up[2][4];
  memset(qaer, 0, sizeof(qaer));
  for (int n = 0; n < 4; ++n) {
    qaer[0][n] = soa[n];
    qaer[2][n] = pom[n];
    up[0][n] = uptk_soag[n];
    up[1][n] = 0.0;
  }
  const int it = mam_soaexch_1subarea(dt, temp, pmid, qgas, qavg, qaer, up);
  *q_soag = qgas[0];
  *q_soag_avg = qavg[0];
  for (int n = 0; n < 4; ++n) soa[n] = qaer[0][n];
  return it;
}

/* H2SO4 block of mam_gasaerexch_1subarea (gasaerexch.hpp:479-537). */
void orc_nonvolatile_update(double dt, double netprod, const double uptkaer[4],
                            double *qgas_cur, double *qgas_avg,
                            double qaer_cur[4]) {
  double qaer_prv[4];
  for (int n = 0; n < 4; ++n) qaer_prv[n] = qaer_cur[n];
  const double qgas_prv = *qgas_cur;
  double tmpa = 0.0;
  for (int n = 0; n < 4; ++n) tmpa += uptkaer[n];
  const double tmp_kxt = tmpa * dt;
  const double tmp_pxt = netprod * dt;
  const double tmp_q1 = qgas_prv;
  if (tmp_kxt >= 1.0e-20) {
    double tmp_q3, tmp_q4;
    if (tmp_kxt > 0.001) {
      const double tmp_pok = tmp_pxt / tmp_kxt;
      tmp_q3 = (tmp_q1 - tmp_pok) * M_exp(-tmp_kxt) + tmp_pok;
      tmp_q4 = (tmp_q1 - tmp_pok) * (1.0 - M_exp(-tmp_kxt)) / tmp_kxt + tmp_pok;
    } else {
      const double k2 = tmp_kxt * tmp_kxt;
      tmp_q3 = tmp_q1 * (1.0 - tmp_kxt + k2 * 0.5) +
               tmp_pxt * (1.0 - tmp_kxt * 0.5 + k2 / 6.0);
      tmp_q4 = tmp_q1 * (1.0 - tmp_kxt * 0.5 + k2 / 6.0) +
               tmp_pxt * (0.5 - tmp_kxt / 6.0 + k2 / 24.0);
    }
    *qgas_cur = tmp_q3;
    const double tmp_qdel_cond = (tmp_q1 + tmp_pxt) - tmp_q3;
    *qgas_avg = tmp_q4;
    for (int n = 0; n < 4; ++n) {
      if (uptkaer[n] <= 0.0) continue;
      const double tmpc = tmp_qdel_cond * (uptkaer[n] / tmpa);
      qaer_cur[n] = qaer_prv[n] + tmpc;
    }
  } else {
    *qgas_cur = tmp_q1 + tmp_pxt;
    *qgas_avg = tmp_q1 + tmp_pxt * 0.5;
  }
}

/* gasaerexch.hpp:354-554 (jtsubstep == 0 only).  Returns SOA iterations. */
static int mam_gasaerexch_1subarea(int jtsubstep, double dtsubstep, double temp,
                                   double pmid, double aircon,
                                   double qgas_cur[2], double qgas_avg[2],
                                   const double qgas_netprod_otrproc[2],
                                   double qaer_cur[7][4], double qnum_cur[4],
                                   const double dgn_awet[4],
                                   double uptkaer[2][4],
                                   double *uptkrate_h2so4) {
  const int mode_aging_optaa[4] = {0, 0, 0, 1};
  double alnsg_aer[4];
  for (int m = 0; m < 4; ++m) alnsg_aer[m] = M_log(MODE_SIGMA[m]);
  const double mw_gas[2] = {1.500000000000000E+02, 9.807840000000000E+01};
  const double vol_molar_gas[2] = {6.563265306122449E+01, 4.288000000000000E+01};
  const double accom_coef_gas[2] = {6.5000E-01, 6.5000E-01};
  double uptkrate[4];
  qgas_avg[0] = 0.0;
  qgas_avg[1] = 0.0;
  if (jtsubstep == 0) {
    const double p_in_atm = pmid / (g_fortran_air ? 1.01325e5 : 1.013e5);
    for (int igas = 0; igas < 2; ++igas) {
      const double dif = gas_diffusivity(temp, p_in_atm, mw_gas[igas], vol_molar_gas[igas]);
      const double spd = mean_molecular_speed(temp, mw_gas[igas]);
      const double freepath = 3.0 * dif / spd;
      gas_aer_uptkrates_1box1gas(accom_coef_gas[igas], dif, freepath, 0.0,
                                 dgn_awet, alnsg_aer, uptkrate);
      const int iaer = igas;
      for (int n = 0; n < 4; ++n) {
        if (LMAP_AER[iaer][n] > 0 || mode_aging_optaa[n] > 0)
          uptkaer[igas][n] = uptkrate[n] * (qnum_cur[n] * aircon);
        else
          uptkaer[igas][n] = 0.0;
      }
    }
    for (int m = 0; m < 4; ++m) uptkaer[0][m] = uptkaer[IGAS_H2SO4][m] * 0.81;
    for (int n = 0; n < 4; ++n) *uptkrate_h2so4 += uptkaer[IGAS_H2SO4][n];
  }
  const int niter = mam_soaexch_1subarea(dtsubstep, temp, pmid, qgas_cur,
                                         qgas_avg, qaer_cur, uptkaer);
  {
    double qa[4], qg = qgas_cur[1], qavg = 0.0;
    for (int n = 0; n < 4; ++n) qa[n] = qaer_cur[A_SO4][n];
    orc_nonvolatile_update(dtsubstep, qgas_netprod_otrproc[1], uptkaer[1], &qg,
                           &qavg, qa);
    qgas_cur[1] = qg;
    qgas_avg[1] = qavg;
    for (int n = 0; n < 4; ++n) qaer_cur[A_SO4][n] = qa[n];
  }
  return niter;
}

/* ======================= nucleation ====================================== */

/* vehkamaki2002.hpp:53-66 */
double orc_vehk_xcrit(double c, double t, double rh) {
  const double lnN = M_log(c), lnrh = M_log(rh);
  return 0.740997 - 0.00266379 * t - 0.00349998 * lnN + 0.0000504022 * t * lnN +
         0.00201048 * lnrh - 0.000183289 * t * lnrh + 0.00157407 * sq(lnrh) -
         0.0000179059 * t * sq(lnrh) + 0.000184403 * cube(lnrh) -
         1.50345e-6 * t * cube(lnrh);
}
/* vehkamaki2002.hpp:74-115 */
double orc_vehk_logj(double c, double t, double rh, double x) {
  const double a = 0.14309 + 2.21956 * t - 0.0273911 * sq(t) + 0.0000722811 * cube(t) + 5.91822 / x;
  const double b = 0.117489 + 0.462532 * t - 0.0118059 * sq(t) + 0.0000404196 * cube(t) + 15.7963 / x;
  const double cc = -0.215554 - 0.0810269 * t + 0.00143581 * sq(t) - 4.7758e-6 * cube(t) - 2.91297 / x;
  const double d = -3.58856 + 0.049508 * t - 0.00021382 * sq(t) + 3.10801e-7 * cube(t) - 0.0293333 / x;
  const double e = 1.14598 - 0.600796 * t + 0.00864245 * sq(t) - 0.0000228947 * cube(t) - 8.44985 / x;
  const double f = 2.15855 + 0.0808121 * t - 0.000407382 * sq(t) - 4.01957e-7 * cube(t) + 0.721326 / x;
  const double g = 1.6241 - 0.0160106 * t + 0.0000377124 * sq(t) + 3.21794e-8 * cube(t) - 0.0113255 / x;
  const double h = 9.71682 - 0.115048 * t + 0.000157098 * sq(t) + 4.00914e-7 * cube(t) + 0.71186 / x;
  const double i = -1.05611 + 0.00903378 * t - 0.0000198417 * sq(t) + 2.46048e-8 * cube(t) - 0.0579087 / x;
  const double j = -0.148712 + 0.00283508 * t - 9.24619e-6 * sq(t) + 5.00427e-9 * cube(t) - 0.0127081 / x;
  const double lnN = M_log(c), lnrh = M_log(rh);
  return a + b * lnrh + cc * sq(lnrh) + d * cube(lnrh) + e * lnN +
         f * lnrh * lnN + g * sq(lnrh) * lnN + h * sq(lnN) +
         i * lnrh * sq(lnN) + j * cube(lnN);
}
/* vehkamaki2002.hpp:130-169 */
double orc_vehk_ntot(double c, double t, double rh, double x) {
  const double A = -0.00295413 - 0.0976834 * t + 0.00102485 * sq(t) - 2.18646e-6 * cube(t) - 0.101717 / x;
  const double B = -0.00205064 - 0.00758504 * t + 0.000192654 * sq(t) - 6.7043e-7 * cube(t) - 0.255774 / x;
  const double C = +0.00322308 + 0.000852637 * t - 0.0000154757 * sq(t) + 5.66661e-8 * cube(t) + 0.0338444 / x;
  const double D = +0.0474323 - 0.000625104 * t + 2.65066e-6 * sq(t) - 3.67471e-9 * cube(t) - 0.000267251 / x;
  const double E = -0.0125211 + 0.00580655 * t - 0.000101674 * sq(t) + 2.88195e-7 * cube(t) + 0.0942243 / x;
  const double F = -0.038546 - 0.000672316 * t + 2.60288e-6 * sq(t) + 1.19416e-8 * cube(t) - 0.00851515 / x;
  const double G = -0.0183749 + 0.000172072 * t - 3.71766e-7 * sq(t) - 5.14875e-10 * cube(t) + 0.00026866 / x;
  const double H = -0.0619974 + 0.000906958 * t - 9.11728e-7 * sq(t) - 5.36796e-9 * cube(t) - 0.00774234 / x;
  const double I = +0.0121827 - 0.00010665 * t + 2.5346e-7 * sq(t) - 3.63519e-10 * cube(t) + 0.000610065 / x;
  const double J = +0.000320184 - 0.0000174762 * t + 6.06504e-8 * sq(t) - 1.4177e-11 * cube(t) + 0.000135751 / x;
  const double lnN = M_log(c), lnrh = M_log(rh);
  return M_exp(A + B * lnrh + C * sq(lnrh) + D * cube(lnrh) + E * lnN +
             F * lnrh * lnN + G * sq(lnrh) * lnN + H * sq(lnN) +
             I * lnrh * sq(lnN) + J * cube(lnN));
}
/* vehkamaki2002.hpp:175-179 */
double orc_vehk_radius(double x, double ntot) {
  return M_exp(-1.6524245 + 0.42316402 * x + 0.3346648 * M_log(ntot));
}
/* vehkamaki2002.hpp:186-193: eq 15, H2SO4 [cm-3] giving J = 1 cm-3 s-1 */
double orc_vehk_threshold(double t, double rh) {
  return M_exp(-279.243 + 11.7344 * rh + 22700.9 / t - 1088.64 * rh / t +
             1.14436 * t - 0.0302331 * rh * t - 0.00130254 * sq(t) -
             6.38697 * M_log(rh) + 854.98 * M_log(rh) / t +
             0.00879662 * t * M_log(rh));
}

/* nucleation.hpp:127-165 */
static void binary_nuc_vehk2002(double temp, double rh, double so4vol,
                                double *ratenucl, double *rateloge,
                                double *cnum_h2so4, double *cnum_tot,
                                double *radius_cluster) {
  const double x = orc_vehk_xcrit(so4vol, temp, rh);
  *rateloge = orc_vehk_logj(so4vol, temp, rh, x);
  *ratenucl = M_exp(kmin(*rateloge, M_log(1e38)));
  *cnum_tot = orc_vehk_ntot(so4vol, temp, rh, x);
  *cnum_h2so4 = *cnum_tot * x;
  *radius_cluster = orc_vehk_radius(x, *cnum_tot);
}

/* nucleation.hpp:37-116 */
static void pbl_nuc_wang2008(double so4vol, double pi, int choice,
                             double adjust, int *actual, double *ratenucl,
                             double *rateloge, double *cnum_tot,
                             double *cnum_h2so4, double *cnum_nh3,
                             double *radius_cluster_nm) {
  const double mw_h2so4_gmol = 98.0, avogadro_mol = 6.023e23,
               density_sulfate_gcm3 = 1.8;
  *actual = 0;
  double tmp_ratenucl;
  if (choice == 11)
    tmp_ratenucl = 1e-6 * so4vol; /* wang2008.hpp:23 */
  else if (choice == 12)
    tmp_ratenucl = 1e-12 * so4vol * so4vol;
  else
    return;
  tmp_ratenucl = tmp_ratenucl * adjust;
  const double tmp_rateloge = M_log(kmax(1.0e-38, tmp_ratenucl));
  if (tmp_rateloge <= *rateloge) return;
  *actual = choice;
  *rateloge = tmp_rateloge;
  *ratenucl = tmp_ratenucl;
  *radius_cluster_nm = 0.5;
  const double tmp_diam = *radius_cluster_nm * 2.0e-7;
  const double tmp_volu = cube(tmp_diam) * (pi / 6.0);
  const double tmp_mass = tmp_volu * density_sulfate_gcm3;
  *cnum_h2so4 = (tmp_mass / mw_h2so4_gmol) * avogadro_mol;
  *cnum_nh3 = 0.0;
  *cnum_tot = *cnum_h2so4;
}

/* nucleation.hpp:346-809.  out: [isize_nuc, qnuma_del, qso4a_del, qnh4a_del,
 * qh2so4_del, qnh3_del, dens_nh4so4a, dnclusterdt] */
void orc_nuc_mosaic_1box(int flag, double dtnuc, double temp_in, double rh_in,
                         double press_in, double zm_in, double pblh_in,
                         double qh2so4_cur, double qh2so4_avg, double qnh3_cur,
                         double h2so4_uptkrate, double mw_so4a_host, int nsize,
                         double dp_lo_sect, double dp_hi_sect, double rgas,
                         double avogad, double mw_nh4a, double mw_so4a,
                         double pi, double out[8]) {
  const double adjust_bin = 1.0, adjust_pbl = 1.0, onethird = 1.0 / 3.0;
  const double accom_coef_h2so4 = 0.65;
  const double dens_ammsulf = 1770.0, dens_ammbisulf = 1770.0,
               dens_sulfacid = 1770.0;
  const double mw_ammsulf = 132.0, mw_ammbisulf = 114.0, mw_sulfacid = 96.0;
  int isize_nuc = 1;
  double qnuma_del = 0.0, qso4a_del = 0.0, qnh4a_del = 0.0, qh2so4_del = 0.0,
         qnh3_del = 0.0, dens_nh4so4a = 0.0, dnclusterdt = 0.0;
#define ORC_NUC_OUT()                                                          \
  do {                                                                         \
    out[0] = isize_nuc; out[1] = qnuma_del; out[2] = qso4a_del;                \
    out[3] = qnh4a_del; out[4] = qh2so4_del; out[5] = qnh3_del;                \
    out[6] = dens_nh4so4a; out[7] = dnclusterdt;                               \
  } while (0)
  if ((flag != 1) && (flag != 2) && (flag != 11) && (flag != 12)) {
    ORC_NUC_OUT();
    return;
  }
  const double cair = press_in / (temp_in * rgas);
  const double so4vol_in = qh2so4_avg * cair * avogad * 1.0e-6;
  const double nh3ppt = qnh3_cur * 1.0e12;
  double ratenuclt = 1.0e-38;
  double rateloge = M_log(ratenuclt);
  double cnum_h2so4 = 0.0, cnum_tot = 0.0, radius_cluster = 0.0, cnum_nh3 = 0.0;
  int flag2 = -1;
  if ((flag != 2) && (nh3ppt >= 0.1)) {
    flag2 = 1; /* ternary branch: unreachable with NH3 = 0 */
  } else {
    if (so4vol_in >= 1.0e4) {
      const double temp_bb = kmax(230.15, kmin(305.15, temp_in));
      const double rh_bb = kmax(1.0e-4, kmin(1.0, rh_in));
      const double so4vol_bb = kmax(1.0e4, kmin(1.0e11, so4vol_in));
      binary_nuc_vehk2002(temp_bb, rh_bb, so4vol_bb, &ratenuclt, &rateloge,
                          &cnum_h2so4, &cnum_tot, &radius_cluster);
    }
    cnum_nh3 = 0.0;
    flag2 = 2;
  }
  rateloge = rateloge + M_log(kmax(1.0e-38, adjust_bin));
  if ((flag == 11) || (flag == 12)) {
    if (zm_in <= kmax(pblh_in, 100.0)) {
      pbl_nuc_wang2008(so4vol_in, pi, flag, adjust_pbl, &flag2, &ratenuclt,
                       &rateloge, &cnum_tot, &cnum_h2so4, &cnum_nh3,
                       &radius_cluster);
    }
  }
  (void)flag2;
  if (rateloge <= -13.82) {
    ORC_NUC_OUT();
    return;
  }
  ratenuclt = M_exp(rateloge);
  const double ratenuclt_bb = ratenuclt * 1.0e6;
  dnclusterdt = ratenuclt_bb;
  double tmpa = kmax(0.10, kmin(0.95, rh_in));
  const double wetvol_dryvol = 1.0 - 0.56 / M_log(tmpa);
  double voldry_clus = (kmax(cnum_h2so4, 1.0) * mw_so4a + cnum_nh3 * mw_nh4a) /
                       (1.0e3 * dens_sulfacid * avogad);
  voldry_clus = voldry_clus * (mw_so4a_host / mw_so4a);
  const double dpdry_clus = M_pow(voldry_clus * 6.0 / pi, onethird);
  int igrow;
  isize_nuc = 1;
  double dpdry_part = dp_lo_sect;
  if (dpdry_clus <= dp_lo_sect) {
    igrow = 1;
  } else if (dpdry_clus >= dp_hi_sect) {
    igrow = 0;
    isize_nuc = nsize;
    dpdry_part = dp_hi_sect;
  } else {
    igrow = 0;
    for (int i = 0; i < nsize; ++i) {
      if (dpdry_clus < dp_hi_sect) {
        /* QUIRK (nucleation.hpp:623-630): early return, no particles. */
        isize_nuc = i;
        ORC_NUC_OUT();
        return;
      }
    }
  }
  const double voldry_part = (pi / 6.0) * M_pow(dpdry_part, 3);
  double n1 = 0.0, n2 = 0.0, n3 = 0.0;
  if (igrow <= 0) {
    n1 = 0.0; n2 = 0.0; n3 = 1.0;
  } else if (qnh3_cur >= qh2so4_cur) {
    n1 = (qnh3_cur / qh2so4_cur) - 1.0;
    n1 = kmax(0.0, kmin(1.0, n1));
    n2 = 1.0 - n1;
    n3 = 0.0;
  } else {
    n1 = 0.0;
    n2 = (qnh3_cur / qh2so4_cur);
    n2 = kmax(0.0, kmin(1.0, n2));
    n3 = 1.0 - n2;
  }
  const double m1 = n1 * mw_ammsulf, m2 = n2 * mw_ammbisulf, m3 = n3 * mw_sulfacid;
  const double dens_part = (m1 + m2 + m3) /
                           ((m1 / dens_ammsulf) + (m2 / dens_ammbisulf) + (m3 / dens_sulfacid));
  dens_nh4so4a = dens_part;
  const double mass_part = voldry_part * dens_part;
  const double molenh4a_per_moleso4a = 2.0 * n1 + n2;
  double kgaero_per_moleso4a = 1.0e-3 * (m1 + m2 + m3);
  kgaero_per_moleso4a = kgaero_per_moleso4a * (mw_so4a_host / mw_so4a);
  double tmpb = 1.0 + molenh4a_per_moleso4a * 17.0 / 98.0;
  const double wet_volfrac_so4a = 1.0 / (wetvol_dryvol * tmpb);
  double factor_kk = 0.0;
  if (igrow <= 0) {
    factor_kk = 1.0;
  } else {
    const double tmp_spd = 14.7 * M_sqrt(temp_in);
    const double gr_kk = 3.0e-9 * tmp_spd * mw_sulfacid * so4vol_in /
                         (dens_part * wet_volfrac_so4a);
    const double dfin_kk = 1.0e9 * dpdry_part * M_pow(wetvol_dryvol, onethird);
    double dnuc_kk = 2.0 * radius_cluster;
    dnuc_kk = kmax(dnuc_kk, 1.0);
    const double gamma_kk = 0.23 * M_pow(dnuc_kk, 0.2) * M_pow(dfin_kk / 3.0, 0.075) *
                            M_pow(dens_part * 1.0e-3, -0.33) *
                            M_pow(temp_in / 293.0, -0.75);
    tmpa = h2so4_uptkrate * 3600.0;
    tmpa = kmax(tmpa, 0.0);
    tmpb = 6.7037e-6 * M_pow(temp_in, 0.75) / cair;
    tmpb = tmpb * 3600.0;
    const double cs_prime_kk = tmpa / (4.0 * pi * tmpb * accom_coef_h2so4);
    const double nu_kk = gamma_kk * cs_prime_kk / gr_kk;
    factor_kk = M_exp((nu_kk / dfin_kk) - (nu_kk / dnuc_kk));
  }
  const double ratenuclt_kk = ratenuclt_bb * factor_kk;
  tmpa = kmax(0.0, (ratenuclt_kk * dtnuc * mass_part));
  const double tmpe = tmpa / (kgaero_per_moleso4a * cair);
  const double qmolso4a_del_max = tmpe;
  double freducea = 1.0;
  if (qmolso4a_del_max > qh2so4_cur) freducea = qh2so4_cur / qmolso4a_del_max;
  double freduceb = 1.0;
  if (molenh4a_per_moleso4a >= 1.0e-10) {
    const double qmolnh4a_del_max = qmolso4a_del_max * molenh4a_per_moleso4a;
    if (qmolnh4a_del_max > qnh3_cur) freduceb = qnh3_cur / qmolnh4a_del_max;
  }
  const double freduce = kmin(freducea, freduceb);
  if ((freduce * ratenuclt_kk) <= 1.0e-12) {
    ORC_NUC_OUT();
    return;
  }
  tmpa = 0.9999;
  qh2so4_del = kmin(tmpa * qh2so4_cur, freduce * qmolso4a_del_max);
  qnh3_del = kmin(tmpa * qnh3_cur, qh2so4_del * molenh4a_per_moleso4a);
  qh2so4_del = -qh2so4_del;
  qnh3_del = -qnh3_del;
  qso4a_del = -qh2so4_del;
  qnh4a_del = -qnh3_del;
  qnuma_del = 1.0e-3 * (qso4a_del * mw_so4a + qnh4a_del * mw_nh4a) / mass_part;
  ORC_NUC_OUT();
#undef ORC_NUC_OUT
}

/* mam4_amicphys.hpp:916-1095 */
static void mam_newnuc_1subarea(double deltat, double temp, double pmid,
                                double zmid, double pblh, double relhum,
                                double uptkrate_h2so4, double qgas_cur[2],
                                const double qgas_avg[2], double qnum_cur[4],
                                double qaer_cur[7][4]) {
  const int flag = 11;
  const double qh2so4_cutoff = 4.0e-16;
  const double qh2so4_cur = qgas_cur[IGAS_H2SO4];
  const double qh2so4_avg = qgas_avg[IGAS_H2SO4];
  const double tmp_uptkrate = uptkrate_h2so4;
  if (qh2so4_avg <= qh2so4_cutoff) return;
  const double qnh3_cur = 0;
  const double dplom = M_exp(0.67 * M_log(MODE_DMIN[NAIT]) + 0.33 * M_log(MODE_DNOM[NAIT]));
  const double dphim = MODE_DMAX[NAIT];
  const double dens_so4a_host = 1770;
  double tmpa = dens_so4a_host * PI / 6.0;
  const double mass1p_aitlo = tmpa * (M_pow(dplom, 3.0));
  const double mass1p_aithi = tmpa * (M_pow(dphim, 3.0));
  const double relhumnn = kmax(0.01, kmin(0.99, relhum));
  const double mw_so4a_host = 115, mwnh4 = 18, mwso4 = 96;
  const double rgas = 8.31446759100000, avogadro = 6.022140000000000E+023;
  double o[8];
  orc_nuc_mosaic_1box(flag, deltat, temp, relhumnn, pmid, zmid, pblh,
                      qh2so4_cur, qh2so4_avg, qnh3_cur, tmp_uptkrate,
                      mw_so4a_host, 1, dplom, dphim, rgas, avogadro, mwnh4,
                      mwso4, PI, o);
  double qnuma_del = o[1] * 1.0e3;
  const double qso4a_del = o[2];
  double dndt_ait = qnuma_del / deltat;
  tmpa = qso4a_del * mw_so4a_host;
  const double tmpb = tmpa;
  const double tmp_frso4 = 1.0;
  double dmdt_ait = kmax(0.0, (tmpb / deltat));
  if (dndt_ait < 1.0e2) {
    dndt_ait = 0.0;
    dmdt_ait = 0.0;
  } else {
    const double mass1p = dmdt_ait / dndt_ait;
    if (mass1p < mass1p_aitlo)
      dndt_ait = dmdt_ait / mass1p_aitlo;
    else if (mass1p > mass1p_aithi)
      dmdt_ait = dndt_ait * mass1p_aithi;
  }
  const double adj = 1;
  dndt_ait = dndt_ait * adj;
  dmdt_ait = dmdt_ait * adj;
  double tmp_q_del = dndt_ait * deltat;
  qnum_cur[NAIT] = qnum_cur[NAIT] + tmp_q_del;
  const double dso4dt_ait = dmdt_ait * tmp_frso4 / mw_so4a_host;
  if (dso4dt_ait > 0.0) {
    tmp_q_del = dso4dt_ait * deltat;
    qaer_cur[A_SO4][NAIT] = qaer_cur[A_SO4][NAIT] + tmp_q_del;
    tmp_q_del = kmin(tmp_q_del, qgas_cur[IGAS_H2SO4]);
    qgas_cur[IGAS_H2SO4] = qgas_cur[IGAS_H2SO4] - tmp_q_del;
  }
}

/* ======================= coagulation ===================================== */

double orc_bm0ij(int n1, int n2a, int n2n) { return WHITBY_BM0IJ[n1][n2a][n2n]; }
double orc_bm3i(int n1, int n2a, int n2n) { return WHITBY_BM3I[n1][n2a][n2n]; }

/* coagulation.hpp:297-328 (table argument swap at :306) */
static double intermodal_0th(double a, double r1, double r2, double rx4,
                             double ri1, double ri2, double ri3, double knc,
                             double kngat, double kngac, double kfmatac,
                             double sqdgat, double esat01, double esat04,
                             double esat09, double esat16, double esac01,
                             double esac04, double esac09, double esac16,
                             int n1, int n2a, int n2n) {
  const double bm0ij = orc_bm0ij(n1, n2n, n2a);
  const double coagnc0 =
      knc * (2.0 + a * (kngat * (esat04 + r2 * esat16 * esac04) +
                        kngac * (esac04 + ri2 * esac16 * esat04)) +
             (r2 + ri2) * esat04 * esac04);
  const double coagfm0 = kfmatac * sqdgat * bm0ij *
                         (esat01 + r1 * esac01 + 2.0 * r2 * esat01 * esac04 +
                          rx4 * esat09 * esac16 + ri3 * esat16 * esac09 +
                          2.0 * ri1 * esat04 + esac01);
  return coagnc0 * coagfm0 / (coagnc0 + coagfm0);
}

/* coagulation.hpp:548-579 (table argument swap at :559) */
static double intermodal_3rd(double a, double r1, double r2, double rx4,
                             double ri1, double ri2, double ri3, double knc,
                             double kngat, double kngac, double dgat3,
                             double kfmatac, double sqdgat7, double esat04,
                             double esat09, double esat16, double esat25,
                             double esat36, double esat49, double esat64,
                             double esac01, double esac04, double esac09,
                             double esac16, double esat100, int n1, int n2a,
                             int n2n) {
  const double bm3i = orc_bm3i(n1, n2n, n2a);
  const double coagnc3 =
      knc * dgat3 *
      (2.0 * esat36 + a * kngat * (esat16 + r2 * esat04 * esac04) +
       a * kngac * (esat36 * esac04 + ri2 * esat64 * esac16) +
       r2 * esat16 * esac04 + ri2 * esat64 * esac04);
  const double coagfm3 = kfmatac * sqdgat7 * bm3i *
                         (esat49 + r1 * esat36 * esac01 +
                          2.0 * r2 * esat25 * esac04 + rx4 * esat09 * esac16 +
                          ri3 * esat100 * esac09 + 2.0 * ri1 * esat64 * esac01);
  return coagnc3 * coagfm3 / (coagnc3 + coagfm3);
}

/* coagulation.hpp:581-609 */
static double intramodal_0th(double a, double knc, double kngxx, double kfmxx,
                             double sqdgxx, double esxx01, double esxx04,
                             double esxx05, double esxx08, double esxx20,
                             double esxx25, int n2x) {
  static const double bm0[10] = {0.707106785165097, 0.726148960080488,
                                 0.766430744110958, 0.814106389441342,
                                 0.861679526483207, 0.903600509090092,
                                 0.936578814219156, 0.960098926735545,
                                 0.975646823342881, 0.985397173215326};
  const double coagnc = knc * (1.0 + esxx08 + a * kngxx * (esxx20 + esxx04));
  const double coagfm = kfmxx * sqdgxx * bm0[n2x] * (esxx01 + esxx25 + 2.0 * esxx05);
  return coagfm * coagnc / (coagfm + coagnc);
}

/* coagulation.hpp:650-748 */
static void getcoags(double lamda, double kfmatac, double kfmat, double kfmac,
                     double knc, double dgatk, double dgacc, double sgatk,
                     double sgacc, double xxlsgat, double xxlsgac, double *qn11,
                     double *qn22, double *qn12, double *qv12) {
  const double a = 1.246;
  const double esat01 = M_exp(0.125 * xxlsgat * xxlsgat);
  const double esac01 = M_exp(0.125 * xxlsgac * xxlsgac);
  const double sqrttwo = M_sqrt(2.0);
  const double dlgsqt2 = 1.0 / M_log(sqrttwo);
  const double esat04 = M_pow(esat01, 4.0);
  const double esac04 = M_pow(esac01, 4.0);
  const double esat05 = esat04 * esat01, esac05 = esac04 * esac01;
  const double esat08 = esat04 * esat04, esac08 = esac04 * esac04;
  const double esat09 = esat08 * esat01, esac09 = esac08 * esac01;
  const double esat16 = esat08 * esat08, esac16 = esac08 * esac08;
  const double esat20 = esat16 * esat04, esac20 = esac16 * esac04;
  const double esat24 = esat20 * esat04;
  const double esat25 = esat20 * esat05, esac25 = esac20 * esac05;
  const double esat36 = esat20 * esat16;
  const double esat49 = esat24 * esat25;
  const double esat64 = esat20 * esat20 * esat24;
  const double esat100 = esat64 * esat36;
  const double dgat3 = dgatk * dgatk * dgatk;
  const double sqdgat = M_sqrt(dgatk), sqdgac = M_sqrt(dgacc);
  const double sqdgat7 = dgat3 * sqdgat;
  const double r1 = sqdgac / sqdgat;
  const double r2 = r1 * r1;
  const double rx4 = r2 * r2;
  const double ri1 = 1.0 / r1;
  const double ri2 = 1.0 / (r1 * r1);
  const double ri3 = 1.0 / (r1 * r1 * r1);
  const double kngat = 2.0 * lamda / dgatk;
  const double kngac = 2.0 * lamda / dgacc;
  const double rat = dgacc / dgatk;
  const int n2n = (int)(kmax(1, kmin(10, round(4.0 * (sgatk - 0.75)))) - 1);
  const int n2a = (int)(kmax(1, kmin(10, round(4.0 * (sgacc - 0.75)))) - 1);
  const int n1 = (int)(kmax(1, kmin(10, 1 + round(dlgsqt2 * M_log(rat)))) - 1);
  *qn12 = intermodal_0th(a, r1, r2, rx4, ri1, ri2, ri3, knc, kngat, kngac,
                         kfmatac, sqdgat, esat01, esat04, esat09, esat16,
                         esac01, esac04, esac09, esac16, n1, n2a, n2n);
  *qv12 = intermodal_3rd(a, r1, r2, rx4, ri1, ri2, ri3, knc, kngat, kngac,
                         dgat3, kfmatac, sqdgat7, esat04, esat09, esat16,
                         esat25, esat36, esat49, esat64, esac01, esac04,
                         esac09, esac16, esat100, n1, n2a, n2n);
  *qn11 = intramodal_0th(a, knc, kngat, kfmat, sqdgat, esat01, esat04, esat05,
                         esat08, esat20, esat25, n2n);
  *qn22 = intramodal_0th(a, knc, kngac, kfmac, sqdgac, esac01, esac04, esac05,
                         esac08, esac20, esac25, n2a);
}

/* coagulation.hpp:750-817.  out = {betaij0, betaij3, betaii0, betajj0} */
void orc_getcoags_wrapper(double airtemp, double airprs, double dgatk,
                          double dgacc, double sgatk, double sgacc,
                          double xxlsgat, double xxlsgac, double pdensat,
                          double pdensac, double out[4]) {
  const double t0 = FREEZING_PT_H2O + 15.0;
  const double sqrt_temp = M_sqrt(airtemp);
  const double lamda = 6.6328e-8 * PRESSURE_STP * airtemp / (t0 * airprs);
  const double amu = 1.458e-6 * airtemp * sqrt_temp / (airtemp + 110.4);
  const double boltzmann = 1.3806500000000000e-023;
  const double knc = (2.0 / 3.0) * boltzmann * airtemp / amu;
  const double kfmat = M_sqrt(3.0 * boltzmann * airtemp / pdensat);
  const double kfmac = M_sqrt(3.0 * boltzmann * airtemp / pdensac);
  const double kfmatac = M_sqrt(6.0 * boltzmann * airtemp / (pdensat + pdensac));
  double qn11, qn22, qn12, qv12;
  getcoags(lamda, kfmatac, kfmat, kfmac, knc, dgatk, dgacc, sgatk, sgacc,
           xxlsgat, xxlsgac, &qn11, &qn22, &qn12, &qv12);
  out[2] = kmax(0.0, qn11);
  out[3] = kmax(0.0, qn22);
  out[0] = kmax(0.0, qn12);
  const double dumatk3 = (cube(dgatk) * M_exp(4.5 * xxlsgat * xxlsgat));
  out[1] = kmax(0.0, qv12 / dumatk3);
}

/* coagulation.hpp:849-928 */
static void mam_coag_aer_update(const double ybetaij3[3], double deltat,
                                const double qnum_tavg[4],
                                const double qaer_bgn[7][4],
                                double qaer_end[7][4],
                                double qaer_del_coag_out[7][1]) {
  for (int s = 0; s < 7; ++s) qaer_del_coag_out[s][0] = 0.0;
  const double bijqnumj1 = kmax(0.0, ybetaij3[0] * qnum_tavg[NACC]);
  const double bijqnumj2 = kmax(0.0, ybetaij3[2] * qnum_tavg[NPCA]);
  double decay_const = bijqnumj1 + bijqnumj2;
  const double epsilonx2 = DBL_EPSILON * 2.0;
  double decay_factor = deltat * decay_const;
  if (decay_factor > epsilonx2) {
    const double prtn2 = bijqnumj2 / decay_const;
    const double prtn1 = 1.0 - prtn2;
    const double tmp_xf = 1.0 - M_exp(-decay_factor);
    for (int s = 0; s < 7; ++s) {
      const double tmp_dq = tmp_xf * qaer_bgn[s][NAIT];
      qaer_end[s][NAIT] -= tmp_dq;
      qaer_end[s][NACC] += tmp_dq * prtn1;
      qaer_end[s][NPCA] += tmp_dq * prtn2;
      qaer_del_coag_out[s][0] += tmp_dq * prtn2;
    }
  }
  decay_const = kmax(0.0, ybetaij3[1] * qnum_tavg[NACC]);
  decay_factor = deltat * decay_const;
  if (decay_factor > epsilonx2) {
    const double tmp_xf = 1.0 - M_exp(-decay_factor);
    for (int s = 0; s < 7; ++s) {
      const double tmp_dq = tmp_xf * qaer_bgn[s][NPCA];
      qaer_end[s][NPCA] -= tmp_dq;
      qaer_end[s][NACC] += tmp_dq;
    }
  }
}

/* coagulation.hpp:937-953 */
static void update_qnum_intra_inter(double bijdtqnumj, double biidt,
                                    double qnumi_bgn, double *qnumi_end) {
  const double eps_denominator = 1.0e-5;
  if (bijdtqnumj < eps_denominator) {
    *qnumi_end = qnumi_bgn / (1.0 + (bijdtqnumj + biidt * qnumi_bgn) *
                                        (1.0 + 0.5 * bijdtqnumj));
  } else {
    const double tmp_exp = M_exp(-bijdtqnumj);
    *qnumi_end = qnumi_bgn * tmp_exp /
                 (1.0 + (biidt * qnumi_bgn / bijdtqnumj) * (1.0 - tmp_exp));
  }
}

/* coagulation.hpp:981-1027 */
static void mam_coag_num_update(const double ybetaij0[3], const double ybetaii0[3],
                                const double ybetajj0[3], double deltat,
                                const double qnum_bgn[4], double qnum_end[4],
                                double qnum_tavg[4]) {
  qnum_end[NACC] = qnum_bgn[NACC] / (1.0 + ybetajj0[0] * deltat * qnum_bgn[NACC]);
  qnum_tavg[NACC] = (qnum_bgn[NACC] + qnum_end[NACC]) * 0.5;
  double bijdtqnumj = kmax(0.0, deltat * ybetaij0[1] * qnum_tavg[NACC]);
  double biidt = kmax(0.0, deltat * ybetaii0[1]);
  update_qnum_intra_inter(bijdtqnumj, biidt, qnum_bgn[NPCA], &qnum_end[NPCA]);
  qnum_tavg[NPCA] = (qnum_bgn[NPCA] + qnum_end[NPCA]) * 0.5;
  double bijqnumj = ybetaij0[0] * qnum_tavg[NACC];
  bijqnumj = bijqnumj + ybetaij0[2] * qnum_tavg[NPCA];
  bijdtqnumj = kmax(0.0, deltat * bijqnumj);
  biidt = kmax(0.0, deltat * ybetaii0[0]);
  update_qnum_intra_inter(bijdtqnumj, biidt, qnum_bgn[NAIT], &qnum_end[NAIT]);
  qnum_tavg[NAIT] = (qnum_bgn[NAIT] + qnum_end[NAIT]) * 0.5;
}

/* coagulation.hpp:1037-1119 */
static void mam_coag_1subarea(double deltat, double temp, double pmid,
                              double aircon, const double dgn_awet[4],
                              const double wetdens[4], double qnum_cur[4],
                              double qaer_cur[7][4],
                              double qaer_del_coag_out[7][1]) {
  double qaer_bgn[7][4], qnum_bgn[4];
  for (int s = 0; s < 7; ++s)
    for (int m = 0; m < 4; ++m) {
      qaer_cur[s][m] = kmax(0.0, qaer_cur[s][m]);
      qaer_bgn[s][m] = qaer_cur[s][m];
    }
  for (int m = 0; m < 4; ++m) {
    qnum_cur[m] = kmax(0.0, qnum_cur[m]);
    qnum_bgn[m] = qnum_cur[m];
  }
  const int src[3] = {NAIT, NPCA, NAIT};
  const int dst[3] = {NACC, NACC, NPCA};
  double ij0[3], ij3[3], ii0[3], jj0[3];
  for (int ip = 0; ip < 3; ++ip) {
    const double ss = MODE_SIGMA[src[ip]], sd = MODE_SIGMA[dst[ip]];
    double o[4];
    orc_getcoags_wrapper(temp, pmid, dgn_awet[src[ip]], dgn_awet[dst[ip]], ss,
                         sd, M_log(ss), M_log(sd), wetdens[src[ip]],
                         wetdens[dst[ip]], o);
    ij0[ip] = o[0];
    ij3[ip] = o[1];
    ii0[ip] = o[2];
    jj0[ip] = o[3];
  }
  for (int ip = 0; ip < 3; ++ip) {
    ij0[ip] *= aircon;
    ij3[ip] *= aircon;
    ii0[ip] *= aircon;
    jj0[ip] *= aircon;
  }
  double qnum_tavg[4] = {0, 0, 0, 0};
  mam_coag_num_update(ij0, ii0, jj0, deltat, qnum_bgn, qnum_cur, qnum_tavg);
  mam_coag_aer_update(ij3, deltat, qnum_tavg, qaer_bgn, qaer_cur,
                      qaer_del_coag_out);
}

/* ======================= aging =========================================== */

/* aging.hpp:69-161 */
static void mam_pcarbon_aging_frac(unsigned nmono, const double dgn_a[4],
                                   const double qaer_cur[7][4],
                                   const double qaer_del_cond[7][4],
                                   const double qaer_del_coag_in[7][1],
                                   double *xferfrac_pcage, double *frac_cond,
                                   double *frac_coag) {
  const double mw_soa = 150 / 1000.0, mw_so4 = 115 / 1000.0;
  const double so4_vol = mw_so4 * 1000.0 / AER_DENS[A_SO4];
  const double soa_vol = mw_soa * 1000.0 / AER_DENS[A_SOA];
  const double hygro_soa = 0.14000000000000001, hygro_so4 = 0.50700000000000001;
  const double fac_m2v = soa_vol * hygro_soa / hygro_so4;
  const double vol_shell = qaer_cur[A_SO4][NPCA] * so4_vol + qaer_cur[A_SOA][NPCA] * fac_m2v;
  double del_cond = qaer_del_cond[A_SO4][NPCA] * so4_vol +
                    qaer_del_cond[A_SOA][NPCA] * fac_m2v;
  const double del_coag = qaer_del_coag_in[A_SO4][0] * so4_vol +
                          qaer_del_coag_in[A_SOA][0] * fac_m2v;
  del_cond = kmax(del_cond, 1e-35);
  *frac_cond = del_cond / (del_cond + kmax(del_coag, 0.0));
  *frac_coag = 1.0 - *frac_cond;
  double vol_core = 0.0;
  for (int s = 0; s < 7; ++s)
    if (LMAP_AER[s][NPCA] > 0) vol_core += qaer_cur[s][NPCA] * MASS_2_VOL[s];
  const double fac_volsfc = M_exp(2.5 * sq(M_log(MODE_SIGMA[NPCA])));
  const double xferfrac_max = 1.0 - 10.0 * DBL_EPSILON;
  const double t1 = vol_shell * dgn_a[NPCA] * fac_volsfc;
  const double dr = nmono * 4.76e-10;
  const double t2 = kmax(6.0 * dr * vol_core, 0.0);
  if (t1 >= t2)
    *xferfrac_pcage = xferfrac_max;
  else
    *xferfrac_pcage = kmin(t1 / t2, xferfrac_max);
}

/* aging.hpp:166-194 */
void orc_transfer_aged_pcarbon_to_accum(int nsrc, int ndest, double xferfrac,
                                        double frac_cond, double frac_coag,
                                        double q[4], double dcond[4],
                                        double dcoag[4]) {
  const double q_tmp = q[nsrc] * xferfrac;
  q[nsrc] -= q_tmp;
  q[ndest] += q_tmp;
  dcond[nsrc] -= q_tmp * frac_cond;
  dcond[ndest] += q_tmp * frac_cond;
  dcoag[nsrc] -= q_tmp * frac_coag;
  dcoag[ndest] += q_tmp * frac_coag;
}

/* aging.hpp:203-231 */
void orc_transfer_cond_coag_mass_to_accum(int nsrc, int ndest, double q[4],
                                          double dcond[4], double dcoag[4]) {
  q[ndest] += q[nsrc];
  dcond[ndest] += dcond[nsrc];
  dcoag[ndest] += dcoag[nsrc];
  q[nsrc] = 0.0;
  dcond[nsrc] = 0.0;
  dcoag[nsrc] = 0.0;
}

/* aging.hpp:237-358 */
void orc_pcarbon_aging_1subarea(unsigned nmono, const double dgn_a[4],
                                double qnum[4], double qnum_del_cond[4],
                                double qnum_del_coag[4], double qaer[7][4],
                                double qaer_del_cond[7][4],
                                double qaer_del_coag[7][4],
                                double qaer_del_coag_in[7][1]) {
  double xferfrac, frac_cond, frac_coag;
  mam_pcarbon_aging_frac(nmono, dgn_a, (const double(*)[4])qaer,
                         (const double(*)[4])qaer_del_cond,
                         (const double(*)[1])qaer_del_coag_in, &xferfrac,
                         &frac_cond, &frac_coag);
  static const int to_accum_aged[3] = {A_POM, A_BC, A_MOM};
  static const int to_accum_all[4] = {A_SOA, A_SO4, A_NCL, A_DST};
  for (int a = 0; a < 3; ++a) {
    const int s = to_accum_aged[a];
    orc_transfer_aged_pcarbon_to_accum(NPCA, NACC, xferfrac, frac_cond,
                                       frac_coag, qaer[s], qaer_del_cond[s],
                                       qaer_del_coag[s]);
  }
  for (int a = 0; a < 4; ++a) {
    const int s = to_accum_all[a];
    orc_transfer_cond_coag_mass_to_accum(NPCA, NACC, qaer[s], qaer_del_cond[s],
                                         qaer_del_coag[s]);
  }
  orc_transfer_aged_pcarbon_to_accum(NPCA, NACC, xferfrac, frac_cond, frac_coag,
                                     qnum, qnum_del_cond, qnum_del_coag);
}

/* ======================= rename ========================================== */

typedef struct {
  double mean_std_dev[4], fmode_dist_tail_fac[4], v2n_lo_rlx[4], v2n_hi_rlx[4],
      ln_diameter_tail_fac[4], diameter_cutoff[4], ln_dia_cutoff[4],
      diameter_threshold[4];
  int num_pairs;
} rename_consts;

/* rename.hpp:392-474 */
static void find_renaming_pairs(const int dest_mode_of_mode[4], rename_consts *c) {
  const double sqrt_half = M_sqrt(0.5);
  const double frelax = 27.0;
  c->num_pairs = 0;
  for (int m = 0; m < 4; ++m) {
    const int dest = dest_mode_of_mode[m];
    if (dest < 0) {
      c->mean_std_dev[m] = 0;
      c->fmode_dist_tail_fac[m] = 0;
      c->v2n_lo_rlx[m] = 0;
      c->v2n_hi_rlx[m] = 0;
      c->ln_diameter_tail_fac[m] = 0;
      c->diameter_cutoff[m] = 0;
      c->ln_dia_cutoff[m] = 0;
      c->diameter_threshold[m] = 0;
    } else {
      const double alnsg = M_log(MODE_SIGMA[m]);
      c->mean_std_dev[m] = MODE_SIGMA[m];
      c->fmode_dist_tail_fac[m] = sqrt_half / alnsg;
      c->v2n_lo_rlx[m] = 1.0 / orc_mean_particle_volume_from_diameter(MODE_DMIN[m], MODE_SIGMA[m]) * frelax;
      c->v2n_hi_rlx[m] = 1.0 / orc_mean_particle_volume_from_diameter(MODE_DMAX[m], MODE_SIGMA[m]) / frelax;
      c->ln_diameter_tail_fac[m] = 3.0 * sq(alnsg);
      c->num_pairs += 1;
      const double alnsg_d = M_log(MODE_SIGMA[dest]);
      c->diameter_cutoff[m] = M_sqrt(MODE_DNOM[m] * M_exp(1.5 * sq(alnsg)) *
                                   MODE_DNOM[dest] * M_exp(1.5 * sq(alnsg_d)));
      c->ln_dia_cutoff[m] = M_log(c->diameter_cutoff[m]);
      c->diameter_threshold[m] = 0.99 * c->diameter_cutoff[m];
    }
  }
}

/* rename.hpp:22-66 */
static void compute_dryvol_change_in_src_mode(const int dest_mode_of_mode[4],
                                              const double q_mmr[4][7],
                                              const double q_del[4][7],
                                              double dryvol[4],
                                              double deldryvol[4]) {
  for (int m = 0; m < 4; ++m) {
    if (dest_mode_of_mode[m] >= 0) {
      double v = 0, dv = 0;
      for (int s = 0; s < 7; ++s) {
        v += q_mmr[m][s] * MASS_2_VOL[s];
        dv += q_del[m][s] * MASS_2_VOL[s];
      }
      dryvol[m] = v - dv;
      deldryvol[m] = dv;
    }
  }
}

/* rename.hpp:142-150 */
static double tail_fraction(double diameter, double log_dia_cutoff,
                            double tail_dist_fac, double log_dia_tail_fac,
                            int with_tail_fac) {
  double logd = M_log(diameter);
  if (with_tail_fac) logd = logd + log_dia_tail_fac;
  const double tail = (log_dia_cutoff - logd) * tail_dist_fac;
  return 0.5 * M_erfc(tail);
}

/* Renaming's tail fraction of the Aitken -> accumulation pair for a mode of
 * geometric-mean diameter `diameter` (number, or volume with volume != 0),
 * and the pair's cutoff diameter (rename.hpp:142-150, 392-474). */
double orc_rename_tail_fraction(double diameter, int volume, double *cutoff) {
  static const int dest_mode_of_mode[4] = {-1, 0, -1, -1};
  rename_consts c;
  find_renaming_pairs(dest_mode_of_mode, &c);
  if (cutoff) *cutoff = c.diameter_cutoff[1];
  return tail_fraction(diameter, c.ln_dia_cutoff[1], c.fmode_dist_tail_fac[1],
                       c.ln_diameter_tail_fac[1], volume);
}

/* rename.hpp:225-389 */
static void do_inter_mode_transfer(const int dest_mode_of_mode[4], int is_cloudy,
                                   double smallest_dryvol, const rename_consts *c,
                                   const double dgnum_amode[4],
                                   const double dryvol_i[4], const double dryvol_c[4],
                                   const double deldryvol_i[4],
                                   const double deldryvol_c[4],
                                   double qmol_i[4][7], double qnum_i[4],
                                   double qmol_c[4][7], double qnum_c[4]) {
  for (int m = 0; m < 4; ++m) {
    const int src = m, dest = dest_mode_of_mode[m];
    if (dest < 0) continue;
    /* compute_before_growth_dryvol_and_num (rename.hpp:96-126) */
    double b4_dryvol = dryvol_i[src];
    if (is_cloudy) b4_dryvol = b4_dryvol + dryvol_c[src];
    const double b4_dryvol_bnd = kmax(b4_dryvol, smallest_dryvol);
    double b4_qnum = qnum_i[src];
    if (is_cloudy) b4_qnum = b4_qnum + qnum_c[src];
    b4_qnum = kmax(0.0, b4_qnum);
    const double b4_qnum_bnd = min_max_bound(b4_dryvol_bnd * c->v2n_hi_rlx[src],
                                             b4_dryvol_bnd * c->v2n_lo_rlx[src], b4_qnum);
    double dryvol_del = deldryvol_i[src];
    if (is_cloudy) dryvol_del = dryvol_del + deldryvol_c[src];
    const double aft_dryvol = b4_dryvol + dryvol_del;
    if (aft_dryvol <= smallest_dryvol) continue;
    const double b4_vol = b4_dryvol_bnd / b4_qnum_bnd;
    double b4_diam = orc_mean_particle_diameter_from_volume(b4_vol, c->mean_std_dev[src]);
    if (b4_diam > c->diameter_threshold[src]) {
      b4_dryvol = b4_dryvol * cube(c->diameter_threshold[src] / b4_diam);
      b4_diam = c->diameter_threshold[src];
    }
    if ((aft_dryvol - b4_dryvol) <= 1.0e-6 * b4_dryvol_bnd) continue;
    const double aft_vol = aft_dryvol / b4_qnum_bnd;
    const double aft_diam = orc_mean_particle_diameter_from_volume(aft_vol, c->mean_std_dev[src]);
    if (aft_diam <= dgnum_amode[src]) continue;
    const double b4_tail_num = tail_fraction(b4_diam, c->ln_dia_cutoff[src],
                                             c->fmode_dist_tail_fac[src], 0, 0);
    const double b4_tail_vol = tail_fraction(b4_diam, c->ln_dia_cutoff[src],
                                             c->fmode_dist_tail_fac[src],
                                             c->ln_diameter_tail_fac[src], 1);
    const double aft_tail_num = tail_fraction(aft_diam, c->ln_dia_cutoff[src],
                                              c->fmode_dist_tail_fac[src], 0, 0);
    const double aft_tail_vol = tail_fraction(aft_diam, c->ln_dia_cutoff[src],
                                              c->fmode_dist_tail_fac[src],
                                              c->ln_diameter_tail_fac[src], 1);
    /* compute_xfer_fractions (rename.hpp:164-198) */
    const double xferfrac_max = 1.0 - 10.0 * DBL_EPSILON;
    const double volume_fraction = aft_tail_vol * aft_dryvol - b4_tail_vol * b4_dryvol;
    if (volume_fraction <= 0.0) continue;
    double xfer_vol = kmin(volume_fraction, aft_dryvol) / aft_dryvol;
    xfer_vol = kmin(xfer_vol, xferfrac_max);
    double xfer_num = aft_tail_num - b4_tail_num;
    xfer_num = kmax(0.0, kmin(xfer_num, xfer_vol));
    /* do_num_and_mass_transfer (rename.hpp:200-222) */
    {
      const double nt = qnum_i[src] * xfer_num;
      qnum_i[src] -= nt;
      qnum_i[dest] += nt;
      for (int s = 0; s < 7; ++s) {
        const double vt = qmol_i[src][s] * xfer_vol;
        qmol_i[src][s] -= vt;
        qmol_i[dest][s] += vt;
      }
    }
    if (is_cloudy) {
      const double nt = qnum_c[src] * xfer_num;
      qnum_c[src] -= nt;
      qnum_c[dest] += nt;
      for (int s = 0; s < 7; ++s) {
        const double vt = qmol_c[src][s] * xfer_vol;
        qmol_c[src][s] -= vt;
        qmol_c[dest][s] += vt;
      }
    }
  }
}

/* rename.hpp:710-786 */
static void mam_rename_1subarea(int is_cloudy, double smallest_dryvol,
                                const int dest_mode_of_mode[4],
                                const rename_consts *c, const double dgnum_amode[4],
                                double qnum_i[4], double qmol_i[4][7],
                                const double qmol_i_del[4][7], double qnum_c[4],
                                double qmol_c[4][7], const double qmol_c_del[4][7]) {
  double dryvol_i[4] = {0, 0, 0, 0}, deldryvol_i[4] = {0, 0, 0, 0};
  compute_dryvol_change_in_src_mode(dest_mode_of_mode, (const double(*)[7])qmol_i,
                                    qmol_i_del, dryvol_i, deldryvol_i);
  double dryvol_c[4] = {0, 0, 0, 0}, deldryvol_c[4] = {0, 0, 0, 0};
  if (is_cloudy)
    compute_dryvol_change_in_src_mode(dest_mode_of_mode, (const double(*)[7])qmol_c,
                                      qmol_c_del, dryvol_c, deldryvol_c);
  do_inter_mode_transfer(dest_mode_of_mode, is_cloudy, smallest_dryvol, c,
                         dgnum_amode, dryvol_i, dryvol_c, deldryvol_i,
                         deldryvol_c, qmol_i, qnum_i, qmol_c, qnum_c);
}

/* ======================= subarea driver ================================== */

/* mam4_amicphys.hpp:1126-1694 (ntsubstep = 1).  delaa arrays [x][4]. */
static int mam_amicphys_1subarea(
    int gaexch_h2so4_uptake_optaa, int do_cond, int do_rename, int do_newnuc,
    int do_coag, double deltat, int ntsubstep, int iscldy, unsigned nmono, double temp,
    double pmid, double zmid, double pblh, double relhum, const double dgn_a[4],
    const double dgn_awet[4], const double wetdens[4], const double qgas1[2],
    const double qgas3[2], double qgas_cur[2], double qgas_delaa[2][4],
    const double qnum3[4], double qnum_cur[4], double qnum_delaa[4][4],
    const double qaer2[7][4], const double qaer3[7][4], double qaer_cur[7][4],
    double qaer_delaa[7][4][4], const double qnumcw3[4], double qnumcw_cur[4],
    double qnumcw_delaa[4][1], const double qaercw2[7][4],
    const double qaercw3[7][4], double qaercw_cur[7][4],
    double qaercw_delaa[7][4][1], double *qnum_cond_only,
    double *qaer_cond_only, double *qnum_coag_aged, double *qaer_coag_aged) {
  int niter = 0;
  const double r_universal = 8314.46759100000;
  const double aircon = pmid / (r_universal * temp);
  for (int g = 0; g < 2; ++g) qgas_cur[g] = qgas3[g];
  for (int s = 0; s < 7; ++s)
    for (int m = 0; m < 4; ++m) qaer_cur[s][m] = qaer3[s][m];
  for (int m = 0; m < 4; ++m) qnum_cur[m] = qnum3[m];
  if (iscldy) {
    for (int m = 0; m < 4; ++m) qnumcw_cur[m] = qnumcw3[m];
    for (int s = 0; s < 7; ++s)
      for (int m = 0; m < 4; ++m) qaercw_cur[s][m] = qaercw3[s][m];
  }
  double netprod[2] = {0.0, 0.0};
  if (do_cond && gaexch_h2so4_uptake_optaa == 2) {
    for (int g = 0; g < 2; ++g)
      if (g == IGAS_H2SO4 || g == IGAS_NH3) {
        netprod[g] = (qgas3[g] - qgas1[g]) / deltat;
        /* QUIRK (:1283): a negative net production zeroes the gas. */
        qgas_cur[g] = (netprod[g] >= 0) ? qgas1[g] : 0;
      }
  }
  memset(qgas_delaa, 0, sizeof(double) * 2 * 4);
  memset(qnum_delaa, 0, sizeof(double) * 4 * 4);
  memset(qaer_delaa, 0, sizeof(double) * 7 * 4 * 4);
  memset(qnumcw_delaa, 0, sizeof(double) * 4);
  memset(qaercw_delaa, 0, sizeof(double) * 7 * 4);
  /* QUIRK (:1319): the sub-step length is an int.  The reference fixes
   * ntsubstep = 1 (:1288); for more, the loop below runs the whole subarea
   * chain ntsubstep times, with the H2SO4 uptake rates and their sum from the
   * first sub-step held for the later ones (the Fortran original's scope of
   * uptkaer / uptkrate_h2so4; the C++ declares both inside the loop, :1344,
   * :1377, and computes them only at jtsubstep == 0, gasaerexch.hpp:428). */
  const int dtsubstep = (int)(deltat / ntsubstep);
  double qgas_sv1[2], qnum_sv1[4], qaer_sv1[7][4];
  double del_h2so4_aeruptk;
  double qgas_avg[2];
  double qnum_delsub_cond[4] = {0, 0, 0, 0}, qnum_delsub_coag[4] = {0, 0, 0, 0};
  double qaer_delsub_cond[7][4], qaer_delsub_coag[7][4],
      qaer_delsub_grow4rnam[7][4], qaercw_delsub_grow4rnam[7][4];
  double qaer_delsub_coag_in[7][1];
  memset(qaer_delsub_cond, 0, sizeof(qaer_delsub_cond));
  memset(qaer_delsub_coag, 0, sizeof(qaer_delsub_coag));
  memset(qaer_delsub_grow4rnam, 0, sizeof(qaer_delsub_grow4rnam));
  memset(qaercw_delsub_grow4rnam, 0, sizeof(qaercw_delsub_grow4rnam));
  memset(qaer_delsub_coag_in, 0, sizeof(qaer_delsub_coag_in));
  double uptkrate_h2so4 = 0;
  double uptkaer[2][4];
  for (int jtsubstep = 0; jtsubstep < ntsubstep; ++jtsubstep) {
    if (do_cond) {
      for (int g = 0; g < 2; ++g) qgas_sv1[g] = qgas_cur[g];
      for (int m = 0; m < 4; ++m) qnum_sv1[m] = qnum_cur[m];
      memcpy(qaer_sv1, qaer_cur, sizeof(qaer_sv1));
      niter += mam_gasaerexch_1subarea(jtsubstep, (double)dtsubstep, temp, pmid, aircon,
                                       qgas_cur, qgas_avg, netprod, qaer_cur,
                                       qnum_cur, dgn_awet, uptkaer, &uptkrate_h2so4);
      for (int g = 0; g < 2; ++g)
        qgas_delaa[g][0] = qgas_delaa[g][0] +
                           (qgas_cur[g] - (qgas_sv1[g] + netprod[g] * dtsubstep));
      for (int m = 0; m < 4; ++m) qnum_delsub_cond[m] = qnum_cur[m] - qnum_sv1[m];
      for (int s = 0; s < 7; ++s)
        for (int m = 0; m < 4; ++m)
          qaer_delsub_cond[s][m] = qaer_cur[s][m] - qaer_sv1[s][m];
      del_h2so4_aeruptk = qgas_cur[IGAS_H2SO4] -
                          (qgas_sv1[IGAS_H2SO4] + netprod[IGAS_H2SO4] * dtsubstep);
      /* condensation's own increments, before aging re-attributes the
       * primary-carbon share (the Fortran's iqtend_cond_only column) */
      if (qnum_cond_only)
        for (int m = 0; m < 4; ++m) qnum_cond_only[m] = qnum_delsub_cond[m];
      if (qaer_cond_only)
        memcpy(qaer_cond_only, qaer_delsub_cond, sizeof(qaer_delsub_cond));
    } else {
      for (int g = 0; g < 2; ++g) qgas_avg[g] = qgas_cur[g];
      memset(qaer_delsub_cond, 0, sizeof(qaer_delsub_cond));
      for (int m = 0; m < 4; ++m) qnum_delsub_cond[m] = 0;
      del_h2so4_aeruptk = 0;
    }
    (void)del_h2so4_aeruptk;

    if (do_rename) {
      static const int dest_mode_of_mode[4] = {-1, 0, -1, -1};
      memcpy(qaer_delsub_grow4rnam, qaer_delsub_cond, sizeof(qaer_delsub_cond));
      if (iscldy) {
        for (int s = 0; s < 7; ++s)
          for (int m = 0; m < 4; ++m) {
            qaer_delsub_grow4rnam[s][m] =
                (qaer3[s][m] - qaer2[s][m]) / ntsubstep + qaer_delsub_grow4rnam[s][m];
            qaercw_delsub_grow4rnam[s][m] = (qaercw3[s][m] - qaercw2[s][m]) / ntsubstep;
          }
      }
      for (int m = 0; m < 4; ++m) qnum_sv1[m] = qnum_cur[m];
      memcpy(qaer_sv1, qaer_cur, sizeof(qaer_sv1));
      double qnumcw_sv1[4], qaercw_sv1[7][4];
      for (int m = 0; m < 4; ++m) qnumcw_sv1[m] = qnumcw_cur[m];
      memcpy(qaercw_sv1, qaercw_cur, sizeof(qaercw_sv1));
      rename_consts rc;
      find_renaming_pairs(dest_mode_of_mode, &rc);
      double dgnum_amode[4];
      for (int m = 0; m < 4; ++m) dgnum_amode[m] = MODE_DNOM[m];
      const double smallest_dryvol_value = 1.0e-25;
      double qa[4][7], qad[4][7], qc[4][7], qcd[4][7];
      for (int s = 0; s < 7; ++s)
        for (int m = 0; m < 4; ++m) {
          qa[m][s] = qaer_cur[s][m];
          qad[m][s] = qaer_delsub_grow4rnam[s][m];
          qc[m][s] = qaercw_cur[s][m];
          qcd[m][s] = qaercw_delsub_grow4rnam[s][m];
        }
      mam_rename_1subarea(iscldy, smallest_dryvol_value, dest_mode_of_mode, &rc,
                          dgnum_amode, qnum_cur, qa, (const double(*)[7])qad,
                          qnumcw_cur, qc, (const double(*)[7])qcd);
      for (int s = 0; s < 7; ++s)
        for (int m = 0; m < 4; ++m) {
          qaer_cur[s][m] = qa[m][s];
          qaercw_cur[s][m] = qc[m][s];
        }
      for (int m = 0; m < 4; ++m) qnum_delaa[m][1] += (qnum_cur[m] - qnum_sv1[m]);
      for (int s = 0; s < 7; ++s)
        for (int m = 0; m < 4; ++m)
          qaer_delaa[s][m][1] += (qaer_cur[s][m] - qaer_sv1[s][m]);
      if (iscldy)
        for (int m = 0; m < 4; ++m)
          qnumcw_delaa[m][0] += (qnumcw_cur[m] - qnumcw_sv1[m]);
      for (int s = 0; s < 7; ++s)
        for (int m = 0; m < 4; ++m)
          qaercw_delaa[s][m][0] += (qaercw_cur[s][m] - qaercw_sv1[s][m]);
    }

    if (do_newnuc) {
      for (int g = 0; g < 2; ++g) qgas_sv1[g] = qgas_cur[g];
      for (int m = 0; m < 4; ++m) qnum_sv1[m] = qnum_cur[m];
      memcpy(qaer_sv1, qaer_cur, sizeof(qaer_sv1));
      mam_newnuc_1subarea((double)dtsubstep, temp, pmid, zmid, pblh, relhum,
                          uptkrate_h2so4, qgas_cur, qgas_avg, qnum_cur, qaer_cur);
      for (int g = 0; g < 2; ++g) qgas_delaa[g][2] += (qgas_cur[g] - qgas_sv1[g]);
      for (int m = 0; m < 4; ++m) qnum_delaa[m][2] += (qnum_cur[m] - qnum_sv1[m]);
      for (int s = 0; s < 7; ++s)
        for (int m = 0; m < 4; ++m)
          qaer_delaa[s][m][2] += (qaer_cur[s][m] - qaer_sv1[s][m]);
    }

    if (do_coag) {
      for (int m = 0; m < 4; ++m) qnum_sv1[m] = qnum_cur[m];
      memcpy(qaer_sv1, qaer_cur, sizeof(qaer_sv1));
      mam_coag_1subarea((double)dtsubstep, temp, pmid, aircon, dgn_awet, wetdens,
                        qnum_cur, qaer_cur, qaer_delsub_coag_in);
      for (int m = 0; m < 4; ++m) qnum_delsub_coag[m] = qnum_cur[m] - qnum_sv1[m];
      for (int s = 0; s < 7; ++s)
        for (int m = 0; m < 4; ++m)
          qaer_delsub_coag[s][m] = qaer_cur[s][m] - qaer_sv1[s][m];
      for (int m = 0; m < 4; ++m) qnum_delaa[m][3] += qnum_delsub_coag[m];
      for (int s = 0; s < 7; ++s)
        for (int m = 0; m < 4; ++m) qaer_delaa[s][m][3] += qaer_delsub_coag[s][m];
    } else {
      memset(qaer_delsub_coag_in, 0, sizeof(qaer_delsub_coag_in));
      memset(qaer_delsub_coag, 0, sizeof(qaer_delsub_coag));
      for (int m = 0; m < 4; ++m) qnum_delsub_coag[m] = 0;
    }

    const int do_aging = (!iscldy) || (iscldy && do_cond);
    if (do_aging)
      orc_pcarbon_aging_1subarea(nmono, dgn_a, qnum_cur, qnum_delsub_cond,
                                 qnum_delsub_coag, qaer_cur, qaer_delsub_cond,
                                 qaer_delsub_coag, qaer_delsub_coag_in);
    /* coagulation's increments after aging re-attributed their primary-carbon
     * share: the Fortran's iqtend_coag column (the reference accumulates its
     * column before aging, :1631-1640) */
    if (qnum_coag_aged)
      for (int m = 0; m < 4; ++m) qnum_coag_aged[m] = qnum_delsub_coag[m];
    if (qaer_coag_aged) memcpy(qaer_coag_aged, qaer_delsub_coag, sizeof(qaer_delsub_coag));
    if (do_cond) {
      for (int m = 0; m < 4; ++m) qnum_delaa[m][0] += qnum_delsub_cond[m];
      for (int s = 0; s < 7; ++s)
        for (int m = 0; m < 4; ++m) qaer_delaa[s][m][0] += qaer_delsub_cond[s][m];
    }
  }
  return niter;
}

/* One call of mam_amicphys_1subarea (mam4_amicphys.hpp:1126-1694) with flat
 * arrays, for the known-answer vectors of one clear and one cloudy subarea.
 * Gases qgas[2] = {SOAG, H2SO4}; masses qaer[7][4] in the amicphys species
 * order (SOA, SO4, POM, BC, NaCl, dust, MOM) x mode (acc, Aitken, coarse,
 * primary carbon); increments qgas_delaa[2][4], qnum_delaa[4][4],
 * qaer_delaa[7][4][4] over (cond, rename, newnuc, coag); cloud-borne
 * increments [4][1] and [7][4][1].  qnum_cond_only[4] / qaer_cond_only[7][4]
 * (may be NULL) receive condensation's increments before aging, and
 * qnum_coag_aged[4] / qaer_coag_aged[7][4] (may be NULL) coagulation's after
 * it.  Returns the SOA sub-step count. */
int orc_amicphys_1subarea(int gaexch_h2so4_uptake_optaa, int do_cond, int do_rename,
                          int do_newnuc, int do_coag, double deltat, int iscldy,
                          unsigned nmono, double temp, double pmid, double zmid,
                          double pblh, double relhum, const double *dgn_a,
                          const double *dgn_awet, const double *wetdens,
                          const double *qgas1, const double *qgas3, double *qgas_cur,
                          double *qgas_delaa, const double *qnum3, double *qnum_cur,
                          double *qnum_delaa, const double *qaer2, const double *qaer3,
                          double *qaer_cur, double *qaer_delaa, const double *qnumcw3,
                          double *qnumcw_cur, double *qnumcw_delaa,
                          const double *qaercw2, const double *qaercw3,
                          double *qaercw_cur, double *qaercw_delaa,
                          double *qnum_cond_only, double *qaer_cond_only,
                          double *qnum_coag_aged, double *qaer_coag_aged) {
  return mam_amicphys_1subarea(
      gaexch_h2so4_uptake_optaa, do_cond, do_rename, do_newnuc, do_coag, deltat, 1, iscldy,
      nmono, temp, pmid, zmid, pblh, relhum, dgn_a, dgn_awet, wetdens, qgas1, qgas3,
      qgas_cur, (double(*)[4])qgas_delaa, qnum3, qnum_cur, (double(*)[4])qnum_delaa,
      (const double(*)[4])qaer2, (const double(*)[4])qaer3, (double(*)[4])qaer_cur,
      (double(*)[4][4])qaer_delaa, qnumcw3, qnumcw_cur, (double(*)[1])qnumcw_delaa,
      (const double(*)[4])qaercw2, (const double(*)[4])qaercw3, (double(*)[4])qaercw_cur,
      (double(*)[4][1])qaercw_delaa, qnum_cond_only, qaer_cond_only, qnum_coag_aged,
      qaer_coag_aged);
}

/* mam4_amicphys.hpp:1699-1991 */
static int mam_amicphys_1gridcell(
    const orc_config *cfg, double deltat, int nsubarea, int ncldy_subarea,
    const int iscldy[MAXSUB], const double afracsub[MAXSUB], unsigned nmono,
    double temp, double pmid, double zmid, double pblh,
    const double relhumsub[MAXSUB], const double dgn_a[4],
    const double dgn_awet[4], const double wetdens[4],
    double qsub1[31][MAXSUB], double qsub2[31][MAXSUB],
    double qqcwsub2[31][MAXSUB], double qsub3[31][MAXSUB],
    double qqcwsub3[31][MAXSUB], double qaerwatsub3[4][MAXSUB],
    double qsub4[31][MAXSUB], double qqcwsub4[31][MAXSUB],
    double qaerwatsub4[4][MAXSUB], double qsub_tendaa[31][4][MAXSUB],
    double qqcwsub_tendaa[31][1][MAXSUB]) {
  (void)ncldy_subarea;
  int niter = 0;
  const int mdo_gaexch_cldy_subarea = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = 1; j < MAXSUB; ++j) {
      qsub4[i][j] = qsub3[i][j];
      qqcwsub4[i][j] = qqcwsub3[i][j];
    }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < MAXSUB; ++j) qaerwatsub4[i][j] = qaerwatsub3[i][j];
  for (int jsub = 1; jsub <= nsubarea; ++jsub) {
    int do_cond, do_rename, do_newnuc, do_coag;
    if (iscldy[jsub]) {
      do_cond = cfg->do_cond;
      do_rename = cfg->do_rename;
      do_newnuc = 0;
      do_coag = 0;
      if (mdo_gaexch_cldy_subarea <= 0) do_cond = 0;
    } else {
      do_cond = cfg->do_cond;
      do_rename = cfg->do_rename;
      do_newnuc = cfg->do_newnuc;
      do_coag = cfg->do_coag;
    }
    const int do_map_gas = do_cond || do_newnuc;
    double qgas1[2] = {0, 0}, qgas3[2] = {0, 0}, qgas4[2] = {0, 0};
    if (do_map_gas)
      for (int g = 0; g < 2; ++g) {
        const int l = LMAP_GAS[g];
        qgas1[g] = qsub1[l][jsub] * fcvt_gas(g);
        qgas3[g] = qsub3[l][jsub] * fcvt_gas(g);
        qgas4[g] = qgas3[g];
      }
    double qaer2[7][4], qaer3[7][4], qaer4[7][4];
    double qnum3[4] = {0, 0, 0, 0}, qnum4[4] = {0, 0, 0, 0};
    double qwtr3[4] = {0, 0, 0, 0};
    memset(qaer2, 0, sizeof(qaer2));
    memset(qaer3, 0, sizeof(qaer3));
    memset(qaer4, 0, sizeof(qaer4));
    for (int m = 0; m < 4; ++m) {
      const int ln = NUMPTR[m];
      qnum3[m] = qsub3[ln][jsub] * FCVT_NUM;
      qnum4[m] = qnum3[m];
      for (int a = 0; a < 7; ++a) {
        const int la = LMAP_AER[a][m];
        if (la > 0) {
          qaer2[a][m] = qsub2[la][jsub] * FCVT_AER[a];
          qaer3[a][m] = qsub3[la][jsub] * FCVT_AER[a];
          qaer4[a][m] = qaer3[a][m];
        }
      }
      qwtr3[m] = qaerwatsub3[m][jsub] * fcvt_wtr();
    }
    double qaercw2[7][4], qaercw3[7][4], qaercw4[7][4];
    double qnumcw3[4] = {0, 0, 0, 0}, qnumcw4[4] = {0, 0, 0, 0};
    memset(qaercw2, 0, sizeof(qaercw2));
    memset(qaercw3, 0, sizeof(qaercw3));
    memset(qaercw4, 0, sizeof(qaercw4));
    if (iscldy[jsub]) {
      for (int m = 0; m < 4; ++m) {
        const int ln = NUMPTR[m];
        qnumcw3[m] = qqcwsub3[ln][jsub] * FCVT_NUM;
        qnumcw4[m] = qnumcw3[m];
      }
      for (int a = 0; a < 7; ++a)
        for (int m = 0; m < 4; ++m) {
          const int la = LMAP_AER[a][m];
          if (la > 0) {
            qaercw2[a][m] = qqcwsub2[la][jsub] * FCVT_AER[a];
            qaercw3[a][m] = qqcwsub3[la][jsub] * FCVT_AER[a];
            qaercw4[a][m] = qaercw3[a][m];
          }
        }
    }
    double qgas_delaa[2][4], qnum_delaa[4][4], qnumcw_delaa[4][1];
    double qaer_delaa[7][4][4], qaercw_delaa[7][4][1];
    /* QUIRK (:1914): the optaa arguments arrive swapped; only the one read
     * as gaexch_h2so4_uptake_optaa matters. */
    niter += mam_amicphys_1subarea(
        cfg->newnuc_h2so4_conc_optaa, do_cond, do_rename, do_newnuc, do_coag,
        deltat, cfg->ntsubstep > 1 ? cfg->ntsubstep : 1, iscldy[jsub], nmono, temp, pmid,
        zmid, pblh, relhumsub[jsub],
        dgn_a, dgn_awet, wetdens, qgas1, qgas3, qgas4, qgas_delaa, qnum3, qnum4,
        qnum_delaa, (const double(*)[4])qaer2, (const double(*)[4])qaer3, qaer4,
        qaer_delaa, qnumcw3, qnumcw4, qnumcw_delaa, (const double(*)[4])qaercw2,
        (const double(*)[4])qaercw3, qaercw4, qaercw_delaa, NULL, NULL, NULL, NULL);
    if (do_map_gas)
      for (int g = 0; g < 2; ++g) {
        const int ll = LMAP_GAS[g];
        qsub4[ll][jsub] = qgas4[g] / fcvt_gas(g);
        for (int jj = 0; jj < 4; ++jj)
          qsub_tendaa[ll][jj][jsub] = qgas_delaa[g][jj] / (fcvt_gas(g) * deltat);
      }
    for (int m = 0; m < 4; ++m) {
      const int ll = NUMPTR[m];
      qsub4[ll][jsub] = qnum4[m] / (FCVT_NUM);
      for (int jj = 0; jj < 4; ++jj)
        qsub_tendaa[ll][jj][jsub] = qnum_delaa[m][jj] / (FCVT_NUM * deltat);
      for (int a = 0; a < 7; ++a) {
        const int la = LMAP_AER[a][m];
        if (la > 0) {
          qsub4[la][jsub] = qaer4[a][m] / FCVT_AER[a];
          for (int jj = 0; jj < 4; ++jj)
            qsub_tendaa[la][jj][jsub] = qaer_delaa[a][m][jj] / (FCVT_AER[a] * deltat);
        }
      }
      qaerwatsub4[m][jsub] = 0.0 / fcvt_wtr(); /* qwtr4 stays qwtr3 = 0 */
      if (iscldy[jsub]) {
        const int lc = NUMPTR[m];
        qqcwsub4[lc][jsub] = qnumcw4[m] / FCVT_NUM;
        qqcwsub_tendaa[lc][0][jsub] = qnumcw_delaa[m][0] / (FCVT_NUM * deltat);
        for (int a = 0; a < 7; ++a) {
          const int lca = LMAP_AER[a][m];
          if (lca > 0) {
            qqcwsub4[lca][jsub] = qaercw4[a][m] / FCVT_AER[a];
            qqcwsub_tendaa[lca][0][jsub] = qaercw_delaa[a][m][0] / (FCVT_AER[a] * deltat);
          }
        }
      }
    }
    (void)qwtr3;
  }
  return niter;
}

/* mam4_amicphys.hpp:1996-2063 */
static void form_gcm_from_subareas(int nsubarea, int ncldy_subarea,
                                   const double afracsub[MAXSUB],
                                   double qsub[31][MAXSUB], double qqcwsub[31][MAXSUB],
                                   const double qqcwgcm_old[31], double qgcm[31],
                                   double qqcwgcm[31]) {
  for (int i = 0; i < 31; ++i) qgcm[i] = 0.0;
  for (int j = 1; j <= nsubarea; ++j)
    for (int i = 0; i < 31; ++i) qgcm[i] += qsub[i][j] * afracsub[j];
  for (int i = 0; i < 31; ++i) qgcm[i] = kmax(0, qgcm[i]);
  if (ncldy_subarea <= 0) {
    for (int i = 0; i < 31; ++i) qqcwgcm[i] = qqcwgcm_old[i];
  } else {
    for (int i = 0; i < 31; ++i) qqcwgcm[i] = 0.0;
    for (int j = 1; j <= nsubarea; ++j)
      for (int i = 0; i < 31; ++i) qqcwgcm[i] += qqcwsub[i][j] * afracsub[j];
  }
}

/* mam4_amicphys.hpp:2119-2376 */
void orc_amicphys_intr(const orc_config *cfg, double deltat, double temp,
                       double pmid, double pdel, double zm, double pblh,
                       double qv, double cld, double qq[31], double qqcw[31],
                       const double q_pregaschem[31],
                       const double q_precldchem[31],
                       const double qqcw_precldchem[31],
                       const double dgncur_a[4], const double dgncur_awet[4],
                       const double wetdens_host[4], double *qgcm_tendaa,
                       double *qqcwgcm_tendaa, int stats[2]) {
  (void)pdel;
  const double epsqs = MW_H2O / MW_DRY_AIR;
  const double ev_sat = vapor_saturation_pressure_magnus(temp, pmid);
  const double qv_sat = epsqs * ev_sat / (pmid - (1 - epsqs) * ev_sat);
  int nsubarea, ncldy_subarea, jclea, jcldy, iscldy[MAXSUB];
  double afracsub[MAXSUB], fcldy, fclea;
  setup_subareas(cld, &nsubarea, &ncldy_subarea, &jclea, &jcldy, iscldy,
                 afracsub, &fclea, &fcldy);
  const double relhumgcm = kmax(0.0, kmin(1.0, qv / qv_sat));
  double relhumsub[MAXSUB];
  set_subarea_rh(ncldy_subarea, jclea, jcldy, afracsub, relhumgcm, relhumsub);
  double qaerwatsub3[4][MAXSUB];
  memset(qaerwatsub3, 0, sizeof(qaerwatsub3));
  double qgcm1[31], qgcm2[31], qgcm3[31], qqcwgcm2[31], qqcwgcm3[31];
  for (int i = 0; i < 31; ++i) {
    qgcm1[i] = kmax(0, q_pregaschem[i]);
    qgcm2[i] = kmax(0, q_precldchem[i]);
    qgcm3[i] = kmax(0, qq[i]);
    qqcwgcm2[i] = kmax(0, qqcw_precldchem[i]);
    qqcwgcm3[i] = kmax(0, qqcw[i]);
  }
  double qsub1[31][MAXSUB], qsub2[31][MAXSUB], qsub3[31][MAXSUB],
      qqcwsub2[31][MAXSUB], qqcwsub3[31][MAXSUB];
  set_subarea_gases_and_aerosols(nsubarea, jclea, jcldy, fclea, fcldy, qgcm1,
                                 qgcm2, qqcwgcm2, qgcm3, qqcwgcm3, qsub1, qsub2,
                                 qqcwsub2, qsub3, qqcwsub3);
  double qsub4[31][MAXSUB], qqcwsub4[31][MAXSUB], qaerwatsub4[4][MAXSUB];
  memset(qsub4, 0, sizeof(qsub4));
  memset(qqcwsub4, 0, sizeof(qqcwsub4));
  memset(qaerwatsub4, 0, sizeof(qaerwatsub4));
  double dgn_a[4], dgn_awet[4], wetdens[4];
  for (int m = 0; m < 4; ++m) {
    dgn_a[m] = dgncur_a[m];
    dgn_awet[m] = dgncur_awet[m];
    wetdens[m] = kmax(1000.0, wetdens_host[m]);
  }
  static double dummy_t;
  double qsub_tendaa[31][4][MAXSUB], qqcwsub_tendaa[31][1][MAXSUB];
  memset(qsub_tendaa, 0, sizeof(qsub_tendaa));
  memset(qqcwsub_tendaa, 0, sizeof(qqcwsub_tendaa));
  (void)dummy_t;
  const int niter = mam_amicphys_1gridcell(
      cfg, deltat, nsubarea, ncldy_subarea, iscldy, afracsub,
      cfg->n_so4_monolayers_pcage, temp, pmid, zm, pblh, relhumsub, dgn_a,
      dgn_awet, wetdens, qsub1, qsub2, qqcwsub2, qsub3, qqcwsub3, qaerwatsub3,
      qsub4, qqcwsub4, qaerwatsub4, qsub_tendaa, qqcwsub_tendaa);
  double qgcm4[31], qqcwgcm4[31];
  form_gcm_from_subareas(nsubarea, ncldy_subarea, afracsub, qsub4, qqcwsub4,
                         qqcwgcm3, qgcm4, qqcwgcm4);
  for (int i = 0; i < 31; ++i) {
    if (LMAPCC_ALL[i] > 0) qq[i] = qgcm4[i];
    if (LMAPCC_ALL[i] >= LMAPCC_AER) qqcw[i] = qqcwgcm4[i];
  }
  /* get_gcm_tend_diags_from_subareas (:2069-2113) */
  if (qgcm_tendaa) {
    for (int i = 0; i < 31 * 4; ++i) qgcm_tendaa[i] = 0.0;
    for (int j = 1; j <= nsubarea; ++j)
      for (int iq = 0; iq < 4; ++iq)
        for (int i = 0; i < 31; ++i)
          qgcm_tendaa[i * 4 + iq] += qsub_tendaa[i][iq][j] * afracsub[j];
  }
  if (qqcwgcm_tendaa) {
    for (int i = 0; i < 31; ++i) qqcwgcm_tendaa[i] = 0.0;
    if (ncldy_subarea > 0)
      for (int j = 1; j <= nsubarea; ++j)
        for (int i = 0; i < 31; ++i)
          qqcwgcm_tendaa[i] += qqcwsub_tendaa[i][0][j] * afracsub[j];
  }
  if (stats) {
    stats[0] = niter;
    stats[1] = nsubarea;
  }
}

/* ======================= calcsize ======================================== */

typedef struct {
  double v2n_nom[4], v2n_min[4], v2n_max[4], dgnmin[4], dgnmax[4], dgnnom[4],
      sigma[4], inv_density[4][7];
  int noxf_acc2ait[7], n_common, ait_spec_in_acc[7], acc_spec_in_ait[7];
} calcsize_consts;

/* CalcSize::init (calcsize.hpp:893-960) */
static void calcsize_init(calcsize_consts *c) {
  int count = 0;
  for (int isp = 0; isp < NSPEC_MODE[NACC]; ++isp) {
    c->noxf_acc2ait[isp] = 1;
    for (int jsp = 0; jsp < NSPEC_MODE[NAIT]; ++jsp) {
      if (MODE_SPEC[NACC][isp] == MODE_SPEC[NAIT][jsp]) {
        c->noxf_acc2ait[isp] = 0;
        c->acc_spec_in_ait[count] = isp;
        c->ait_spec_in_acc[count] = jsp;
        count++;
        break;
      }
    }
  }
  c->n_common = count;
  for (int m = 0; m < 4; ++m) {
    c->dgnnom[m] = MODE_DNOM[m];
    c->dgnmin[m] = MODE_DMIN[m];
    c->dgnmax[m] = MODE_DMAX[m];
    c->sigma[m] = MODE_SIGMA[m];
    c->v2n_nom[m] = 1.0 / orc_mean_particle_volume_from_diameter(c->dgnnom[m], MODE_SIGMA[m]);
    c->v2n_min[m] = 1.0 / orc_mean_particle_volume_from_diameter(c->dgnmax[m], MODE_SIGMA[m]);
    c->v2n_max[m] = 1.0 / orc_mean_particle_volume_from_diameter(c->dgnmin[m], MODE_SIGMA[m]);
    for (int s = 0; s < 7; ++s) c->inv_density[m][s] = 0.0;
    for (int s = 0; s < NSPEC_MODE[m]; ++s)
      c->inv_density[m][s] = 1.0 / AER_DENS[MODE_SPEC[m][s]];
  }
}

static inline double safe_denominator(double x) {
  const double tol = DBL_EPSILON;
  return x / (x * x + tol * tol);
}

/* calcsize.hpp:55-81 */
static void update_diameter_and_vol2num(double drv, double num, double v2nmin,
                                        double v2nmax, double dgnmin,
                                        double dgnmax, double sigma,
                                        double *dgncur, double *v2ncur) {
  if (!(drv > 0.0)) return;
  if (num <= drv * v2nmin) {
    *dgncur = dgnmax;
    *v2ncur = v2nmin;
  } else if (num >= drv * v2nmax) {
    *dgncur = dgnmin;
    *v2ncur = v2nmax;
  } else {
    *v2ncur = num / drv;
    *dgncur = orc_mean_particle_diameter_from_volume(1.0 / *v2ncur, sigma);
  }
}

/* calcsize.hpp:94-311 */
static void adjust_num_sizes(double drv_i, double drv_c, double init_num_i,
                             double init_num_c, double dt, double v2nmin,
                             double v2nmax, double adj_tscale_inv,
                             double *num_i, double *num_c, double *dqdt,
                             double *dqqcwdt) {
  const double zero = 0.0, one = 1.0;
  const double frac = kmax(zero, kmin(one, dt * adj_tscale_inv));
  const double dtinv = safe_denominator(dt);
  const double relax = 27.0;
  const double v2nminrl = v2nmin / relax, v2nmaxrl = v2nmax * relax;
  if (drv_i <= zero && drv_c <= zero) {
    *num_i = zero;
    *num_c = zero;
  } else if (drv_c <= zero) {
    *num_c = zero;
    const double nb = min_max_bound(drv_i * v2nmin, drv_i * v2nmax, *num_i);
    *num_i = *num_i + (nb - *num_i) * frac;
  } else if (drv_i <= zero) {
    const double nb = min_max_bound(drv_c * v2nmin, drv_c * v2nmax, *num_c);
    *num_c = *num_c + (nb - *num_c) * frac;
    *num_i = zero;
  } else {
    const double ni1 = *num_i, nc1 = *num_c;
    double nb = min_max_bound(drv_i * v2nminrl, drv_i * v2nmaxrl, ni1);
    const double dni2 = (nb - ni1) * frac;
    double ni2 = ni1 + dni2;
    nb = min_max_bound(drv_c * v2nminrl, drv_c * v2nmaxrl, nc1);
    const double dnc2 = (nb - nc1) * frac;
    double nc2 = nc1 + dnc2;
    if (dni2 == zero && dnc2 != zero)
      ni2 = min_max_bound(drv_i * v2nminrl, drv_i * v2nmaxrl, ni1 - dnc2);
    else if (dni2 != zero && dnc2 == zero)
      nc2 = min_max_bound(drv_c * v2nminrl, drv_c * v2nmaxrl, nc1 - dni2);
    const double total_drv = drv_i + drv_c;
    const double total_num = ni2 + nc2;
    double dni3 = zero, dnc3 = zero;
    const double minb = total_drv * v2nmin, maxb = total_drv * v2nmax;
    if (total_num < minb) {
      const double dt3 = (minb - total_num) * frac;
      if ((ni2 < drv_i * v2nmin) && (nc2 < drv_c * v2nmin)) {
        dni3 = dt3 * (ni2 / total_num);
        dnc3 = dt3 * (nc2 / total_num);
      } else if (nc2 < drv_c * v2nmin) {
        dnc3 = dt3;
      } else if (ni2 < drv_i * v2nmin) {
        dni3 = dt3;
      }
    } else if (total_num > maxb) {
      const double dt3 = (maxb - total_num) * frac;
      if ((ni2 > drv_i * v2nmax) && (nc2 > drv_c * v2nmax)) {
        dni3 = dt3 * (ni2 / total_num);
        dnc3 = dt3 * (nc2 / total_num);
      } else if (nc2 > drv_c * v2nmax) {
        dnc3 = dt3;
      } else if (ni2 > drv_i * v2nmax) {
        dni3 = dt3;
      }
    }
    *num_i = ni2 + dni3;
    *num_c = nc2 + dnc3;
  }
  *dqdt = (*num_i - init_num_i) * dtinv;
  *dqqcwdt = (*num_c - init_num_c) * dtinv;
}

/* calcsize.hpp:480-512 */
static void compute_new_sz_after_transfer(double drv, double num, double v2n_hi,
                                          double v2n_lo, double v2n_nom,
                                          double dgn_hi, double dgn_lo,
                                          double dgn_nom, double sigma,
                                          double *dgncur, double *v2ncur) {
  if (drv > 0.0) {
    if (num <= drv * v2n_hi) {
      *dgncur = dgn_hi;
      *v2ncur = v2n_hi;
    } else if (num >= drv * v2n_lo) {
      *dgncur = dgn_lo;
      *v2ncur = v2n_lo;
    } else {
      *v2ncur = num / drv;
      *dgncur = orc_mean_particle_diameter_from_volume(1.0 / *v2ncur, sigma);
    }
  } else {
    *dgncur = dgn_nom;
    *v2ncur = v2n_nom;
  }
}

/* calcsize.hpp:528-590 (tendencies accumulate: QUIRK, caller zeroes) */
static void update_tends_flx(int jmode, int src, int dst, int ncommon,
                             const int *src_idx, const int *dst_idx,
                             const double xfertend_num[2][2], double xfercoef,
                             const double q_i[4][7], const double q_c[4][7],
                             double dn_i[4], double dn_c[4], double dq_i[4][7],
                             double dq_c[4][7]) {
  dn_i[src] -= xfertend_num[jmode][0];
  dn_i[dst] += xfertend_num[jmode][0];
  dn_c[src] -= xfertend_num[jmode][1];
  dn_c[dst] += xfertend_num[jmode][1];
  for (int i = 0; i < ncommon; ++i) {
    const int ss = src_idx[i], sd = dst_idx[i];
    const double ti = kmax(0.0, q_i[src][ss]) * xfercoef;
    dq_i[src][ss] -= ti;
    dq_i[dst][sd] += ti;
    const double tc = kmax(0.0, q_c[src][ss]) * xfercoef;
    dq_c[src][ss] -= tc;
    dq_c[dst][sd] += tc;
  }
}

/* calcsize.hpp:594-826 */
static void aitken_accum_exchange(const calcsize_consts *c, double adj_tscale_inv,
                                  double dt, const double q_i[4][7],
                                  const double q_c[4][7], double drv_i_ait,
                                  double num_i_ait, double drv_c_ait,
                                  double num_c_ait, double drv_i_acc,
                                  double num_i_acc, double drv_c_acc,
                                  double num_c_acc, double dgn_i[4],
                                  double dgn_c[4], double dn_i[4], double dn_c[4],
                                  double dq_i[4][7], double dq_c[4][7], int e3sm) {
  const double zero = 0, one = 1;
  double v2ncur;
  const double voltonum_ait = c->v2n_nom[NAIT];
  const double voltonum_acc = c->v2n_nom[NACC];
  int ait2acc = 0, acc2ait = 0;
  double xfertend_num[2][2] = {{0, 0}, {0, 0}};
  double xfn_a2c = 0, xfv_a2c = 0, xfn_c2a = 0, xfv_c2a = 0;
  double drv_i_noxf = 0, drv_c_noxf = 0;
  const double geomean = M_sqrt(voltonum_ait) * M_sqrt(voltonum_acc);
  /* compute_coef_ait_acc_transfer (calcsize.hpp:313-378) */
  {
    double fn = zero, fv = zero;
    const double drv_t = drv_i_ait + drv_c_ait;
    const double num_t = num_i_ait + num_c_ait;
    if (drv_t > zero) {
      if (num_t < drv_t * geomean) {
        ait2acc = 1;
        if (num_t < drv_t * voltonum_acc) {
          fn = one;
          fv = one;
        } else {
          fv = ((num_t / drv_t) - geomean) / (voltonum_acc - geomean);
          fn = fv * (drv_t * voltonum_acc / num_t);
          if ((fn <= zero) || (fv <= zero)) {
            fn = zero;
            fv = zero;
          } else if ((fn >= one) || (fv >= one)) {
            fn = one;
            fv = one;
          }
        }
        xfn_a2c = fn * adj_tscale_inv;
        xfv_a2c = fv * adj_tscale_inv;
        xfertend_num[0][0] = num_i_ait * xfn_a2c;
        xfertend_num[0][1] = num_c_ait * xfn_a2c;
      }
    }
  }
  /* compute_coef_acc_ait_transfer (calcsize.hpp:380-478) */
  {
    double drv_t_noxf = zero, num_t0 = zero, num_t_noxf = zero;
    double fn = zero, fv = zero;
    double drv_t = drv_i_acc + drv_c_acc;
    double num_t = num_i_acc + num_c_acc;
    if (drv_t > zero) {
      if (num_t > drv_t * geomean) {
        for (int s = 0; s < NSPEC_MODE[NACC]; ++s)
          if (c->noxf_acc2ait[s]) {
            drv_i_noxf += kmax(zero, q_i[NACC][s]) * c->inv_density[NACC][s];
            drv_c_noxf += kmax(zero, q_c[NACC][s]) * c->inv_density[NACC][s];
          }
        drv_t_noxf = drv_i_noxf + drv_c_noxf;
        /* QUIRK (calcsize.hpp:438): the *min* num2vol ratio. */
        /* the number that cannot move: num2vol_ratio_min (calcsize.hpp:436-438);
         * modal_aero_calcsize_sub uses num2vol_ratio_max
         * (modal_aero_calcsize.hpp:219-222) */
        num_t_noxf = drv_t_noxf * (e3sm ? c->v2n_max[NACC] : c->v2n_min[NACC]);
        num_t0 = num_t;
        num_t = kmax(zero, num_t - num_t_noxf);
        drv_t = kmax(zero, drv_t - drv_t_noxf);
      }
    }
    if (drv_t > zero) {
      if (num_t > drv_t * geomean) {
        acc2ait = 1;
        if (num_t > drv_t * voltonum_ait) {
          fn = one;
          fv = one;
        } else {
          fv = ((num_t / drv_t) - geomean) / (voltonum_ait - geomean);
          fn = fv * (drv_t * voltonum_ait / num_t);
          if ((fn <= zero) || (fv <= zero)) {
            fn = zero;
            fv = zero;
          } else if ((fn >= one) || (fv >= one)) {
            fn = one;
            fv = one;
          }
        }
        fn = fn * num_t * safe_denominator(num_t0);
        xfn_c2a = fn * adj_tscale_inv;
        xfv_c2a = fv * adj_tscale_inv;
        xfertend_num[1][0] = num_i_acc * xfn_c2a;
        xfertend_num[1][1] = num_c_acc * xfn_c2a;
      }
    }
  }
  if (ait2acc + acc2ait > 0) {
    const double ndi = (xfertend_num[0][0] - xfertend_num[1][0]) * dt;
    const double num_i = kmax(zero, num_i_ait - ndi);
    const double num_i_a = kmax(zero, num_i_acc + ndi);
    const double vdi = (drv_i_ait * xfv_a2c - (drv_i_acc - drv_i_noxf) * xfv_c2a) * dt;
    const double drv_i = kmax(zero, drv_i_ait - vdi);
    const double drv_i_a = kmax(zero, drv_i_acc + vdi);
    const double ndc = (xfertend_num[0][1] - xfertend_num[1][1]) * dt;
    const double num_c = kmax(zero, num_c_ait - ndc);
    const double num_c_a = kmax(zero, num_c_acc + ndc);
    const double vdc = (drv_c_ait * xfv_a2c - (drv_c_acc - drv_c_noxf) * xfv_c2a) * dt;
    const double drv_c = kmax(zero, drv_c_ait - vdc);
    const double drv_c_a = kmax(zero, drv_c_acc + vdc);
    compute_new_sz_after_transfer(drv_i, num_i, c->v2n_min[NAIT], c->v2n_max[NAIT],
                                  c->v2n_nom[NAIT], c->dgnmax[NAIT], c->dgnmin[NAIT],
                                  c->dgnnom[NAIT], c->sigma[NAIT], &dgn_i[NAIT], &v2ncur);
    compute_new_sz_after_transfer(drv_c, num_c, c->v2n_min[NAIT], c->v2n_max[NAIT],
                                  c->v2n_nom[NAIT], c->dgnmax[NAIT], c->dgnmin[NAIT],
                                  c->dgnnom[NAIT], c->sigma[NAIT], &dgn_c[NAIT], &v2ncur);
    compute_new_sz_after_transfer(drv_i_a, num_i_a, c->v2n_min[NACC], c->v2n_max[NACC],
                                  c->v2n_nom[NACC], c->dgnmax[NACC], c->dgnmin[NACC],
                                  c->dgnnom[NACC], c->sigma[NACC], &dgn_i[NACC], &v2ncur);
    compute_new_sz_after_transfer(drv_c_a, num_c_a, c->v2n_min[NACC], c->v2n_max[NACC],
                                  c->v2n_nom[NACC], c->dgnmax[NACC], c->dgnmin[NACC],
                                  c->dgnnom[NACC], c->sigma[NACC], &dgn_c[NACC], &v2ncur);
    if (ait2acc > 0)
      update_tends_flx(0, NAIT, NACC, c->n_common, c->ait_spec_in_acc,
                       c->acc_spec_in_ait, (const double(*)[2])xfertend_num, xfv_a2c,
                       q_i, q_c, dn_i, dn_c, dq_i, dq_c);
    if (acc2ait > 0)
      update_tends_flx(1, NACC, NAIT, c->n_common, c->acc_spec_in_ait,
                       c->ait_spec_in_acc, (const double(*)[2])xfertend_num, xfv_c2a,
                       q_i, q_c, dn_i, dn_c, dq_i, dq_c);
  }
}

static calcsize_consts g_cs;
static int g_cs_ready = 0;

static void ensure_calcsize_consts(void) {
  if (__atomic_load_n(&g_cs_ready, __ATOMIC_ACQUIRE)) return;
#ifdef _OPENMP
#pragma omp critical(orc_cs_init)
#endif
  {
    if (!g_cs_ready) {
      calcsize_init(&g_cs);
      __atomic_store_n(&g_cs_ready, 1, __ATOMIC_RELEASE);
    }
  }
}

/* CalcSize::compute_tendencies body for one level (calcsize.hpp:999-1216) */
static void calcsize_level(const orc_config *cfg, double dt, const double n_i[4],
                           const double n_c[4], const double q_i[4][7],
                           const double q_c[4][7], double dn_i[4], double dn_c[4],
                           double dq_i[4][7], double dq_c[4][7], double dgn_i[4],
                           double dgn_c[4], int e3sm);

void orc_calcsize_level(const orc_config *cfg, double dt, const double n_i[4],
                        const double n_c[4], const double q_i[4][7],
                        const double q_c[4][7], double dn_i[4], double dn_c[4],
                        double dq_i[4][7], double dq_c[4][7], double dgn_i[4],
                        double dgn_c[4]) {
  calcsize_level(cfg, dt, n_i, n_c, q_i, q_c, dn_i, dn_c, dq_i, dq_c, dgn_i, dgn_c, 0);
}

/* modal_aero_calcsize_sub (modal_aero_calcsize.hpp:748-928) for one level:
 * the E3SM-layout calcsize.  Same algorithm as CalcSize::compute_tendencies
 * on the same quantities (its state_q / qqcw slots are the AeroProcess
 * fields in mode_aero_species order, ndrop.hpp:55-90; its ptend / dqqcwdt
 * are dn_i/dq_i and dn_c/dq_c), with two differences:
 *   - size_adjustment divides the Aitken mode's num2vol_ratio_min and
 *     multiplies the accumulation mode's num2vol_ratio_max by
 *     szadj_block_fac = 1e6 before updating the diameters (:373-382);
 *   - the accumulation -> Aitken transfer counts the non-transferable
 *     species' number with num2vol_ratio_max (:219-222).
 * do_adjust and do_aitacc_transfer are always on (CalcsizeData, :138-139). */
void orc_calcsize_sub_level(double dt, const double n_i[4], const double n_c[4],
                            const double q_i[4][7], const double q_c[4][7], double dn_i[4],
                            double dn_c[4], double dq_i[4][7], double dq_c[4][7],
                            double dgn_i[4], double dgn_c[4]) {
  orc_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.do_calcsize = 1;
  cfg.calcsize_do_aitacc_transfer = 1;
  cfg.calcsize_do_adjust = 1;
  calcsize_level(&cfg, dt, n_i, n_c, q_i, q_c, dn_i, dn_c, dq_i, dq_c, dgn_i, dgn_c, 1);
}

static void calcsize_level(const orc_config *cfg, double dt, const double n_i[4],
                           const double n_c[4], const double q_i[4][7],
                           const double q_c[4][7], double dn_i[4], double dn_c[4],
                           double dq_i[4][7], double dq_c[4][7], double dgn_i[4],
                           double dgn_c[4], int e3sm) {
  ensure_calcsize_consts();
  const calcsize_consts *c = &g_cs;
  const int do_xfer = cfg->calcsize_do_aitacc_transfer;
  const int do_adjust = cfg->calcsize_do_adjust;
  double dvi_ait = 0, ni_ait = 0, dvc_ait = 0, nc_ait = 0;
  double dvi_acc = 0, ni_acc = 0, dvc_acc = 0, nc_acc = 0;
  const double adj_tscale = kmax(86400.0, dt);
  const double adj_tscale_inv = safe_denominator(adj_tscale);
  for (int m = 0; m < 4; ++m) {
    dgn_i[m] = c->dgnnom[m];
    double v2n_i = c->v2n_nom[m];
    dgn_c[m] = c->dgnnom[m];
    double v2n_c = c->v2n_nom[m];
    double dv_i = 0, dv_c = 0;
    for (int s = 0; s < NSPEC_MODE[m]; ++s) {
      dv_i += kmax(0.0, q_i[m][s]) * c->inv_density[m][s];
      dv_c += kmax(0.0, q_c[m][s]) * c->inv_density[m][s];
    }
    const double init_num_i = n_i[m];
    double num_i = init_num_i < 0 ? 0.0 : init_num_i;
    const double init_num_c = n_c[m];
    double num_c = init_num_c < 0 ? 0.0 : init_num_c;
    const int is_aa = (m == NACC) || (m == NAIT);
    const int skip_adjust = is_aa && do_xfer;
    if (do_adjust && !skip_adjust)
      adjust_num_sizes(dv_i, dv_c, init_num_i, init_num_c, dt, c->v2n_min[m],
                       c->v2n_max[m], adj_tscale_inv, &num_i, &num_c, &dn_i[m],
                       &dn_c[m]);
    double v2nmin = c->v2n_min[m], v2nmax = c->v2n_max[m];
    if (e3sm) { /* szadj_block_fac (modal_aero_calcsize.hpp:373-382) */
      if (m == NAIT) v2nmin /= 1.0e6;
      if (m == NACC) v2nmax *= 1.0e6;
    }
    update_diameter_and_vol2num(dv_i, num_i, v2nmin, v2nmax, c->dgnmin[m], c->dgnmax[m],
                                c->sigma[m], &dgn_i[m], &v2n_i);
    update_diameter_and_vol2num(dv_c, num_c, v2nmin, v2nmax, c->dgnmin[m], c->dgnmax[m],
                                c->sigma[m], &dgn_c[m], &v2n_c);
    if (do_xfer) {
      if (m == NAIT) {
        dvi_ait = dv_i; ni_ait = num_i; dvc_ait = dv_c; nc_ait = num_c;
      } else if (m == NACC) {
        dvi_acc = dv_i; ni_acc = num_i; dvc_acc = dv_c; nc_acc = num_c;
      }
    }
  }
  if (do_xfer)
    aitken_accum_exchange(c, adj_tscale_inv, dt, q_i, q_c, dvi_ait, ni_ait,
                          dvc_ait, nc_ait, dvi_acc, ni_acc, dvc_acc, nc_acc,
                          dgn_i, dgn_c, dn_i, dn_c, dq_i, dq_c, e3sm);
}

/* ======================= composed step ==================================== */

/* State field order (include/mam4x.h): n_i[4], n_c[4], q_i[21], q_c[21],
 * h2so4, soag.  Species within a mode follow mode_aero_species order. */
static const int MODE_OFF[4] = {0, 7, 11, 18};
#define F_NI(m) (m)
#define F_NC(m) (4 + (m))
#define F_QI(m, s) (8 + MODE_OFF[m] + (s))
#define F_QC(m, s) (29 + MODE_OFF[m] + (s))
#define F_H2SO4 50
#define F_SOAG 51

/* gas_pcnst slot and kg/kg -> mol/mol factor of each state field */
static void field_maps(int pcnst[ORC_NSTATE], double to_vmr[ORC_NSTATE],
                       double to_mmr[ORC_NSTATE]) {
  const double mwair = 1000.0 * MW_DRY_AIR; /* 28.966 g/mol */
  for (int m = 0; m < 4; ++m) {
    pcnst[F_NI(m)] = NUMPTR[m];
    pcnst[F_NC(m)] = NUMPTR[m];
    for (int s = 0; s < NSPEC_MODE[m]; ++s) {
      pcnst[F_QI(m, s)] = LMASSPTR[s][m];
      pcnst[F_QC(m, s)] = LMASSPTR[s][m];
    }
  }
  pcnst[F_H2SO4] = 2;
  pcnst[F_SOAG] = 5;
  for (int f = 0; f < ORC_NSTATE; ++f) {
    if (f < 8) { /* numbers: #/kg -> #/kmol */
      to_vmr[f] = mwair;
      to_mmr[f] = 1.0 / mwair;
    } else {
      to_vmr[f] = mwair / ADV_MASS[pcnst[f]];
      to_mmr[f] = ADV_MASS[pcnst[f]] / mwair;
    }
  }
}

long orc_chain_step(const orc_config *cfg, int ncol, int nlev, double dt,
                    const double *state, const double *temperature,
                    const double *pressure, const double *height,
                    const double *qv, const double *cloud_fraction,
                    const double *pblh, const double *dgncur_a,
                    const double *dgncur_awet, const double *wetdens,
                    const double *pregaschem_gas, const double *precldchem,
                    double *tend, double *dgn_i, double *dgn_c, int nthreads) {
  return orc_chain_step_proc(cfg, ncol, nlev, dt, state, temperature, pressure,
                             height, qv, cloud_fraction, pblh, dgncur_a,
                             dgncur_awet, wetdens, pregaschem_gas, precldchem,
                             tend, dgn_i, dgn_c, NULL, nthreads);
}

long orc_chain_step_proc(const orc_config *cfg, int ncol, int nlev, double dt,
                         const double *state, const double *temperature,
                         const double *pressure, const double *height,
                         const double *qv, const double *cloud_fraction,
                         const double *pblh, const double *dgncur_a,
                         const double *dgncur_awet, const double *wetdens,
                         const double *pregaschem_gas, const double *precldchem,
                         double *tend, double *dgn_i, double *dgn_c,
                         double *proc, int nthreads) {
  int pcnst[ORC_NSTATE];
  double to_vmr[ORC_NSTATE], to_mmr[ORC_NSTATE];
  field_maps(pcnst, to_vmr, to_mmr);
  const long ncell = (long)ncol * nlev;
  long total_iter = 0;
  ensure_calcsize_consts();
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(+ : total_iter)
#endif
  for (int icol = 0; icol < ncol; ++icol) {
    for (int k = 0; k < nlev; ++k) {
      const long c = (long)k * ncol + icol;
      if (g_noise_seed) noise_cell(c);
#define AT(ptr, f) (ptr)[(long)(f) * ncell + c]
      double qq[31], qqcw[31], q1[31], q2[31], qcw2[31];
      for (int i = 0; i < 31; ++i) qq[i] = qqcw[i] = q1[i] = q2[i] = qcw2[i] = 0.0;
      for (int f = 0; f < ORC_NSTATE; ++f) {
        const int isc = (f >= 4 && f < 8) || (f >= 29 && f < 50);
        const double v = AT(state, f) * to_vmr[f];
        if (isc)
          qqcw[pcnst[f]] = v;
        else
          qq[pcnst[f]] = v;
      }
      if (precldchem) {
        for (int f = 0; f < ORC_NSTATE; ++f) {
          const int isc = (f >= 4 && f < 8) || (f >= 29 && f < 50);
          const double v = AT(precldchem, f) * to_vmr[f];
          if (isc)
            qcw2[pcnst[f]] = v;
          else
            q2[pcnst[f]] = v;
        }
      } else {
        memcpy(q2, qq, sizeof(q2));
        memcpy(qcw2, qqcw, sizeof(qcw2));
      }
      memcpy(q1, q2, sizeof(q1));
      if (pregaschem_gas) {
        q1[2] = AT(pregaschem_gas, 0) * to_vmr[F_H2SO4];
        q1[5] = AT(pregaschem_gas, 1) * to_vmr[F_SOAG];
      } else {
        q1[2] = qq[2];
        q1[5] = qq[5];
      }
      double dga[4], dgw[4], wd[4];
      for (int m = 0; m < 4; ++m) {
        dga[m] = AT(dgncur_a, m);
        dgw[m] = AT(dgncur_awet, m);
        wd[m] = AT(wetdens, m);
      }
      int stats[2];
      double qgcm_tendaa[31 * 4], qqcwgcm_tendaa[31];
      orc_amicphys_intr(cfg, dt, temperature[c], pressure[c], 0.0, height[c],
                        pblh[icol], qv[c], cloud_fraction[c], qq, qqcw, q1, q2,
                        qcw2, dga, dgw, wd, proc ? qgcm_tendaa : NULL,
                        proc ? qqcwgcm_tendaa : NULL, stats);
      total_iter += stats[0];
      if (proc) /* the five diagnostic blocks (mam4_amicphys.hpp:2365-2374) */
        for (int i = 0; i < 31; ++i) {
          for (int iq = 0; iq < 4; ++iq)
            proc[((long)iq * 31 + i) * ncell + c] = qgcm_tendaa[i * 4 + iq];
          proc[((long)4 * 31 + i) * ncell + c] = qqcwgcm_tendaa[i];
        }
      /* back to prognostic units; amicphys tendency */
      double post[ORC_NSTATE];
      for (int f = 0; f < ORC_NSTATE; ++f) {
        const int isc = (f >= 4 && f < 8) || (f >= 29 && f < 50);
        post[f] = (isc ? qqcw[pcnst[f]] : qq[pcnst[f]]) * to_mmr[f];
        AT(tend, f) = (post[f] - AT(state, f)) / dt;
      }
      double di[4] = {0, 0, 0, 0}, dc[4] = {0, 0, 0, 0};
      if (cfg->do_calcsize) {
        double n_i[4], n_c[4], q_i[4][7], q_c[4][7];
        double dn_i[4] = {0, 0, 0, 0}, dn_c[4] = {0, 0, 0, 0};
        double dq_i[4][7], dq_c[4][7];
        memset(q_i, 0, sizeof(q_i));
        memset(q_c, 0, sizeof(q_c));
        memset(dq_i, 0, sizeof(dq_i));
        memset(dq_c, 0, sizeof(dq_c));
        for (int m = 0; m < 4; ++m) {
          n_i[m] = post[F_NI(m)];
          n_c[m] = post[F_NC(m)];
          for (int s = 0; s < NSPEC_MODE[m]; ++s) {
            q_i[m][s] = post[F_QI(m, s)];
            q_c[m][s] = post[F_QC(m, s)];
          }
        }
        orc_calcsize_level(cfg, dt, n_i, n_c, (const double(*)[7])q_i,
                           (const double(*)[7])q_c, dn_i, dn_c, dq_i, dq_c, di, dc);
        for (int m = 0; m < 4; ++m) {
          AT(tend, F_NI(m)) += dn_i[m];
          AT(tend, F_NC(m)) += dn_c[m];
          for (int s = 0; s < NSPEC_MODE[m]; ++s) {
            AT(tend, F_QI(m, s)) += dq_i[m][s];
            AT(tend, F_QC(m, s)) += dq_c[m][s];
          }
        }
      } else {
        for (int m = 0; m < 4; ++m) {
          di[m] = dga[m];
          dc[m] = dga[m];
        }
      }
      for (int m = 0; m < 4; ++m) {
        AT(dgn_i, m) = di[m];
        AT(dgn_c, m) = dc[m];
      }
#undef AT
    }
  }
  return total_iter;
}

/* ======================= water uptake ===================================== */
/* modal_aero_water_uptake_dr (water_uptake.hpp:481-505) and the functions it
 * calls, per cell.  The reference's complex arithmetic is Kokkos::complex
 * (Kokkos, a dependency not vendored in the reference tree); the helpers
 * below restate its published algorithms:
 *   z1 * z2   (a c - b d, a d + b c)
 *   z1 / z2   both scaled by s = |re(z2)| + |im(z2)|, then
 *             (z1/s) * conj(z2/s) / |z2/s|^2
 *   sqrt(z)   re == 0: t = sqrt(|im|/2), (t, sign(im) t); else
 *             t = sqrt(2 (|z| + |re|)), u = t/2:
 *             re > 0 ? (u, im/t) : (|im|/t, sign(im) u)
 *   pow(z, y) polar(|z|^y, y atan2(im, re)),  |z| = hypot(re, im)
 * The restatement is pinned by the quartic residual and the reference's own
 * Mathematica roots of the Kohler polynomial (tests/test_water_uptake.py). */
/* Kokkos::max/min and utils::min_max_bound (mam4_math.hpp:33-34,
 * utils.hpp:24-26) */
static double wu_max(double x, double y) { return x < y ? y : x; }
static double wu_min(double x, double y) { return y < x ? y : x; }
static double wu_mmb(double lo, double hi, double v) { return wu_max(lo, wu_min(hi, v)); }

typedef struct {
  double re, im;
} cplx;
static cplx cx(double re, double im) {
  cplx z = {re, im};
  return z;
}
static cplx c_add(cplx a, cplx b) { return cx(a.re + b.re, a.im + b.im); }
static cplx c_sub(cplx a, cplx b) { return cx(a.re - b.re, a.im - b.im); }
static cplx c_mul(cplx a, cplx b) {
  return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static cplx c_divr(cplx a, double r) { return cx(a.re / r, a.im / r); }
static cplx c_div(cplx x, cplx y) {
  const double s = fabs(y.re) + fabs(y.im);
  const cplx xs = cx(x.re / s, x.im / s);
  const cplx ys = cx(y.re / s, -y.im / s);
  const double a = ys.re * ys.re + ys.im * ys.im;
  return c_divr(c_mul(xs, ys), a);
}
static double c_abs(cplx z) { return hypot(z.re, z.im); }
static cplx c_sqrt(cplx z) {
  const double r = z.re, i = z.im;
  if (r == 0.0) {
    const double t = M_sqrt(fabs(i) / 2);
    return cx(t, i < 0.0 ? -t : t);
  }
  const double t = M_sqrt(2 * (c_abs(z) + fabs(r)));
  const double u = t / 2;
  return r > 0.0 ? cx(u, i / t) : cx(fabs(i) / t, i < 0.0 ? -u : u);
}
static cplx c_pow(cplx z, double y) {
  const double rho = M_pow(c_abs(z), y), th = y * atan2(z.im, z.re);
  return cx(rho * cos(th), rho * sin(th));
}

/* makoh_quartic (water_uptake.hpp:124-169): roots of
 * x^4 + p3 x^3 + p2 x^2 + p1 x + p0 */
void orc_makoh_quartic(double p3, double p2, double p1, double p0, double re[4],
                       double im[4]) {
  cplx r[4];
  const double qq = -p2 * p2 / 36.0 + (p3 * p1 - 4.0 * p0) / 12.0;
  const double rr = -cube(p2 / 6.0) + p2 * (p3 * p1 - 4.0 * p0) / 48.0 +
                    (4.0 * p0 * p2 - p0 * p3 * p3 - p1 * p1) / 16.0;
  cplx crad = c_sqrt(cx(rr * rr + qq * qq * qq, 0.0));
  cplx cb = c_sub(cx(rr, 0.0), crad);
  if (cb.re == 0.0 && cb.im == 0.0) { /* insoluble particle */
    const double c = M_cbrt(-p1);
    for (int n = 0; n < 4; ++n) r[n] = cx(c, 0.0);
  } else {
    cb = c_pow(cb, 1.0 / 3.0);
    const cplx cy = c_add(c_add(cx(-cb.re, -cb.im), c_div(cx(qq, 0.0), cb)), cx(p2 / 6.0, 0.0));
    const cplx cb0 = c_sqrt(c_sub(c_mul(cy, cy), cx(p0, 0.0)));
    const cplx cb1 = c_div(c_sub(c_mul(cx(p3, 0.0), cy), cx(p1, 0.0)), c_mul(cx(2.0, 0.0), cb0));
    cb = c_add(cx(p3 / 2.0, 0.0), cb1);
    crad = c_sqrt(c_sub(c_mul(cb, cb), c_mul(cx(4.0, 0.0), c_add(cy, cb0))));
    r[0] = c_divr(c_add(cx(-cb.re, -cb.im), crad), 2.0);
    r[1] = c_divr(c_sub(cx(-cb.re, -cb.im), crad), 2.0);
    cb = c_sub(cx(p3 / 2.0, 0.0), cb1);
    crad = c_sqrt(c_sub(c_mul(cb, cb), c_mul(cx(4.0, 0.0), c_sub(cy, cb0))));
    r[2] = c_divr(c_add(cx(-cb.re, -cb.im), crad), 2.0);
    r[3] = c_divr(c_sub(cx(-cb.re, -cb.im), crad), 2.0);
  }
  for (int n = 0; n < 4; ++n) {
    re[n] = r[n].re;
    im[n] = r[n].im;
  }
}

#define WU_EPS 1e-4 /* Water_Uptake::eps (water_uptake.hpp:38) */

/* find_real_solution (water_uptake.hpp:93-117): the smallest real root in
 * [rdry (1 - eps), 1000 rdry] */
static double find_real_solution(double rdry, const double re[4], const double im[4]) {
  double rwet = 1000.0 * rdry;
  for (int n = 0; n < 4; ++n) {
    const double xr = re[n], xi = im[n];
    if (fabs(xi) > fabs(xr) * WU_EPS) continue;
    if (xr > rwet) continue;
    if (xr < rdry * (1.0 - WU_EPS)) continue;
    if (isnan(xr)) continue;
    rwet = xr;
  }
  return rwet;
}

/* Kohler quartic solve for a given Kelvin coefficient aa [micron]:
 * the body of modal_aero_kohler (water_uptake.hpp:178-229) from the
 * coefficients on; rdry [micron], returns rwet [micron] before the bound. */
double orc_kohler_root(double aa, double rdry, double hygro, double rh) {
  const double vol = cube(rdry);
  const double bb = vol * hygro;
  const double ss = wu_mmb(1.e-10, 1.0 - WU_EPS, rh);
  const double slog = M_log(ss);
  const double p43 = -aa / slog;
  const double p42 = 0.0;
  const double p41 = bb / slog - vol;
  const double p40 = aa * vol / slog;
  const double pp = fabs(-bb / aa) / (rdry * rdry);
  if (pp < WU_EPS) /* approximate solution for small particles */
    return rdry * (1.0 + pp * (1.0 / 3.0) / (1.0 - slog * rdry / aa));
  double re[4], im[4];
  orc_makoh_quartic(p43, p42, p41, p40, re, im);
  return find_real_solution(rdry, re, im);
}

/* modal_aero_kohler (water_uptake.hpp:178-229): rdry_in [m] -> rwet [m] */
static double modal_aero_kohler(double rdry_in, double hygro, double rh) {
  const double rhow = 1.0, surften = 76.0, mw = 18.0, tair = 273.0, ugascon = 8.3e7;
  const double aa = 2.0e4 * mw * surften / (ugascon * tair * rhow);
  double rwet = orc_kohler_root(aa, rdry_in * 1.e6, hygro, rh);
  rwet = wu_min(rwet, 30.0); /* rmax: upper bound based on 1 day lifetime */
  return rwet * 1.e-6;
}

/* get_e3sm_parameters (water_uptake.hpp:312-360): E3SM species type (1-based)
 * of the ispec-th species of each mode, and density / hygroscopicity by type.
 * The ispec-th E3SM species of a mode is the ispec-th species of the mode in
 * mode_aero_species order (state_q layout, utils.hpp:224-258). */
static const int WU_LSPECTYPE[4][7] = {{1, 4, 5, 6, 8, 7, 9},
                                       {1, 5, 7, 9, 0, 0, 0},
                                       {8, 7, 1, 6, 4, 5, 9},
                                       {4, 6, 9, 0, 0, 0, 0}};
/* types 2 and 3 are unused by the four modes; their entries are written
 * "0.1797693135 + 309" in the reference */
static const double WU_SPECDENS[9] = {1770.0, 0.1797693135 + 309, 0.1797693135 + 309,
                                      1000.0, 1000.0, 1700.0, 1900.0, 2600.0, 1601.0};
static const double WU_SPECHYGRO[9] = {0.507, 0.1797693135 + 309, 0.1797693135 + 309,
                                       0.1000000083e-09, 0.14, 0.1000000013e-09,
                                       1.16, 0.068, 0.1000000015};
#define WU_RHCRYSTAL 0.35 /* aero_modes.hpp:76-77 */
#define WU_RHDELIQUES 0.8

/* wv_sat_qsat_water (wv_sat_methods.hpp:13-133): Goff & Gratch SVP over
 * water, then saturation specific humidity */
static void qsat_water(double t, double p, double *es, double *qs) {
  const double tboil = 373.16, ten = 10, one = 1;
  const double e = M_pow(ten, -7.90298 * (tboil / t - one) + 5.02808 * log10(tboil / t) -
                                 1.3816e-7 * (M_pow(ten, 11.344 * (one - t / tboil)) - one) +
                                 8.1328e-3 * (M_pow(ten, -3.49149 * (tboil / t - one)) - one) +
                                 log10(1013.246)) *
                   ten * ten;
  const double epsilo = MW_H2O / MW_DRY_AIR;
  const double omeps = one - epsilo;
  *qs = (p - e) <= 0.0 ? one : epsilo * e / (p - omeps * e);
  *es = wu_min(e, p);
}

void orc_qsat_water(double t, double p, double *es, double *qs) { qsat_water(t, p, es, qs); }

/* One cell of modal_aero_water_uptake_dr with wet density
 * (water_uptake.hpp:362-505).  qaer[m][s]: interstitial mass mixing ratios
 * [kg/kg] in mode_aero_species order; qnum is not used by the reference (the
 * number comes from the dry volume and dgncur_a).  qv [kg/kg]. */
void orc_water_uptake_cell(const double qaer[4][7], double temperature, double pmid,
                           double qv, double cldn, const double dgncur_a[4],
                           double dgncur_awet[4], double qaerwat[4], double wetdens[4]) {
  double hygro[4], naer[4], dryrad[4], dryvol[4], drymass[4], specdens_1[4];
  /* modal_aero_water_uptake_dryaer (:362-437) */
  for (int m = 0; m < 4; ++m) {
    double dryvolmr = 0.0, maer = 0.0;
    hygro[m] = 0.0;
    const double alnsg = M_log(MODE_SIGMA[m]);
    specdens_1[m] = WU_SPECDENS[WU_LSPECTYPE[m][0] - 1];
    const double spechygro_1 = WU_SPECHYGRO[WU_LSPECTYPE[m][0] - 1];
    for (int s = 0; s < NSPEC_MODE[m]; ++s) {
      const int t = WU_LSPECTYPE[m][s] - 1;
      const double raer = qaer[m][s];
      const double vol_tmp = raer / WU_SPECDENS[t];
      maer = maer + raer;
      dryvolmr += vol_tmp;
      hygro[m] += vol_tmp * WU_SPECHYGRO[t];
    }
    hygro[m] = dryvolmr > 1e-30 ? hygro[m] / dryvolmr : spechygro_1;
    const double v2ncur_a = 1.0 / ((PI / 6.0) * cube(dgncur_a[m]) * M_exp(4.5 * sq(alnsg)));
    naer[m] = dryvolmr * v2ncur_a;
    const double drydens = maer > 1e-31 ? maer / dryvolmr : 1.0;
    dryvol[m] = 1.0 / v2ncur_a;
    drymass[m] = drydens * dryvol[m];
    dryrad[m] = M_cbrt(dryvol[m] / (PI * 4.0 / 3.0));
  }
  /* modal_aero_water_uptake_rh_clearair (:259-285) */
  double es, qs, rh;
  qsat_water(temperature, pmid, &es, &qs);
  rh = qs > qv ? qv / qs : 0.98;
  rh = wu_mmb(0.0, 0.98, rh);
  if (cldn < 1.0) rh = (rh - cldn) / (1.0 - cldn);
  rh = wu_max(rh, 0.0);
  /* modal_aero_water_uptake_wetaer (:241-257) */
  double wetvol[4], wtrvol[4];
  for (int m = 0; m < 4; ++m) {
    const double hystfac = 1.0 / wu_max(1.0e-5, WU_RHDELIQUES - WU_RHCRYSTAL);
    double wetrad = modal_aero_kohler(dryrad[m], hygro[m], rh);
    wetrad = wu_max(wetrad, dryrad[m]);
    wetvol[m] = (PI * 4.0 / 3.0) * cube(wetrad);
    wetvol[m] = wu_max(wetvol[m], dryvol[m]);
    wtrvol[m] = wetvol[m] - dryvol[m];
    wtrvol[m] = wu_max(wtrvol[m], 0.0);
    if (rh < WU_RHCRYSTAL) {
      wetrad = dryrad[m];
      wetvol[m] = dryvol[m];
      wtrvol[m] = 0.0;
    } else if (rh < WU_RHDELIQUES) {
      wtrvol[m] = wtrvol[m] * hystfac * (rh - WU_RHCRYSTAL);
      wtrvol[m] = wu_max(wtrvol[m], 0.0);
      wetvol[m] = dryvol[m] + wtrvol[m];
      wetrad = M_cbrt(wetvol[m] / (4.0 / 3.0 * PI));
    }
    dgncur_awet[m] = dgncur_a[m] * (wetrad / dryrad[m]);
    qaerwat[m] = 1.0e3 * naer[m] * wtrvol[m];
  }
  /* modal_aero_water_uptake_wetdens (:67-90) */
  for (int m = 0; m < 4; ++m)
    wetdens[m] = wetvol[m] > 1.0e-30 ? (drymass[m] + 1.0e3 * wtrvol[m]) / wetvol[m]
                                     : specdens_1[m];
}

/* Water uptake over [field][level][column] buffers (column fastest): state
 * as for orc_chain_step (only the interstitial mass fields are read),
 * dgncur_a [4][nlev][ncol]; outputs [4][nlev][ncol] each (wetdens may be
 * NULL). */
void orc_water_uptake(int ncol, int nlev, const double *state, const double *temperature,
                      const double *pressure, const double *qv, const double *cloud_fraction,
                      const double *dgncur_a, double *dgncur_awet, double *qaerwat,
                      double *wetdens) {
  const long ncell = (long)ncol * nlev;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (long c = 0; c < ncell; ++c) {
    double qaer[4][7] = {{0}}, dga[4], dgw[4], qw[4], wd[4];
    for (int m = 0; m < 4; ++m) {
      dga[m] = dgncur_a[m * ncell + c];
      for (int s = 0; s < NSPEC_MODE[m]; ++s) qaer[m][s] = state[(long)F_QI(m, s) * ncell + c];
    }
    orc_water_uptake_cell((const double(*)[7])qaer, temperature[c], pressure[c], qv[c],
                          cloud_fraction[c], dga, dgw, qw, wd);
    for (int m = 0; m < 4; ++m) {
      dgncur_awet[m * ncell + c] = dgw[m];
      qaerwat[m * ncell + c] = qw[m];
      if (wetdens) wetdens[m * ncell + c] = wd[m];
    }
  }
}

/* compute_calcsize_and_water_uptake_dr (modal_aero_opt.hpp:510-533) over
 * [field][level][column] buffers: modal_aero_calcsize_sub on the state, then
 * water uptake on its interstitial dry diameters.  Outputs [4][nlev][ncol]
 * (dgn_c, wetdens may be NULL) and the calcsize tendencies [ORC_NSTATE]
 * (may be NULL; gas fields 0). */
void orc_wet_size(int ncol, int nlev, double dt, const double *state,
                  const double *temperature, const double *pressure, const double *qv,
                  const double *cloud_fraction, double *dgn_i, double *dgn_c,
                  double *dgncur_awet, double *qaerwat, double *wetdens, double *tend) {
  const long ncell = (long)ncol * nlev;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (long c = 0; c < ncell; ++c) {
    double n_i[4], n_c[4], q_i[4][7] = {{0}}, q_c[4][7] = {{0}};
    double dn_i[4] = {0}, dn_c[4] = {0}, dq_i[4][7] = {{0}}, dq_c[4][7] = {{0}};
    double di[4], dc[4], dgw[4], qw[4], wd[4];
    for (int m = 0; m < 4; ++m) {
      n_i[m] = state[(long)F_NI(m) * ncell + c];
      n_c[m] = state[(long)F_NC(m) * ncell + c];
      for (int s = 0; s < NSPEC_MODE[m]; ++s) {
        q_i[m][s] = state[(long)F_QI(m, s) * ncell + c];
        q_c[m][s] = state[(long)F_QC(m, s) * ncell + c];
      }
    }
    orc_calcsize_sub_level(dt, n_i, n_c, (const double(*)[7])q_i, (const double(*)[7])q_c, dn_i,
                           dn_c, dq_i, dq_c, di, dc);
    orc_water_uptake_cell((const double(*)[7])q_i, temperature[c], pressure[c], qv[c],
                          cloud_fraction[c], di, dgw, qw, wd);
    for (int m = 0; m < 4; ++m) {
      dgn_i[m * ncell + c] = di[m];
      if (dgn_c) dgn_c[m * ncell + c] = dc[m];
      dgncur_awet[m * ncell + c] = dgw[m];
      qaerwat[m * ncell + c] = qw[m];
      if (wetdens) wetdens[m * ncell + c] = wd[m];
    }
    if (tend) {
      for (int f = 0; f < ORC_NSTATE; ++f) tend[(long)f * ncell + c] = 0.0;
      for (int m = 0; m < 4; ++m) {
        tend[(long)F_NI(m) * ncell + c] = dn_i[m];
        tend[(long)F_NC(m) * ncell + c] = dn_c[m];
        for (int s = 0; s < NSPEC_MODE[m]; ++s) {
          tend[(long)F_QI(m, s) * ncell + c] = dq_i[m][s];
          tend[(long)F_QC(m, s) * ncell + c] = dq_c[m][s];
        }
      }
    }
  }
}

/* ======================= aqueous sulfur chemistry ======================== */
/* mo_setsox.hpp: setsox_single_level (:943-1259) and the routines it calls,
 * per cell, on the gas_pcnst tracer arrays qin (gases, mol/mol) and qcw
 * (cloud-borne aerosol, mol/mol and #/kmol) at the reference's indices
 * (lptr_so4_cw_amode - loffset = 6, 14, 21; numptrcw_amode - loffset = 13,
 * 18, 26, 30; O3 0, H2O2 1, H2SO4 2, SO2 3). */
#define SOX_RA (8314.0 / 101325.0)
#define SOX_XKW 1.0e-14
#define SOX_CONST0 (1.0e3 / 6.023e23)
#define SOX_CO2G 330.0e-6
#define SOX_P0 101300.0
static const int SOX_SO4_CW[3] = {6, 14, 21};
static const int SOX_NUM_CW[4] = {13, 18, 26, 30};

/* sox_cldaero_create_obj (:88-131) */
void orc_sox_create_obj(double cldfrc, const double qcw[31], double lwc, double cfact,
                        double out[3]) {
  double xlwc = 0.0;
  if (cldfrc > 0.0) {
    xlwc = lwc * cfact;
    xlwc = xlwc / cldfrc;
  }
  out[0] = qcw[SOX_SO4_CW[0]] + qcw[SOX_SO4_CW[1]] + qcw[SOX_SO4_CW[2]];
  out[1] = xlwc;
  out[2] = 1.0; /* so4_fact */
}

/* henry_factor_{so2,co2,h2o2,o3} (:133-186): species 0 SO2 (xk, xe, x2),
 * 1 CO2, 2 H2O2 (xk, xe), 3 O3 (xk) */
void orc_sox_henry(int species, double t_factor, double out[3]) {
  out[0] = out[1] = out[2] = 0.0;
  switch (species) {
    case 0:
      out[0] = 1.230 * M_exp(3120.0 * t_factor);
      out[1] = 1.7e-2 * M_exp(2090.0 * t_factor);
      out[2] = 6.0e-8 * M_exp(1120.0 * t_factor);
      break;
    case 1:
      out[0] = 3.1e-2 * M_exp(2423.0 * t_factor);
      out[1] = 4.3e-7 * M_exp(-913.0 * t_factor);
      break;
    case 2:
      out[0] = 7.4e4 * M_exp(6621.0 * t_factor);
      out[1] = 2.2e-12 * M_exp(-3730.0 * t_factor);
      break;
    default:
      out[0] = 1.15e-2 * M_exp(2560.0 * t_factor);
  }
}

/* calc_ynetpos (:189-223): [H+] of pH yph and the net positive charge */
double orc_sox_ynetpos(double yph, const double f[4], double Eco2, double Eh2o, double Eso4,
                       double so4_fact, double *xph) {
  *xph = M_pow(10.0, -yph);
  const double Eso2 = f[0] / (1.0 + f[1] * (1.0 + (f[2] / *xph) * (1.0 + f[3] / *xph)));
  const double tmp_hso3 = Eso2 / *xph;
  const double tmp_so3 = tmp_hso3 * 2.0 * f[3] / *xph;
  const double tmp_hco3 = Eco2 / *xph;
  const double tmp_oh = Eh2o / *xph;
  const double tmp_so4 = so4_fact * Eso4;
  return *xph - (tmp_oh + tmp_hco3 + tmp_hso3 + tmp_so3 + tmp_so4);
}

/* calc_ph_values (:225-413): bisection on pH in [2, 7].  f (out, optional):
 * fact1-4_so2, then Eco2, Eh2o, Eso4.  Returns [H+]; *converged as the
 * reference sets it. */
double orc_sox_ph(double temperature, double patm, double xlwc, double t_factor,
                  double xso2, double xso4, double xhnm, double so4_fact, int *converged,
                  double *fout) {
  double h[3];
  orc_sox_henry(0, t_factor, h);
  double f[4];
  f[0] = h[0] * h[1] * patm * xso2;
  f[1] = h[0] * SOX_RA * temperature * xlwc;
  f[2] = h[1];
  f[3] = h[2];
  const double Eh2o = SOX_XKW;
  orc_sox_henry(1, t_factor, h);
  const double Eco2 = h[0] * h[1] * SOX_CO2G * patm;
  const double Eso4 = xso4 * xhnm * SOX_CONST0 / xlwc;
  if (fout) {
    for (int i = 0; i < 4; ++i) fout[i] = f[i];
    fout[4] = Eco2;
    fout[5] = Eh2o;
    fout[6] = Eso4;
  }
  double xph, ynetpos;
  *converged = 0;
  double yph_lo = 2.00, yph_hi = yph_lo, yph = yph_lo;
  ynetpos = orc_sox_ynetpos(yph, f, Eco2, Eh2o, Eso4, so4_fact, &xph);
  if (ynetpos <= 0.0) *converged = 1;
  yph_hi = 7.0;
  yph = yph_hi;
  ynetpos = orc_sox_ynetpos(yph, f, Eco2, Eh2o, Eso4, so4_fact, &xph);
  if (ynetpos >= 0.0) *converged = 1;
  for (int i = 3; i < 20; ++i) {
    yph = 0.5 * (yph_lo + yph_hi);
    ynetpos = orc_sox_ynetpos(yph, f, Eco2, Eh2o, Eso4, so4_fact, &xph);
    if (ynetpos >= 0.0)
      yph_lo = yph;
    else
      yph_hi = yph;
    if (fabs(yph_hi - yph_lo) <= 0.005) {
      yph = 0.5 * (yph_hi + yph_lo);
      xph = M_pow(10.0, -yph);
      *converged = 1;
      return xph;
    }
  }
  return xph;
}

/* compute_aer_factor (:524-595) */
void orc_sox_aer_factor(const double tmr[31], double faq[4]) {
  double sumf = 0.0, qnum_c[4];
  for (int m = 0; m < 4; ++m) qnum_c[m] = faq[m] = 0.0;
  for (int m = 0; m < 4; ++m) {
    qnum_c[m] = kmax(0.0, tmr[SOX_NUM_CW[m]]);
    if (m == NACC) qnum_c[m] = kmax(1.0e-10, qnum_c[m]);
    if (m < 3) { /* lptr_so4_cw_amode > 0 for acc, ait, coarse */
      faq[m] = qnum_c[m];
      sumf += faq[m];
    }
  }
  if (sumf > 0.0)
    for (int m = 0; m < 4; ++m) faq[m] /= sumf;
}

/* cldaero_uptakerate (:597-695) */
double orc_sox_uptakerate(double xl, double cldnum, double cfact, double cldfrc, double tfld,
                          double press) {
  const double pi_x4 = 12.56637, one_third = 0.3333333, cm3_to_L = 1.0e-3;
  double num_cd = cm3_to_L * cldnum * cfact / cldfrc;
  num_cd = kmax(num_cd, 0.0);
  const double volx34pi_cd = (xl * 0.75) / 3.14159265358979323846;
  double radxnum_cd = M_pow(volx34pi_cd * sq(num_cd), one_third);
  double rad_cd;
  if (radxnum_cd <= volx34pi_cd * 4.0e4) {
    radxnum_cd = volx34pi_cd * 4.0e4;
    rad_cd = 50.0e-4;
  } else if (radxnum_cd >= volx34pi_cd * 4.0e8) {
    radxnum_cd = volx34pi_cd * 4.0e8;
    rad_cd = 0.5e-4;
  } else {
    rad_cd = radxnum_cd / num_cd;
  }
  const double gasdiffus = 0.557 * M_pow(tfld, 1.75) / press;
  const double gasspeed = 1.455e4 * M_sqrt(tfld / 98.0);
  const double knudsen = 3.0 * gasdiffus / (gasspeed * rad_cd);
  const double fs = (0.4875 * (1.0 + knudsen)) / (knudsen * (1.184 + knudsen) + 0.4875);
  return pi_x4 * radxnum_cd * gasdiffus * fs;
}

/* setsox_single_level (:943-1259) with calc_sox_aqueous (:415-522) and
 * sox_cldaero_update (:732-941) inline.  qcw/qin updated in place;
 * dqdt_aqso4/dqdt_aqh2so4 [31] zeroed and filled; ph_out (optional): the
 * [H+] used. */
void orc_setsox_single_level(double dt, double press, double tfld, double lwc, double cldfrc,
                             double cldnum, double xhnm, double qcw[31], double qin[31],
                             double dqdt_aqso4[31], double dqdt_aqh2so4[31], double *ph_out) {
  const double t298K = 298.0;
  double xph = M_pow(10, -5.0); /* ph0 */
  const double cfact = xhnm * 1.0e6 * 1.38e-23 / 287.0 * 1.0e-3;
  double cc[3];
  orc_sox_create_obj(cldfrc, qcw, lwc, cfact, cc);
  double xso4 = 0.0;
  double xso2 = qin[3], xh2o2 = qin[1], xo3 = qin[0], xh2so4 = qin[2];
  const double xso4c = cc[0];
  double xlwc = cc[1];
  if (xlwc >= 1.0e-8) {
    const double t_factor = (1.0 / tfld) - (1.0 / t298K);
    const double patm = press / SOX_P0;
    if (cldfrc > 0.0) xso4 = xso4c / cldfrc; /* cloud_borne */
    int converged = 0;
    xph = orc_sox_ph(tfld, patm, xlwc, t_factor, xso2, xso4, xhnm, cc[2], &converged, NULL);
  } else {
    xph = 1.0e-7;
  }
  if (ph_out) *ph_out = xph;
  const double t_factor = (1.0 / tfld) - (1.0 / t298K);
  xlwc = cc[1];
  const double patm = press / SOX_P0;
  double h[3];
  orc_sox_henry(2, t_factor, h);
  const double heh2o2 = h[0] * (1.0 + h[1] / xph);
  orc_sox_henry(0, t_factor, h);
  const double heso2 = h[0] * (1.0 + h[1] / xph * (1.0 + h[2] / xph));
  orc_sox_henry(3, t_factor, h);
  const double heo3 = h[0];
  const double tz = tfld;
  double px = heh2o2 * SOX_RA * tz * xlwc;
  const double h2o2g = xh2o2 / (1.0 + px);
  px = heso2 * SOX_RA * tz * xlwc;
  const double so2g = xso2 / (1.0 + px);
  px = heo3 * SOX_RA * tz * xlwc;
  const double o3g = xo3 / (1.0 + px);
  const double rah2o2 = 8.e4 * M_exp(-3650.0 * t_factor) / (0.1 + xph);
  const double rao3 = 4.39e11 * M_exp(-4131.0 / tz) + 2.56e3 * M_exp(-996.0 / tz) / xph;
  double xdelso4hp = 0.0, xso4_init = 0.0;
  if (xlwc >= 1.0e-8) {
    /* calc_sox_aqueous (:415-522) */
    double pso4 = rah2o2 * 7.4e4 * M_exp(6621.0 * t_factor) * h2o2g * patm * 1.23 *
                  M_exp(3120.0 * t_factor) * so2g * patm;
    pso4 = pso4 * xlwc / SOX_CONST0 / xhnm;
    double delta_s = kmax(pso4 * dt, 1.0e-30);
    xso4_init = xso4;
    if ((delta_s <= xso2) && (delta_s <= xh2o2)) {
      xso4 = xso4 + delta_s;
      xh2o2 = xh2o2 - delta_s;
      xso2 = xso2 - delta_s;
    } else if (xh2o2 > xso2) {
      xso4 = xso4 + xso2;
      xh2o2 = xh2o2 - xso2;
      xso2 = 1.0e-20;
    } else {
      xso4 = xso4 + xh2o2;
      xso2 = xso2 - xh2o2;
      xh2o2 = 1.0e-20;
    }
    xdelso4hp = xso4 - xso4_init; /* modal_aerosols */
    pso4 = rao3 * heo3 * o3g * patm * heso2 * so2g * patm;
    pso4 = pso4 * xlwc / SOX_CONST0 / xhnm;
    delta_s = kmax(pso4 * dt, 1.0e-30);
    xso4_init = xso4;
    if (delta_s > xso2) {
      xso4 = xso4 + xso2;
      xso2 = 1.0e-20;
    } else {
      xso4 = xso4 + delta_s;
      xso2 = xso2 - delta_s;
    }
  }
  /* sox_cldaero_update (:732-941) */
  for (int i = 0; i < 31; ++i) dqdt_aqso4[i] = dqdt_aqh2so4[i] = 0.0;
  if ((cldfrc >= 1.0e-5) && (xlwc >= 1.0e-8)) {
    double faq[4];
    orc_sox_aer_factor(qcw, faq);
    double uptkrate = orc_sox_uptakerate(xlwc, cldnum, cfact, cldfrc, tfld, press);
    uptkrate = (1.0 - M_exp(-1.0 * kmin(100.0, dt * uptkrate))) / dt;
    const double dso4dt_gasuptk = xh2so4 * uptkrate;
    const double delso4_o3rxn = xso4 - xso4_init;
    const double dso4dt_aqrxn = (delso4_o3rxn + xdelso4hp) / dt;
    const double dso4dt_hprxn = xdelso4hp / dt;
    for (int m = 0; m < 3; ++m) {
      const int l = SOX_SO4_CW[m];
      dqdt_aqso4[l] = faq[m] * dso4dt_aqrxn * cldfrc;
      dqdt_aqh2so4[l] = faq[m] * dso4dt_gasuptk * cldfrc;
      const double dqdt_aq = dqdt_aqso4[l] + dqdt_aqh2so4[l];
      qcw[l] += (dqdt_aq + 0.0) * dt;
    }
    qin[2] = qin[2] - dso4dt_gasuptk * dt * cldfrc;
    qin[3] += (-dso4dt_aqrxn * cldfrc + 0.0) * dt;
    qin[1] += (-dso4dt_hprxn * cldfrc + 0.0) * dt;
  }
  for (int m = 0; m < 3; ++m) qcw[SOX_SO4_CW[m]] = kmax(qcw[SOX_SO4_CW[m]], 1.0e-20);
  qin[3] = kmax(qin[3], 1.0e-20);
}

/* setsox over a domain of planes [.][ncell]: met planes press, tfld, lwc,
 * cldfrc, cldnum, xhnm; qcw/qin [31][ncell] updated in place; dqdt_*
 * [31][ncell] (NULL to skip). */
void orc_setsox(long ncell, double dt, const double *press, const double *tfld, const double *lwc,
                const double *cldfrc, const double *cldnum, const double *xhnm, double *qcw,
                double *qin, double *dqdt_aqso4, double *dqdt_aqh2so4) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (long c = 0; c < ncell; ++c) {
    if (g_noise_seed) noise_cell(c);
    double qc[31], qi[31], d1[31], d2[31];
    for (int i = 0; i < 31; ++i) {
      qc[i] = qcw[(long)i * ncell + c];
      qi[i] = qin[(long)i * ncell + c];
    }
    orc_setsox_single_level(dt, press[c], tfld[c], lwc[c], cldfrc[c], cldnum[c], xhnm[c], qc, qi,
                            d1, d2, NULL);
    for (int i = 0; i < 31; ++i) {
      qcw[(long)i * ncell + c] = qc[i];
      qin[(long)i * ncell + c] = qi[i];
      if (dqdt_aqso4) dqdt_aqso4[(long)i * ncell + c] = d1[i];
      if (dqdt_aqh2so4) dqdt_aqh2so4[(long)i * ncell + c] = d2[i];
    }
  }
}

/* ======================= gas-phase chemistry ============================= */
/* gas_phase_chemistry (gas_phase_chemistry.hpp:37-116) with the
 * pp_linoz_mam4_resus_mom_soag mechanism (gas_chem_mechanism.hpp) and the
 * implicit solver imp_sol / newton_raphson_iter (gas_chem.hpp:14-425), per
 * cell.  qq[31] is the gas_pcnst array (mol/mol); the implicit class is
 * species 1..30 (clsmap_4), permute_4 the identity.  The mechanism's sparse
 * LU has 32 entries: a diagonal plus the SO2 -> H2SO4 and DMS -> SO2
 * couplings. */
#define GC_RXNTOT 7
#define GC_CLS 30
#define GC_NZ 32
#define GC_EXTCNT 9
#define GC_NFS 8

/* setrxt (gas_chem_mechanism.hpp:40-45), usrrxt (gas_chem.hpp:20-63) with
 * the indices of gas_phase_chemistry.hpp:28-29, adjrxt (:63-71) */
void orc_gc_rates(double temp, const double inv[GC_NFS], double rxt[GC_RXNTOT]) {
  for (int i = 0; i < GC_RXNTOT; ++i) rxt[i] = 0.0;
  rxt[2] = 2.9000000000e-12 * M_exp(-160.000000 / temp);
  rxt[4] = 9.6000000000e-12 * M_exp(-234.000000 / temp);
  rxt[6] = 1.9000000000e-13 * M_exp(520.000000 / temp);
  const double mtot = inv[0];
  { /* HO2 + HO2 -> H2O2 (index 1; water vapour invariant 3) */
    const double ko = 3.5e-13 * M_exp(430.0 / temp);
    const double kinf = 1.7e-33 * mtot * M_exp(1000. / temp);
    const double fc = 1.0 + 1.4e-21 * inv[3] * M_exp(2200. / temp);
    rxt[1] = (ko + kinf) * fc;
  }
  { /* DMS + OH -> 0.5 SO2 (index 5) */
    const double ko = 1.0 + 5.5e-31 * M_exp(7460. / temp) * mtot * 0.21;
    rxt[5] = 1.7e-42 * M_exp(7810. / temp) * mtot * 0.21 / ko;
  }
  { /* SO2 + OH -> SO4 (index 3) */
    const double fc = 3.0e-31 * M_pow(300. / temp, 3.3);
    const double ko = fc * mtot / (1.0 + fc * mtot / 1.5e-12);
    const double lg = log10(fc * mtot / 1.5e-12);
    rxt[3] = ko * M_pow(0.6, 1.0 / (1.0 + lg * lg));
  }
  rxt[2] *= inv[4];
  rxt[3] *= inv[4];
  rxt[4] *= inv[4];
  rxt[5] *= inv[4];
  rxt[6] *= inv[5];
  rxt[1] *= inv[6] * inv[6] / mtot;
}

/* imp_prod_loss (gas_chem_mechanism.hpp:82-99) */
static void gc_prod_loss(double prod[GC_CLS], double loss[GC_CLS], const double y[31],
                         const double rxt[GC_RXNTOT], const double het[31]) {
  loss[0] = (het[1] + rxt[0] + rxt[2]) * y[1];
  prod[0] = 0.0;
  loss[1] = het[2] * y[2];
  prod[1] = rxt[3] * y[3];
  loss[2] = (het[3] + rxt[3]) * y[3];
  prod[2] = (rxt[4] + 0.500000 * rxt[5] + rxt[6]) * y[4];
  loss[3] = (het[4] + rxt[4] + rxt[5] + rxt[6]) * y[4];
  prod[3] = 0.0;
  for (int i = 4; i < GC_CLS; ++i) {
    loss[i] = het[i + 1] * y[i + 1];
    prod[i] = 0.0;
  }
}

/* indprd, class 4 (gas_chem_mechanism.hpp:100-143): the external forcing
 * enters species H2O2 (via rxt[1]), SO2, DMS, SOAG and six aerosol tracers */
static void gc_indprd(double prod[GC_CLS], const double rxt[GC_RXNTOT],
                      const double extfrc[GC_EXTCNT]) {
  for (int i = 0; i < GC_CLS; ++i) prod[i] = 0.0;
  prod[0] = +rxt[1];
  prod[2] = +extfrc[0];
  prod[4] = +extfrc[8];
  prod[5] = +extfrc[1];
  prod[12] = +extfrc[5];
  prod[13] = +extfrc[2];
  prod[17] = +extfrc[6];
  prod[26] = +extfrc[3];
  prod[27] = +extfrc[4];
  prod[29] = +extfrc[7];
}

/* linmat + nlnmat + lu_fac (gas_chem_mechanism.hpp:145-300): the factored
 * system matrix; entries 2 and 4 are the couplings */
static void gc_sys(double lu[GC_NZ], const double rxt[GC_RXNTOT], const double het[31],
                   double dti) {
  double m[GC_NZ];
  m[0] = -(+rxt[0] + rxt[2] + het[1]);
  m[1] = -(+het[2]);
  m[2] = +rxt[3];
  m[3] = -(+rxt[3] + het[3]);
  m[4] = +rxt[4] + 0.500000 * rxt[5] + rxt[6];
  m[5] = -(+rxt[4] + rxt[5] + rxt[6] + het[4]);
  for (int i = 6; i < GC_NZ; ++i) m[i] = -(+het[i - 1]);
  for (int i = 0; i < GC_NZ; ++i) lu[i] = (i == 2 || i == 4) ? m[i] : m[i] - dti;
  for (int i = 0; i < GC_NZ; ++i)
    if (i != 2 && i != 4) lu[i] = 1.0 / lu[i];
}

/* lu_slv (gas_chem_mechanism.hpp:183-228) */
static void gc_lu_slv(const double lu[GC_NZ], double b[GC_CLS]) {
  for (int k = 29; k >= 3; --k) b[k] *= lu[k + 2];
  b[2] -= lu[4] * b[3];
  b[2] *= lu[3];
  b[1] -= lu[2] * b[2];
  b[1] *= lu[1];
  b[0] *= lu[0];
}

/* imp_sol (gas_chem.hpp:230-425) with newton_raphson_iter (:73-228).
 * base_sol [31] advanced over delt in place; returns the number of time
 * steps taken. */
int orc_gc_imp_sol(double base_sol[31], const double rxt[GC_RXNTOT], const double het[31],
                   const double extfrc[GC_EXTCNT], double delt) {
  double ind_prd[GC_CLS], solution[GC_CLS], iter_inv[GC_CLS], prod[GC_CLS], loss[GC_CLS];
  double lu[GC_NZ], forcing[GC_CLS];
  gc_indprd(ind_prd, rxt, extfrc);
  double dt = delt, interval_done = 0.0;
  int cut_cnt = 0, fail_cnt = 0, stp_con_cnt = 0, nstep = 0;
  for (int i = 0; i < 1000; ++i) {
    ++nstep;
    const double dti = 1.0 / dt;
    for (int k = 0; k < GC_CLS; ++k) solution[k] = base_sol[k + 1];
    for (int k = 0; k < GC_CLS; ++k) iter_inv[k] = dti * solution[k] + ind_prd[k];
    int convergence = 0;
    for (int it = 0; it < 11; ++it) {
      gc_sys(lu, rxt, het, dti); /* factor[] is all true */
      gc_prod_loss(prod, loss, base_sol, rxt, het);
      for (int k = 0; k < GC_CLS; ++k)
        forcing[k] = solution[k] * dti - (iter_inv[k] + prod[k] - loss[k]);
      gc_lu_slv(lu, forcing);
      for (int k = 0; k < GC_CLS; ++k) solution[k] += forcing[k];
      for (int k = 0; k < GC_CLS; ++k)
        if (solution[k] < 0.0) solution[k] = 0.0;
      for (int k = 0; k < GC_CLS; ++k) base_sol[k + 1] = solution[k];
      if (it > 0) {
        convergence = 1;
        for (int k = 0; k < GC_CLS; ++k) {
          int conv = 1;
          if (fabs(forcing[k]) > 1.0e-40) conv = fabs(forcing[k]) <= 1.0e-3 * fabs(solution[k]);
          if (!conv) convergence = 0;
        }
        if (convergence) break;
      }
    }
    if (!convergence) {
      fail_cnt += fail_cnt; /* as the reference (stays 0) */
      stp_con_cnt = 0;
      if (cut_cnt < 5) {
        cut_cnt += 1;
        if (cut_cnt < 5)
          dt *= 0.5;
        else
          dt *= 0.1;
      }
    }
    interval_done += dt;
    if (fabs(delt - interval_done) <= 0.0001) break;
    if (convergence) stp_con_cnt += 1;
    if (stp_con_cnt >= 2) {
      dt *= 2.0;
      stp_con_cnt = 0;
    }
    dt = kmin(dt, delt - interval_done);
  }
  for (int k = 0; k < GC_CLS; ++k) base_sol[k + 1] = solution[k];
  return nstep;
}

/* gas_phase_chemistry for one cell: photo_rate = rxt[0] (phtcnt = 1);
 * extfrc [1/cm3/s] scaled by the total density; qq[31] updated in place.
 * Returns the time steps taken. */
int orc_gas_phase_chemistry(double temp, double dt, double photo_rate,
                            const double extfrc[GC_EXTCNT], const double inv[GC_NFS],
                            const double het[31], double qq[31]) {
  double rxt[GC_RXNTOT], ext[GC_EXTCNT];
  orc_gc_rates(temp, inv, rxt);
  for (int m = 0; m < GC_EXTCNT; ++m) ext[m] = extfrc[m] / inv[0];
  rxt[0] = photo_rate;
  return orc_gc_imp_sol(qq, rxt, het, ext, dt);
}

/* over a domain: temp, photo [ncell]; extfrc [9][ncell], inv [8][ncell],
 * het [31][ncell], qq [31][ncell] in place; steps (optional) [ncell] */
void orc_gas_phase_chemistry_domain(long ncell, double dt, const double *temp,
                                    const double *photo, const double *extfrc,
                                    const double *inv, const double *het, double *qq,
                                    int *steps) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (long c = 0; c < ncell; ++c) {
    if (g_noise_seed) noise_cell(c);
    double e[GC_EXTCNT], v[GC_NFS], h[31], q[31];
    for (int i = 0; i < GC_EXTCNT; ++i) e[i] = extfrc[(long)i * ncell + c];
    for (int i = 0; i < GC_NFS; ++i) v[i] = inv[(long)i * ncell + c];
    for (int i = 0; i < 31; ++i) {
      h[i] = het[(long)i * ncell + c];
      q[i] = qq[(long)i * ncell + c];
    }
    const int n = orc_gas_phase_chemistry(temp[c], dt, photo[c], e, v, h, q);
    for (int i = 0; i < 31; ++i) qq[(long)i * ncell + c] = q[i];
    if (steps) steps[c] = n;
  }
}

/* ================ stratospheric ozone (linoz) and invariants =============
 * compute_o3_column_density (compute_o3_column_density.hpp:59-83),
 * lin_strat_chem_solve_kk / psc_activation / compute_steady_state_ozone
 * (lin_strat_chem.hpp:31-222), lin_strat_sfcsink(_kk) (:263-316) and
 * setinv_single_level (mo_setinv.hpp:45-86). */
#define LNZ_PI 3.14159265358979323846264
static const double lnz_rad2deg = 180. / LNZ_PI;

/* compute_o3_column_density.hpp:59-83: o3_col_dens(k) = col0 + sum of the
 * levels above + half of level k's own slab [mol/cm2] */
void orc_o3_column_density(int nlev, const double *pdel, const double *vmr_o3, double col0,
                           double *o3_col_dens) {
  const double xfactor = 2.8704e21 / (9.80616 * 1.38044);
  double partial_sum = 0.0;
  for (int kk = 0; kk < nlev; ++kk) {
    const double delta_kk = xfactor * pdel[kk] * vmr_o3[kk];
    o3_col_dens[kk] = col0 + partial_sum + 0.5 * delta_kk;
    partial_sum += delta_kk;
  }
}

/* lin_strat_chem.hpp:31-95 */
static void psc_activation(double lats, double temp, double pmid, double sza,
                           double linoz_cariolle_psc, double delta_t, int excess_chlorine,
                           double o3_old, double psc_T, double chlorine_loading,
                           double *o3_new, double *do3_linoz_psc) {
  const double lats_threshold = 40.0, chlorine_loading_1987 = 2.5977;
  if (fabs(lats) > lats_threshold) {
    if (excess_chlorine) {
      if (temp <= psc_T) {
        const double max_sza = 90.0 + M_sqrt(kmax(16.0 * M_log10(1e5 / pmid), 0.0));
        if ((sza * lnz_rad2deg) <= max_sza) {
          const double r = chlorine_loading / chlorine_loading_1987;
          const double psc_loss = M_exp(-linoz_cariolle_psc * (r * r) * delta_t);
          *o3_new = o3_old * psc_loss;
          *do3_linoz_psc = (*o3_new - o3_old) / delta_t;
        }
      }
    }
  }
}

/* lin_strat_chem.hpp:97-222.  clim: o3_clim, t_clim, o3col_clim, PmL_clim,
 * dPmL_dO3, dPmL_dT, dPmL_dO3col, cariolle_psc.  diag[3] = do3_linoz,
 * do3_linoz_psc (0 unless the PSC loss applies), ss_o3. */
void orc_linoz_solve_kk(double o3col, double temperature, double sza, double pmid,
                        double delta_t, double rlats, const double clim[8],
                        double chlorine_loading, double psc_T, double *o3_vmr, double diag[3]) {
  const double convert_to_du = 1.0 / 2.687e16;
  const double lats = rlats * lnz_rad2deg;
  const int excess_chlorine = (chlorine_loading - 0.0000) > 0;
  const double o3_old = *o3_vmr;
  const double o3col_du = o3col * convert_to_du;
  /* compute_steady_state_ozone */
  const double delta_temperature = temperature - clim[1];
  const double delta_o3col = o3col_du - clim[2];
  const double ss_o3 =
      clim[0] - (clim[3] + delta_o3col * clim[6] + delta_temperature * clim[5]) / clim[4];
  const double delta_o3 = (ss_o3 - o3_old) * (1.0 - M_exp(clim[4] * delta_t));
  double o3_new = o3_old + delta_o3;
  double do3_psc = 0.0;
  psc_activation(lats, temperature, pmid, sza, clim[7], delta_t, excess_chlorine, o3_old, psc_T,
                 chlorine_loading, &o3_new, &do3_psc);
  *o3_vmr = o3_new;
  diag[0] = delta_o3 / delta_t;
  diag[1] = do3_psc;
  diag[2] = ss_o3;
}

/* lin_strat_chem.hpp:263-316: returns o3l_sfcsink [Tg/yr/m2] */
double orc_linoz_sfcsink(int nlev, double delta_t, const double *pdel, double *o3l_vmr,
                         double o3_sfc, int o3_lbl, double o3_tau) {
  const double mwo3 = 48., mwdry = 0.028966 * 1e3, rgrav = 1.0 / 9.80616;
  double do3mass = 0.0;
  for (int kk = nlev - 1; kk > nlev - o3_lbl - 1; --kk) {
    const double efactor = 1.0 - M_exp(-delta_t / o3_tau);
    const double mass = pdel[kk] * rgrav;
    const double do3 = (o3_sfc - o3l_vmr[kk]) * efactor;
    o3l_vmr[kk] = o3l_vmr[kk] + do3;
    do3mass += do3 * mass * mwo3 / mwdry;
  }
  return do3mass / delta_t * 1.0e-9 * (86400. * 365.0);
}

/* one column, levels contiguous: the column density from the ozone before
 * the step, the solve above the tropopause (k < ltrop), then the surface
 * sink (o3_lbl > 0).  Optional outputs: o3col, do3, do3_psc, ss [nlev]
 * (0 at k >= ltrop), sfcsink. */
void orc_linoz_column(int nlev, double dt, double sza, double rlats, int ltrop, double col0,
                      double chlorine_loading, double psc_T, double o3_sfc, double o3_tau,
                      int o3_lbl, const double *temp, const double *pmid, const double *pdel,
                      const double *clim /* [8][nlev] */, double *o3, double *o3col,
                      double *do3, double *do3_psc, double *ss, double *sfcsink) {
  double *col = (double *)malloc(sizeof(double) * (nlev > 0 ? nlev : 1));
  orc_o3_column_density(nlev, pdel, o3, col0, col);
  for (int k = 0; k < nlev; ++k) {
    double d[3] = {0.0, 0.0, 0.0};
    if (k < ltrop) {
      double c[8];
      for (int j = 0; j < 8; ++j) c[j] = clim[(long)j * nlev + k];
      orc_linoz_solve_kk(col[k], temp[k], sza, pmid[k], dt, rlats, c, chlorine_loading, psc_T,
                         &o3[k], d);
    }
    if (o3col) o3col[k] = col[k];
    if (do3) do3[k] = d[0];
    if (do3_psc) do3_psc[k] = d[1];
    if (ss) ss[k] = d[2];
  }
  const double s = o3_lbl > 0 ? orc_linoz_sfcsink(nlev, dt, pdel, o3, o3_sfc, o3_lbl, o3_tau)
                              : 0.0;
  if (sfcsink) *sfcsink = s;
  free(col);
}

/* mo_setinv.hpp:45-86 with the default Config: inv[8] [#/cm3] */
void orc_setinv_single_level(double tfld, double h2ovmr, double pmid,
                             const double cnst_offline[4], double inv[8]) {
  const double Pa_xfac = 10.0, boltz_cgs = 0.13806500000000001E-015;
  const int coff_idx_list[4] = {7, 4, 5, 6};
  for (int i = 0; i < 8; ++i) inv[i] = 0.0;
  inv[0] = Pa_xfac * pmid / (boltz_cgs * tfld);
  const double inv_m = inv[0];
  inv[1] = 0.79 * inv_m;
  inv[2] = 0.21 * inv_m;
  inv[3] = h2ovmr * inv_m;
  for (int i = 0; i < 4; ++i) inv[coff_idx_list[i]] = cnst_offline[i] * inv_m;
}

/* setinv over cells: T, p, qv (mass mixing ratio; vmr_from_mmr with the
 * water molecular weight, mo_setinv.hpp:103) [ncell], cnst_offline
 * [4][ncell], inv [8][ncell] */
void orc_setinv(long ncell, const double *tfld, const double *pmid, const double *qv,
                const double *cnst_offline, double *inv) {
  for (long c = 0; c < ncell; ++c) {
    double co[4], v[8];
    for (int i = 0; i < 4; ++i) co[i] = cnst_offline[(long)i * ncell + c];
    const double h2ovmr = qv[c] * 0.028966 / 0.018016;
    orc_setinv_single_level(tfld[c], h2ovmr, pmid[c], co, v);
    for (int i = 0; i < 8; ++i) inv[(long)i * ncell + c] = v[i];
  }
}

/* orc_linoz_column over a domain of [nlev][ncol] planes (column fastest);
 * clim [8][nlev][ncol]; per column sza, rlats, col0, ltrop; optional
 * outputs as orc_linoz_column (planes) and sfcsink [ncol].  OpenMP over
 * columns. */
void orc_linoz_domain(int ncol, int nlev, double dt, double chlorine_loading, double psc_T,
                      double o3_sfc, double o3_tau, int o3_lbl, const double *temp,
                      const double *pmid, const double *pdel, const double *clim,
                      const double *sza, const double *rlats, const double *col0,
                      const int *ltrop, double *o3, double *o3col, double *do3,
                      double *do3_psc, double *ss, double *sfcsink) {
  const long plane = (long)nlev * ncol;
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *buf = (double *)malloc(sizeof(double) * (size_t)nlev * 16 + 8);
    double *t = buf, *pm = t + nlev, *pd = pm + nlev, *o = pd + nlev, *oc = o + nlev,
           *d1 = oc + nlev, *d2 = d1 + nlev, *sv = d2 + nlev, *cl = sv + nlev;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int j = 0; j < ncol; ++j) {
      for (int k = 0; k < nlev; ++k) {
        const long i = (long)k * ncol + j;
        t[k] = temp[i];
        pm[k] = pmid[i];
        pd[k] = pdel[i];
        o[k] = o3[i];
        for (int c = 0; c < 8; ++c) cl[(long)c * nlev + k] = clim[c * plane + i];
      }
      double s = 0.0;
      orc_linoz_column(nlev, dt, sza[j], rlats[j], ltrop[j], col0 ? col0[j] : 0.0,
                       chlorine_loading, psc_T, o3_sfc, o3_tau, o3_lbl, t, pm, pd, cl, o, oc, d1,
                       d2, sv, &s);
      for (int k = 0; k < nlev; ++k) {
        const long i = (long)k * ncol + j;
        o3[i] = o[k];
        if (o3col) o3col[i] = oc[k];
        if (do3) do3[i] = d1[k];
        if (do3_psc) do3_psc[i] = d2[k];
        if (ss) ss[i] = sv[k];
      }
      if (sfcsink) sfcsink[j] = s;
    }
    free(buf);
  }
}

/* ============================ washout (sethet) ============================
 * mo_sethet::sethet_detail (mo_sethet.hpp:217-500) for one column with the
 * wrapper's indices (:549-556): H2O2 (1), H2SO4 (2), SO2 (3) are the wet-
 * deposited species.  Levels contiguous; level 0 is the model top. */
#define SH_MASS_H2O (0.018016 * 1000)
static const double sh_avo = 6.022214076e23;
static const double sh_boltz_cgs = 1.380649e-23 * 1.e7;

/* mo_sethet.hpp:36-52 */
static double sh_calc_het_rates(double satf, double rain, double xhen, double tmp_hetrates,
                                double work1, double work2) {
  const double h2o_mol = 1.0e3 / SH_MASS_H2O;
  const double work3 = satf * kmax(rain / (h2o_mol * (work1 + 1.0 / (xhen * work2))), 0.0);
  return work3 + tmp_hetrates;
}

/* mo_sethet.hpp:59-103 */
void orc_precip_rescale(int nlev, const double *cmfdqr, const double *nrain,
                        const double *nevapr, double *precip) {
  double total_rain = 0, total_pos = 0;
  for (int kk = 0; kk < nlev; ++kk) {
    precip[kk] = cmfdqr[kk] + nrain[kk] - nevapr[kk];
    total_rain += precip[kk];
  }
  for (int kk = 0; kk < nlev; ++kk) {
    if (precip[kk] < 0.0) precip[kk] = 0.0;
    total_pos += precip[kk];
  }
  for (int kk = 0; kk < nlev; ++kk)
    precip[kk] = total_rain <= 0.0 ? 0.0 : precip[kk] * total_rain / total_pos;
}

/* mo_sethet.hpp:107-170: in-place washout of xgas by the rain of every
 * raining level kk >= ktop onto the levels below it */
void orc_gas_washout(int nlev, int ktop, double xkgm, const double *xliq, const double *rain,
                     const double *xhen_i, const double *tfld_i, const double *delz_i,
                     double *xgas) {
  const double const0 = sh_boltz_cgs * 1.0e-6, geo_fac = 6.0, xrm = .189, xum = 748.0;
  const double cm3_2_m3 = 1.0e-6, avo2 = sh_avo * 1.0e-3 * cm3_2_m3;
  double *scratch = (double *)malloc(sizeof(double) * (nlev > 0 ? nlev : 1));
  for (int kk = ktop; kk < nlev; ++kk) {
    if (rain[kk] != 0.0) {
      double accumulator = 0.0;
      for (int k = kk; k < nlev; ++k) {
        const double xca = geo_fac * xkgm * xgas[k] / (xrm * xum) * delz_i[k] * xliq[kk] *
                           cm3_2_m3;
        accumulator += xca;
        scratch[k] = accumulator;
      }
      for (int k = kk; k < nlev; ++k) {
        const double xeqca = xgas[k] /
                             (xliq[kk] * avo2 + 1.0 / (xhen_i[k] * const0 * tfld_i[k])) *
                             xliq[kk] * avo2;
        if (scratch[k] < xeqca) {
          const double xca = geo_fac * xkgm * xgas[k] / (xrm * xum) * delz_i[k] * xliq[kk] *
                             cm3_2_m3;
          xgas[k] = kmax(xgas[k] - xca, 0.0);
        }
      }
    }
  }
  free(scratch);
}

/* mo_sethet.hpp:173-197: the lowest level above p_limit (0 if none) */
int orc_find_ktop(int nlev, double rlat, const double *press) {
  const double d2r = 3.14159265358979323846264 / 180.0;
  const double p_limit = fabs(rlat) > 60.0 * d2r ? 300.0e2 : 100.0e2;
  for (int kk = nlev - 1; kk >= 0; kk--)
    if (press[kk] < p_limit) return kk;
  return 0;
}

/* the rain drops' mass-transfer coefficient (mo_sethet.hpp:330-331) */
double orc_sethet_xkgm(void) {
  const double xrm = .189, xum = 748.0, xvv = 6.18e-2, xdg = .112;
  return xdg / xrm * 2.0 + xdg / xrm * .6 * M_sqrt(xrm * xum / xvv) * M_pow((xvv / xdg), (1.0 / 3.0));
}

/* one column: press [Pa], zmid (the reference's zmid argument, see
 * sethet_detail), tfld, cmfdqr, nrain, nevapr [kg/kg/s], inv_m [#/cm3],
 * qin_h2o2 / qin_so2 [vmr] (levels contiguous); het [31][nlev] out */
void orc_sethet_column(int nlev, double rlat, double phis, double delt, const double *press,
                       const double *zmid, const double *tfld, const double *cmfdqr,
                       const double *nrain, const double *nevapr, const double *inv_m,
                       const double *qin_h2o2, const double *qin_so2, double *het) {
  const double xrm = .189, xum = 748.0, t0 = 298.0, xph0 = 1.0e-5, satf_hno3 = .016,
               satf_h2o2 = .016, const0 = sh_boltz_cgs * 1.0e-6, hno3_diss = 15.4,
               mass_air = 29.0, km2cm = 1.0e5, m2km = 1.0e-3, m3_2_cm3 = 1.0e6,
               MISSING = -999999.0, large_value_lifetime = 1.0e29;
  const double avo2 = sh_avo * 1.0e-3 * 1.0e-6, rga = 1.0 / 9.80616;
  (void)xrm;
  const size_t n = (size_t)(nlev > 0 ? nlev : 1);
  double *w = (double *)calloc(n * 15, sizeof(double));
  double *t_factor = w, *xk0_hno3 = w + n, *xk0_so2 = w + 2 * n, *so2_diss = w + 3 * n,
         *xgas2 = w + 4 * n, *xgas3 = w + 5 * n, *delz = w + 6 * n, *xh2o2 = w + 7 * n,
         *xso2 = w + 8 * n, *xliq = w + 9 * n, *rain = w + 10 * n, *precip = w + 11 * n,
         *xhen_h2o2 = w + 12 * n, *xhen_hno3 = w + 13 * n, *xhen_so2 = w + 14 * n;
  double *tmp1 = (double *)calloc(n, sizeof(double)), *tmp2 = (double *)calloc(n, sizeof(double));
  const int wetdep_map[3] = {1, 2, 3};
  for (int kk = 0; kk < nlev; ++kk)
    for (int mm = 0; mm < 31; ++mm) het[(long)mm * nlev + kk] = 0.0;
  for (int kk = 0; kk < nlev; ++kk)
    for (int mm = 0; mm < 3; ++mm) het[(long)wetdep_map[mm] * nlev + kk] = MISSING;
  const double xkgm = orc_sethet_xkgm();
  const int ktop = orc_find_ktop(nlev, rlat, press);
  orc_precip_rescale(nlev, cmfdqr, nrain, nevapr, precip);
  for (int kk = 0; kk < nlev; ++kk) {
    rain[kk] = mass_air * precip[kk] * inv_m[kk] / SH_MASS_H2O;
    xliq[kk] = precip[kk] * delt * inv_m[kk] / sh_avo * mass_air * m3_2_cm3;
    xh2o2[kk] = qin_h2o2[kk] * inv_m[kk];
    xso2[kk] = qin_so2[kk] * inv_m[kk];
  }
  const double zsurf = m2km * phis * rga;
  for (int kk = ktop; kk < nlev - 1; ++kk) delz[kk] = fabs((zmid[kk] - zmid[kk + 1]) * km2cm);
  delz[nlev - 1] = fabs((zmid[nlev - 1] - zsurf) * km2cm);
  for (int kk = ktop; kk < nlev; ++kk) {
    t_factor[kk] = (t0 - tfld[kk]) / (t0 * tfld[kk]);
    xhen_h2o2[kk] = 7.45e4 * M_exp(6620.0 * t_factor[kk]);
    xk0_hno3[kk] = 2.1e5 * M_exp(8700.0 * t_factor[kk]);
    xhen_hno3[kk] = xk0_hno3[kk] * (1.0 + hno3_diss / xph0);
    xk0_so2[kk] = 1.23 * M_exp(3120.0 * t_factor[kk]);
    so2_diss[kk] = 1.23e-2 * M_exp(1960.0 * t_factor[kk]);
    xhen_so2[kk] = xk0_so2[kk] * (1.0 + so2_diss[kk] / xph0);
  }
  for (int kk = 0; kk < nlev; ++kk) {
    xgas2[kk] = xh2o2[kk];
    xgas3[kk] = xso2[kk];
  }
  orc_gas_washout(nlev, ktop, xkgm, xliq, rain, xhen_h2o2, tfld, delz, xgas2);
  orc_gas_washout(nlev, ktop, xkgm, xliq, rain, xhen_so2, tfld, delz, xgas3);
  for (int kk = ktop; kk < nlev; ++kk) {
    double stay = 1.0;
    if (rain[kk] != 0.0) {
      stay = ((zmid[kk] - zsurf) * km2cm) / (xum * delt);
      stay = kmin(stay, 1.0);
    }
    const double xdtm = delz[kk] / xum;
    const double xxx2 = (xh2o2[kk] - xgas2[kk]);
    const double yh2o2 = xxx2 != 0.0 ? xh2o2[kk] / xxx2 * xdtm : large_value_lifetime;
    tmp1[kk] = kmax(1.0 / yh2o2, 0.0) * stay;
    const double xxx3 = (xso2[kk] - xgas3[kk]);
    const double yso2 = xxx3 != 0.0 ? xso2[kk] / xxx3 * xdtm : large_value_lifetime;
    tmp2[kk] = kmax(1.0 / yso2, 0.0) * stay;
  }
  for (int kk = ktop; kk < nlev; ++kk) {
    if (rain[kk] <= 0.0) {
      for (int mm = 0; mm < 31; ++mm) het[(long)mm * nlev + kk] = 0.0;
      continue;
    }
    const double work1 = avo2 * xliq[kk], work2 = const0 * tfld[kk];
    het[1L * nlev + kk] = sh_calc_het_rates(satf_h2o2, rain[kk], xhen_h2o2[kk], tmp1[kk], work1,
                                            work2);
    het[3L * nlev + kk] = het[1L * nlev + kk]; /* SO2 takes the H2O2 rate */
    het[2L * nlev + kk] = sh_calc_het_rates(satf_hno3, rain[kk], xhen_hno3[kk], 0.0, work1,
                                            work2);
  }
  for (int kk = 0; kk < ktop + 1 && kk < nlev; ++kk)
    for (int mm = 0; mm < 3; ++mm) het[(long)wetdep_map[mm] * nlev + kk] = 0.0;
  free(w);
  free(tmp1);
  free(tmp2);
}

/* orc_sethet_column over [nlev][ncol] planes (column fastest); rlat, phis
 * [ncol]; het [31][nlev][ncol] */
void orc_sethet_domain(int ncol, int nlev, double delt, const double *rlat, const double *phis,
                       const double *press, const double *zmid, const double *tfld,
                       const double *cmfdqr, const double *nrain, const double *nevapr,
                       const double *inv_m, const double *qin_h2o2, const double *qin_so2,
                       double *het) {
  const long plane = (long)nlev * ncol;
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *b = (double *)malloc(sizeof(double) * (size_t)(nlev > 0 ? nlev : 1) * (9 + 31));
    double *in[9];
    for (int i = 0; i < 9; ++i) in[i] = b + (long)i * nlev;
    double *h = b + 9L * nlev;
    const double *src[9] = {press, zmid, tfld, cmfdqr, nrain, nevapr, inv_m, qin_h2o2, qin_so2};
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int j = 0; j < ncol; ++j) {
      for (int i = 0; i < 9; ++i)
        for (int k = 0; k < nlev; ++k) in[i][k] = src[i][(long)k * ncol + j];
      orc_sethet_column(nlev, rlat[j], phis[j], delt, in[0], in[1], in[2], in[3], in[4], in[5],
                        in[6], in[7], in[8], h);
      for (int m = 0; m < 31; ++m)
        for (int k = 0; k < nlev; ++k) het[m * plane + (long)k * ncol + j] = h[(long)m * nlev + k];
    }
    free(b);
  }
}

/* ====================== external forcing (setext) =========================
 * mo_setext::extfrc_set (mo_setext.hpp:22-54) for one column: each forced
 * species' rate is the sum of its sectors, read bottom-up where the data
 * are stored that way (file_alt_data).  sectors[mm][isec] -> [nlev];
 * frcing [9][nlev]. */
void orc_setext_column(int nlev, const int frc_ndx[9], const int file_alt_data[9],
                       const int nsectors[9], const double *const sectors[9][4],
                       double *frcing) {
  for (int kk = 0; kk < nlev; ++kk)
    for (int mm = 0; mm < 9; ++mm) {
      const int nn = frc_ndx[mm] - 1;
      frcing[(long)nn * nlev + kk] = 0.0;
      for (int isec = 0; isec < nsectors[mm]; ++isec) {
        if (file_alt_data[mm])
          frcing[(long)nn * nlev + kk] += sectors[mm][isec][nlev - 1 - kk];
        else
          frcing[(long)nn * nlev + kk] += sectors[mm][isec][kk];
      }
    }
}

/* ================ photolysis (mo_photo.hpp) ==============================
 * table_photo (mo_photo.hpp:798-893) for one column: cloud_mod (:169-427)
 * forms the effective albedo and the cloud multiplier, jlong (:647-796)
 * interpolates the radiative source function (interpolate_rsf, :498-645,
 * with find_index :428-450 and calc_sum_wght :452-496) and the cross
 * sections, and the rate of the one photolysis reaction (phtcnt = 1) is
 * photo = cld_mult esfact (photo + pht_alias_mult jlong[lng_index]).
 * Tables are row-major: rsf_tab [nw][nump][numsza][numcolo3][numalb],
 * xsqy [numj][nw][nt][np_xs].  Level 0 is the model top. */

/* mo_photo.hpp:428-450: the index before the first entry above vmin
 * (clamped to [0, len - 2]); idx0 when no entry is above vmin */
int orc_photo_find_index(const double *var, int len, double vmin, int idx0) {
  for (int ii = 0; ii < len; ++ii)
    if (var[ii] > vmin) {
      const int m = ii < len - 1 ? ii : len - 1;
      return m - 1 > 0 ? m - 1 : 0;
    }
  return idx0;
}

static double pht_bound(double lo, double hi, double x) { return kmax(lo, kmin(hi, x)); }

/* mo_photo.hpp:169-427, the column's scans in level order (the serial
 * loops its comments give) */
void orc_photo_cloud_mod(int nlev, double zen_angle, const double *clouds, const double *lwc,
                         const double *delp, double srf_alb, double *eff_alb,
                         double *cld_mult) {
  const double rgrav = 1.0 / 9.80616, f_lwp2tau = .155, tau_min = 5.0;
  const double C1 = 11.905, C2 = 9.524;
  double *w = (double *)malloc(sizeof(double) * 5 * (size_t)(nlev > 0 ? nlev : 1));
  double *del_tau = w, *above_tau = w + nlev, *above_cld = w + 2 * nlev,
         *below_tau = w + 3 * nlev, *below_cld = w + 4 * nlev;
  for (int k = 0; k < nlev; ++k) {
    if (clouds[k] != 0.0) {
      const double del_lwp = rgrav * lwc[k] * delp[k] * 1000.0 / clouds[k];
      del_tau[k] = del_lwp * f_lwp2tau * M_pow(clouds[k], 1.5);
    } else {
      del_tau[k] = 0.0;
    }
  }
  /* top down */
  above_tau[0] = 0.0;
  above_cld[0] = 0.0;
  for (int k = 0; k < nlev - 1; ++k) {
    above_tau[k + 1] = del_tau[k] + above_tau[k];
    above_cld[k + 1] = clouds[k] * del_tau[k] + above_cld[k];
  }
  for (int k = 1; k < nlev; ++k)
    if (above_tau[k] != 0.0) above_cld[k] /= above_tau[k];
  for (int k = 1; k < nlev; ++k)
    if (above_tau[k] == 0.0) above_cld[k] = above_cld[k - 1];
  /* bottom up */
  below_tau[nlev - 1] = 0.0;
  below_cld[nlev - 1] = 0.0;
  for (int k = nlev - 2; k >= 0; --k) {
    below_tau[k] = del_tau[k + 1] + below_tau[k + 1];
    below_cld[k] = clouds[k + 1] * del_tau[k + 1] + below_cld[k + 1];
  }
  for (int k = nlev - 2; k >= 0; --k)
    if (below_tau[k] != 0.0) below_cld[k] /= below_tau[k];
  for (int k = nlev - 2; k >= 0; --k)
    if (below_tau[k] == 0.0) below_cld[k] = below_cld[k + 1];
  /* optical depths of the cloudy part (jfm) */
  for (int k = 1; k < nlev; ++k)
    if (above_cld[k] != 0.0) above_tau[k] /= above_cld[k];
  for (int k = 1; k < nlev; ++k)
    if (above_tau[k] < tau_min) above_cld[k] = 0.0;
  for (int k = 0; k < nlev - 1; ++k)
    if (below_cld[k] != 0.0) below_tau[k] /= below_cld[k];
  for (int k = 0; k < nlev - 1; ++k)
    if (below_tau[k] < tau_min) below_cld[k] = 0.0;
  /* transmission factors */
  const double coschi = kmax(M_cos(zen_angle), 0.5);
  for (int k = 0; k < nlev; ++k) {
    const double below_tra = C1 / (C2 + below_tau[k]);
    eff_alb[k] = srf_alb + below_cld[k] * (1.0 - below_tra) * (1.0 - srf_alb);
    double del_lwp = 0.0;
    if (clouds[k] != 0.0) del_lwp = rgrav * lwc[k] * delp[k] * 1000.0 / clouds[k];
    double fac1 = 0.0;
    if (del_lwp * f_lwp2tau >= tau_min) fac1 = 1.4 * coschi - 1.0;
    const double above_tra = C1 / (C2 + above_tau[k]);
    const double fac2 = kmin(0.0, 1.6 * coschi * above_tra - 1.0);
    cld_mult[k] = kmax(.05, 1.0 + fac1 * clouds[k] + fac2 * above_cld[k]);
  }
  free(w);
}

/* mo_photo.hpp:452-496: the eight-corner weighted sum for one wavelength */
static double pht_sum_wght(const orc_photo_table *t, const double dels[3], double wrk0, int iz,
                           int is, int iv, int ial, int wn) {
  const long s_al = 1, s_v = t->numalb, s_s = (long)t->numcolo3 * s_v, s_z = (long)t->numsza * s_s,
             s_w = (long)t->nump * s_z;
  const double *r = t->rsf_tab + wn * s_w + iz * s_z;
  double wrk1 = (1.0 - dels[1]) * (1.0 - dels[2]);
  const double w000 = wrk0 * wrk1, w100 = dels[0] * wrk1;
  wrk1 = (1.0 - dels[1]) * dels[2];
  const double w001 = wrk0 * wrk1, w101 = dels[0] * wrk1;
  wrk1 = dels[1] * (1.0 - dels[2]);
  const double w010 = wrk0 * wrk1, w110 = dels[0] * wrk1;
  wrk1 = dels[1] * dels[2];
  const double w011 = wrk0 * wrk1, w111 = dels[0] * wrk1;
  const long a0 = is * s_s, a1 = (is + 1) * s_s, v0 = iv * s_v, v1 = (iv + 1) * s_v;
  const long l0 = ial * s_al, l1 = (ial + 1) * s_al;
  return w000 * r[a0 + v0 + l0] + w001 * r[a0 + v0 + l1] + w010 * r[a0 + v1 + l0] +
         w011 * r[a0 + v1 + l1] + w100 * r[a1 + v0 + l0] + w101 * r[a1 + v0 + l1] +
         w110 * r[a1 + v1 + l0] + w111 * r[a1 + v1 + l1];
}

/* mo_photo.hpp:498-645: rsf [nw][nlev] */
void orc_photo_interpolate_rsf(const orc_photo_table *t, int nlev, const double *alb_in,
                               double sza_in, const double *p_in, const double *colo3_in,
                               double *rsf) {
  const int is = orc_photo_find_index(t->sza, t->numsza, sza_in, 0);
  double dels[3];
  dels[0] = pht_bound(0.0, 1.0, (sza_in - t->sza[is]) * t->del_sza[is]);
  const double wrk0 = 1.0 - dels[0];
  for (int kk = 0; kk < nlev; ++kk) {
    const int albind = orc_photo_find_index(t->alb, t->numalb, alb_in[kk], 0);
    int pind;
    double wght1;
    if (p_in[kk] > t->press[0]) {
      pind = 1;
      wght1 = 1.0;
    } else if (p_in[kk] <= t->press[t->nump - 1]) {
      pind = t->nump - 1;
      wght1 = 0.0;
    } else {
      int iz;
      for (iz = 1; iz < t->nump; ++iz)
        if (t->press[iz] < p_in[kk]) break;
      iz = iz < t->nump - 1 ? iz : t->nump - 1;
      pind = iz > 1 ? iz : 1;
      wght1 = pht_bound(0.0, 1.0, (p_in[kk] - t->press[pind]) * t->del_p[pind - 1]);
    }
    const double v3ratu = colo3_in[kk] / t->colo3[pind - 1];
    const int ratindu = orc_photo_find_index(t->o3rat, t->numcolo3, v3ratu, 0);
    double v3ratl;
    int ratindl;
    if (t->colo3[pind] != 0.0) {
      v3ratl = colo3_in[kk] / t->colo3[pind];
      ratindl = orc_photo_find_index(t->o3rat, t->numcolo3, v3ratl, 0);
    } else {
      ratindl = ratindu;
      v3ratl = t->o3rat[ratindu];
    }
    const int ial = albind;
    dels[2] = pht_bound(0.0, 1.0, (alb_in[kk] - t->alb[ial]) * t->del_alb[ial]);
    const double d1l = pht_bound(0.0, 1.0, (v3ratl - t->o3rat[ratindl]) * t->del_o3rat[ratindl]);
    const double d1u = pht_bound(0.0, 1.0, (v3ratu - t->o3rat[ratindu]) * t->del_o3rat[ratindu]);
    for (int wn = 0; wn < t->nw; ++wn) {
      double d[3] = {dels[0], d1l, dels[2]};
      const double psum_l = pht_sum_wght(t, d, wrk0, pind, is, ratindl, ial, wn);
      d[1] = d1u;
      const double psum_u = pht_sum_wght(t, d, wrk0, pind - 1, is, ratindu, ial, wn);
      rsf[(long)wn * nlev + kk] = (psum_l + wght1 * (psum_u - psum_l)) * t->etfphot[wn];
    }
  }
}

/* mo_photo.hpp:647-796: j_long [numj][nlev] */
void orc_photo_jlong(const orc_photo_table *t, int nlev, double sza_in, const double *alb_in,
                     const double *p_in, const double *t_in, const double *colo3_in,
                     double *j_long) {
  double *rsf = (double *)malloc(sizeof(double) * (size_t)t->nw * (nlev > 0 ? nlev : 1));
  orc_photo_interpolate_rsf(t, nlev, alb_in, sza_in, p_in, colo3_in, rsf);
  const long s_t = t->np_xs, s_w = (long)t->nt * s_t, s_j = (long)t->nw * s_w;
  for (int kk = 0; kk < nlev; ++kk) {
    const int t_index = (int)(kmin(201.0, kmax(t_in[kk] - 148.5, 1.0)) - 1.0);
    const double ptarget = p_in[kk];
    for (int i = 0; i < t->numj; ++i) {
      double suma = 0.0;
      for (int wn = 0; wn < t->nw; ++wn) {
        const double *x = t->xsqy + i * s_j + wn * s_w + t_index * s_t;
        double xswk;
        if (ptarget >= t->prs[0]) {
          xswk = x[0];
        } else if (ptarget <= t->prs[t->np_xs - 1]) {
          xswk = x[t->np_xs - 1];
        } else {
          double delp = 0.0;
          int pndx = 0;
          for (int km = 1; km < t->np_xs; ++km)
            if (ptarget >= t->prs[km]) {
              pndx = km - 1;
              delp = (t->prs[pndx] - ptarget) * t->dprs[pndx];
              break;
            }
          xswk = x[pndx] + delp * (x[pndx + 1] - x[pndx]);
        }
        suma += xswk * rsf[(long)wn * nlev + kk];
      }
      j_long[(long)i * nlev + kk] = suma;
    }
  }
  free(rsf);
}

/* mo_photo.hpp:798-893 for one column; photo [nlev] in/out (unchanged at
 * night, sza >= 88.85 degrees, and when lng_index < 0) */
void orc_table_photo(const orc_photo_table *t, int nlev, double *photo, const double *pmid,
                     const double *pdel, const double *temper, const double *colo3_in,
                     double zen_angle, double srf_alb, const double *lwc, const double *clouds,
                     double esfact) {
  const double Pa2mb = 1.e-2, r2d = 180.0 / PI, max_zen_angle = 88.85;
  const double sza_in = zen_angle * r2d;
  if (!(sza_in >= 0.0 && sza_in < max_zen_angle)) return;
  const size_t n = (size_t)(nlev > 0 ? nlev : 1);
  double *w = (double *)malloc(sizeof(double) * n * (3 + (size_t)t->numj));
  double *eff_alb = w, *cld_mult = w + n, *parg = w + 2 * n, *jl = w + 3 * n;
  orc_photo_cloud_mod(nlev, zen_angle, clouds, lwc, pdel, srf_alb, eff_alb, cld_mult);
  for (int k = 0; k < nlev; ++k) parg[k] = pmid[k] * Pa2mb;
  orc_photo_jlong(t, nlev, sza_in, eff_alb, parg, temper, colo3_in, jl);
  const int ind = t->lng_index;
  if (ind > -1)
    for (int k = 0; k < nlev; ++k)
      photo[k] = cld_mult[k] * esfact * (photo[k] + t->pht_alias_mult * jl[(long)ind * nlev + k]);
  free(w);
}

/* orc_table_photo over a domain of [nlev][ncol] planes (column fastest);
 * per column zen_angle, srf_alb.  OpenMP over columns. */
void orc_table_photo_domain(const orc_photo_table *t, int ncol, int nlev, double *photo,
                            const double *pmid, const double *pdel, const double *temper,
                            const double *colo3_in, const double *zen_angle,
                            const double *srf_alb, const double *lwc, const double *clouds,
                            double esfact) {
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *buf = (double *)malloc(sizeof(double) * (size_t)(nlev > 0 ? nlev : 1) * 7);
    double *ph = buf, *pm = ph + nlev, *pd = pm + nlev, *te = pd + nlev, *co = te + nlev,
           *lw = co + nlev, *cl = lw + nlev;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int j = 0; j < ncol; ++j) {
      for (int k = 0; k < nlev; ++k) {
        const long i = (long)k * ncol + j;
        ph[k] = photo[i];
        pm[k] = pmid[i];
        pd[k] = pdel[i];
        te[k] = temper[i];
        co[k] = colo3_in[i];
        lw[k] = lwc[i];
        cl[k] = clouds[i];
      }
      orc_table_photo(t, nlev, ph, pm, pd, te, co, zen_angle[j], srf_alb[j], lw, cl, esfact);
      for (int k = 0; k < nlev; ++k) photo[(long)k * ncol + j] = ph[k];
    }
    free(buf);
  }
}
