// Constants of the MAM4 microphysics chain, shared by the host set-up code and
// the gfx950 kernels.
//
// Two kinds of constants live here:
//  * literal tables (mode diameters and widths, species densities, the
//    gas_pcnst maps) -- constexpr, folded into the kernels;
//  * derived constants that the reference recomputes with libm on every call
//    (log(sigma), lognormal moment factors, coagulation e^{k ln^2 sigma}
//    powers, rename cut-offs, calcsize num2vol limits, ...).  They depend only
//    on the mode table, so the host evaluates them once per handle
//    (chain_consts_init) and the kernels read them from a device buffer.
#pragma once

// Plain C++ compilers see the device annotations as no-ops, so the test
// suite can compile the per-cell code for the CPU (tests/cpp/cell_host.cpp).
#if !defined(__HIP__)
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

constexpr int NMODE = 4;
constexpr int NAER = 7;
constexpr int NSTATE = 52;
enum { NACC = 0, NAIT = 1, NCOA = 2, NPCA = 3 };
// AeroId order (aero_modes.hpp:126-135)
enum { A_SOA = 0, A_SO4 = 1, A_POM = 2, A_BC = 3, A_NCL = 4, A_DST = 5, A_MOM = 6 };

constexpr double PI = 3.14159265358979323846264;  // mam4_constants.hpp:29
constexpr double MW_H2O = 0.018016;                // mam4_constants.hpp:51
constexpr double MW_DRY_AIR = 0.028966;            // mam4_constants.hpp:60
constexpr double MWAIR_G = 1000.0 * MW_DRY_AIR;    // 28.966 g/mol

// aero_modes.hpp:78-93
constexpr double MODE_DMIN[4] = {5.35e-8, 8.7e-9, 1e-6, 1e-8};
constexpr double MODE_DNOM[4] = {1.1e-7, 2.6e-8, 2e-6, 5e-8};
constexpr double MODE_DMAX[4] = {4.4e-7, 5.2e-8, 4e-6, 1e-7};
constexpr double MODE_SIGMA[4] = {1.8, 1.6, 1.8, 1.60000002384186};
// aero_modes.hpp:154-161 (AeroId order)
constexpr double AER_DENS[7] = {1000.0, 1770.0, 1000.0, 1700.0, 1900.0, 2600.0, 1601.0};
constexpr int NSPEC_MODE[4] = {7, 4, 7, 3};
constexpr int MODE_OFF[4] = {0, 7, 11, 18};

// species (AeroId) of the s-th species of mode m (aero_modes.hpp:214-243)
__host__ __device__ constexpr int mode_spec(int m, int s) {
  return m == 0 ? (s == 0 ? A_SO4 : s == 1 ? A_POM : s == 2 ? A_SOA : s == 3 ? A_BC
                   : s == 4 ? A_DST : s == 5 ? A_NCL : A_MOM)
       : m == 1 ? (s == 0 ? A_SO4 : s == 1 ? A_SOA : s == 2 ? A_NCL : A_MOM)
       : m == 2 ? (s == 0 ? A_DST : s == 1 ? A_NCL : s == 2 ? A_SO4 : s == 3 ? A_BC
                   : s == 4 ? A_POM : s == 5 ? A_SOA : A_MOM)
                : (s == 0 ? A_POM : s == 1 ? A_BC : A_MOM);
}
// index of AeroId a inside mode m, or -1
__host__ __device__ constexpr int spec_in_mode(int a, int m) {
  return mode_spec(m, 0) == a ? 0
       : (NSPEC_MODE[m] > 1 && mode_spec(m, 1) == a) ? 1
       : (NSPEC_MODE[m] > 2 && mode_spec(m, 2) == a) ? 2
       : (NSPEC_MODE[m] > 3 && mode_spec(m, 3) == a) ? 3
       : (NSPEC_MODE[m] > 4 && mode_spec(m, 4) == a) ? 4
       : (NSPEC_MODE[m] > 5 && mode_spec(m, 5) == a) ? 5
       : (NSPEC_MODE[m] > 6 && mode_spec(m, 6) == a) ? 6 : -1;
}
// aerosol species the mode carries as a prognostic
__host__ __device__ constexpr bool valid(int a, int m) { return spec_in_mode(a, m) >= 0; }
// species that can be non-zero inside a clear subarea: the prognostic ones,
// plus SO4/SOA condensed onto and SO4/SOA/NaCl coagulated into the
// primary-carbon mode before aging moves them to accumulation.
__host__ __device__ constexpr bool live(int a, int m) {
  return valid(a, m) || (m == NPCA && (a == A_SO4 || a == A_SOA || a == A_NCL));
}

// gas_pcnst slot of the state fields (mam4_amicphys.hpp:35-63)
constexpr int NUMPTR[4] = {13, 18, 26, 30};
// E3SM adv_mass (gas_chem_mechanism.hpp:31-37)
constexpr double ADV_MASS[31] = {
    47.998200,  34.013600, 98.078400,     64.064800, 62.132400, 12.011000,
    115.107340, 12.011000, 12.011000,     12.011000, 135.064039, 58.442468,
    250092.672000, 1.007400, 115.107340,  12.011000, 58.442468, 250092.672000,
    1.007400,   135.064039, 58.442468,    115.107340, 12.011000, 12.011000,
    12.011000,  250092.672000, 1.007400,  12.011000, 12.011000, 250092.672000,
    1.007400};
__host__ __device__ constexpr int mass_slot(int m, int s) {
  return (m == 0 ? 6 : m == 1 ? 14 : m == 2 ? 19 : 27) + s;
}
// kg/kg -> mol/mol of species s of mode m, and back
__host__ __device__ constexpr double to_vmr_mass(int m, int s) {
  return MWAIR_G / ADV_MASS[mass_slot(m, s)];
}
__host__ __device__ constexpr double to_mmr_mass(int m, int s) {
  return ADV_MASS[mass_slot(m, s)] / MWAIR_G;
}
constexpr double TO_VMR_NUM = MWAIR_G;
constexpr double TO_MMR_NUM = 1.0 / MWAIR_G;
constexpr double TO_VMR_H2SO4 = MWAIR_G / 98.078400;
constexpr double TO_MMR_H2SO4 = 98.078400 / MWAIR_G;
constexpr double TO_VMR_SOAG = MWAIR_G / 12.011000;
constexpr double TO_MMR_SOAG = 12.011000 / MWAIR_G;

// mam4_amicphys.hpp:131-135, 141-154
__host__ __device__ constexpr double fcvt_aer(int a) {
  return (a == A_SOA || a == A_POM) ? 8.000000000000000E-002 : 1.0;
}
constexpr double FCVT_SOAG = 12.011000 / 150.0;

// mass_2_vol [m3/kmol] (mam4_amicphys.hpp:1512-1518), AeroId order
__host__ __device__ constexpr double mass_2_vol(int a) {
  return a == 0 ? 0.15 : a == 1 ? 6.4971751412429377e-002 : a == 2 ? 0.15
       : a == 3 ? 7.0588235294117650e-003 : a == 4 ? 3.0789473684210526e-002
       : a == 5 ? 5.1923076923076926e-002 : 156.20986883198000;
}

// state field indices (include/mam4x.h)
__host__ __device__ constexpr int f_ni(int m) { return m; }
__host__ __device__ constexpr int f_nc(int m) { return 4 + m; }
__host__ __device__ constexpr int f_qi(int m, int s) { return 8 + MODE_OFF[m] + s; }
__host__ __device__ constexpr int f_qc(int m, int s) { return 29 + MODE_OFF[m] + s; }
constexpr int F_NUM_C = 4;
constexpr int F_Q_I = 8;
constexpr int F_Q_C = 29;
constexpr int F_H2SO4 = 50;
constexpr int F_SOAG = 51;
// mode and mode-local species of a species-mass field (interstitial or
// cloud-borne)
__host__ __device__ constexpr int field_mode(int f) {
  return (f - (f >= F_Q_C ? F_Q_C : F_Q_I)) < MODE_OFF[1]   ? 0
       : (f - (f >= F_Q_C ? F_Q_C : F_Q_I)) < MODE_OFF[2] ? 1
       : (f - (f >= F_Q_C ? F_Q_C : F_Q_I)) < MODE_OFF[3] ? 2 : 3;
}
__host__ __device__ constexpr int field_spec(int f) {
  return f - (f >= F_Q_C ? F_Q_C : F_Q_I) - MODE_OFF[field_mode(f)];
}

// Derived constants, evaluated once on the host with the same libm calls the
// reference makes per cell.
struct ChainConsts {
  double alnsg[4];         // log(sigma)
  double dgn_from_vol[4];  // exp(-1.5 ln^2 sigma)          (conversions.hpp:302)
  double vol_from_dgn[4];  // exp(4.5 ln^2 sigma)           (conversions.hpp:315)
  // gas diffusivity of H2SO4 (gasaerexch.hpp:222-236)
  double h2so4_sqrt_mw;    // sqrt(1/mw + 0.035)
  double h2so4_vd_term;    // pow(pow(vd, 1/3) + 2.7189, 2)
  double tworootpi;        // 2 sqrt(pi)
  double root2;            // sqrt(2)
  // nucleation (mam4_amicphys.hpp:991-1003)
  double dplom, dphim, mass1p_aitlo, mass1p_aithi;
  double ln_1e38;          // log(1e38)  (nucleation.hpp:155)
  double ln_1em38;         // log(1e-38) (nucleation.hpp:548)
  // coagulation, per pair {ait->acc, pca->acc, ait->pca}: the sigma-only
  // factors of getcoags (coagulation.hpp:650-748)
  double dlgsqt2;          // 1/log(sqrt(2))
  struct CoagPair {
    double esat01, esat04, esat05, esat08, esat09, esat16, esat20, esat25,
        esat36, esat49, esat64, esat100;
    double esac01, esac04, esac05, esac08, esac09, esac16, esac20, esac25;
    double bm0ij[10];      // bm0ij[n1][n2n][n2a] for the pair's n2n, n2a
    double bm3i[10];
    double bm0_at, bm0_ac; // intramodal bm0[n2n], bm0[n2a]
    double dumatk_fac;     // exp(4.5 ln^2 sigma_src)  (coagulation.hpp:814)
  } pair[3];
  // aging (aging.hpp:99-151)
  double so4_vol, fac_m2v_eqvhyg, fac_volsfc, xferfrac_max;
  // rename, Aitken -> accumulation (rename.hpp:392-474)
  double rn_tail_fac, rn_v2n_lo_rlx, rn_v2n_hi_rlx, rn_ln_tail_fac, rn_ln_cutoff,
      rn_threshold;
  // calcsize (calcsize.hpp:893-960)
  double v2n_nom[4], v2n_min[4], v2n_max[4];
  double inv_density[4][7];
  double v2n_geomean;      // sqrt(v2n_nom_ait) * sqrt(v2n_nom_acc)
  // SOA (gasaerexch_soaexch.hpp:146-162): p0 = 1e-10 exp(-c1 (1/T - 1/298))
  double soa_c1;           // delh_vap_soa / rgas
  double soa_inv298;       // 1/298
};

// Fill the derived constants with host libm.  Defined in chain_consts.cpp.
void chain_consts_init(ChainConsts *c);

}  // namespace mam4x
