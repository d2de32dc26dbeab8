// Per-cell MAM4 microphysics chain for gfx950 (one lane = one cell).
//
// The reference (eagles-project/mam4xx, mam4_amicphys.hpp) carries each
// subarea through ~8 KB of [31][3] and [31][4][3] scratch arrays per cell.
// Here a subarea is the 30 values that can actually change -- 2 gases, 4
// numbers and the 24 (species, mode) masses that are prognostic or that
// condensation/coagulation can put into the primary-carbon mode before aging
// empties it (chain_consts.hpp: live()) -- so the whole chain stays in
// registers.  Per-process increment bookkeeping the reference keeps only for
// its optional history diagnostics is dropped; the few increments that feed
// back into the physics (Aitken growth for renaming, primary-carbon
// condensation/coagulation increments for aging) are kept.
//
// min/max are Kokkos' ternaries, not fmin/fmax (NaN semantics differ).
#pragma once

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#else
#include <cmath>
#endif

#include "chain_consts.hpp"

namespace mam4x {
namespace dev {

__device__ __forceinline__ double kmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double kmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double sq(double x) { return x * x; }
__device__ __forceinline__ double cube(double x) { return x * x * x; }
__device__ __forceinline__ double mmb(double lo, double hi, double x) {  // utils.hpp:23
  return kmax(lo, kmin(hi, x));
}

// Aerosol of one subarea in chain units: numbers [#/kmol], masses [mol/mol]
// indexed [AeroId][mode]; only live() entries are ever read or written.
struct Aer {
  double n[4];
  double q[7][4];
};

struct Cfg {
  int do_cond, do_rename, do_newnuc, do_coag;
  int gaexch_h2so4_uptake_optaa;
  unsigned nmono;
};

// ------------------------------------------------------------------ uptake
// gas_aer_uptkrates_1box1gas for H2SO4 (gasaerexch.hpp:267-350, accom 0.65,
// beta from the Knudsen number) times number*aircon (gasaerexch.hpp:427-435).
// The SOAG rates are 0.81 x these (:458-465), so SOAG's own transport
// properties, which the reference computes and discards, are not evaluated.
__device__ __forceinline__ void h2so4_uptake(const ChainConsts &K, double temp,
                                             double pmid, double aircon,
                                             const double dgn_awet[4],
                                             const double qnum[4], double uptk[4]) {
  const double p_in_atm = pmid / 1.013e5;
  const double dgas = (1.0e-3 * pow(temp, 1.75) * K.h2so4_sqrt_mw) /
                      (p_in_atm * K.h2so4_vd_term);
  const double gasdiffus = dgas * 1.0e-4;
  const double speed = 145.5 * sqrt(temp / 9.807840000000000E+01);
  const double freepath = 3.0 * gasdiffus / speed;
  const double a283 = 6.5000E-01 * 0.283, a75 = 6.5000E-01 * 0.75;
  const double xghq0 = 7.071067690849304E-01, wghq = 8.862269520759583E-01;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const double lnsg = K.alnsg[n];
    const double lndpgn = log(dgn_awet[n]);
    const double kn0 = 2.0 * freepath / dgn_awet[n];
    const double tmpa = 1.0 / (1.0 + kn0) -
                        (2.0 * kn0 + 1.0 + a283) / (kn0 * (kn0 + 1.0 + a283) + a75);
    double beta = 1.0 - kn0 * tmpa;
    beta = kmax(1.0, kmin(2.0, beta));
    const double constant = K.tworootpi * exp(beta * lndpgn + 0.5 * sq(beta * lnsg));
    double sumghq = 0.0;
#pragma unroll
    for (int iq = 0; iq < 2; ++iq) {
      const double x = iq == 0 ? xghq0 : -xghq0;
      const double lndp = lndpgn + beta * lnsg * lnsg + K.root2 * lnsg * x;
      const double D_p = exp(lndp);
      const double kn = 2.0 * freepath / D_p;
      const double hh = (a75 * (1.0 + kn)) / (kn * (kn + 1.0 + a283) + a75);
      sumghq += wghq * D_p * hh / pow(D_p, beta);
    }
    uptk[n] = (constant * gasdiffus * sumghq) * (qnum[n] * aircon);
  }
}

// ------------------------------------------------------------------ SOA
// mam_soaexch_1subarea (gasaerexch_soaexch.hpp:79-341), one SOA species, as
// a resumable state machine: soa_tick advances one adaptive sub-step, so the
// persistent work-queue kernel can hand a lane a new cell as soon as its
// current one has converged.
struct SoaIn {        // inputs of the SOA loop of one clear subarea
  double g;           // SOAG [mol/mol, fcvt applied]
  double a[4];        // SOA in the four modes
  double uptk[4];     // H2SO4 uptake rates [1/s]; SOAG's are 0.81x
  double aopoa[4];    // oxygenated POM (0.1 x POM, 0 in the pcarbon mode)
  double g0;          // equilibrium SOAG mixing ratio
};

struct SoaState {
  double g, a[4], uk[4], aopoa[4], g0;
  double tcur, avg, dtsum;
  int niter;
  bool skip[4];
};

__device__ __forceinline__ double soa_g0(const ChainConsts &K, double temp, double pmid) {
  const double p0_soa = 1.0e-10 * exp(-(K.soa_c1) * ((1.0 / temp) - K.soa_inv298));
  return 1.01325e5 * p0_soa / pmid;
}

__device__ __forceinline__ void soa_init(SoaState &s, const SoaIn &in) {
  s.g = in.g;
  s.g0 = in.g0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const double uk = in.uptk[n] * 0.81;  // gasaerexch.hpp:458-465
    s.skip[n] = !(uk > 1.0e-15);
    s.uk[n] = s.skip[n] ? 0.0 : uk;
    s.a[n] = in.a[n];
    s.aopoa[n] = in.aopoa[n];
  }
  s.tcur = 0.0;
  s.avg = 0.0;
  s.dtsum = 0.0;
  s.niter = 0;
}

// rate of the first sub-step times dtfull / alpha: the sub-step count the loop
// would need at its initial rate (used to order the work queue longest-first)
__device__ __forceinline__ double soa_predict(const SoaState &s, double dtfull) {
  const double g_soa = kmax(s.g, 0.0);
  double tmpb = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    if (s.skip[n]) continue;
    const double a_soa = kmax(s.a[n], 0.0);
    const double sat = s.g0 / kmax(s.aopoa[n] + a_soa, 1.0e-20);
    const double gstar = sat * a_soa;
    tmpb += s.uk[n] * fabs((g_soa - gstar) / kmax(g_soa, kmax(gstar, 1.0e-20)));
  }
  return dtfull * tmpb / 0.05;
}

// One pass of the while loop (gasaerexch_soaexch.hpp:170-335).  Returns false
// once the loop has ended (time reached or 1000 sub-steps taken).
//
// Written without branches on the mode: a skipped mode (uptake <= 1e-15) has
// uk = 0 and a_soa = 0, so its terms add exactly +0 to every sum and its
// denominators are exactly 1 -- the same values the reference's skip produces
// -- and the divisions of all four modes issue as straight-line code.
__device__ __forceinline__ bool soa_tick(SoaState &s, double dtfull) {
  const double alpha_astem = 0.05, a_min1 = 1.0e-20, g_min1 = 1.0e-20;
  if (!(s.tcur < dtfull - 1.0e-3)) return false;
  s.niter++;
  if (s.niter > 1000) return false;
  const double gprv = s.g;
  double g_soa = kmax(gprv, 0.0);
  double tot = g_soa;
  double a_soa[4], sat[4], gstar[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    a_soa[n] = s.skip[n] ? 0.0 : kmax(s.a[n], 0.0);
    tot = s.skip[n] ? tot : tot + a_soa[n];
  }
  double tmpb = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    sat[n] = s.g0 / kmax(s.aopoa[n] + a_soa[n], a_min1);
    gstar[n] = sat[n] * a_soa[n];
    const double phi = (g_soa - gstar[n]) / kmax(g_soa, kmax(gstar[n], g_min1));
    tmpb = s.skip[n] ? tmpb : tmpb + s.uk[n] * fabs(phi);
  }
  const double tmpa = kmax(0.0, tmpb);
  const double dtmax = dtfull - s.tcur;
  const bool last = dtmax * tmpa <= alpha_astem;
  const double dtcur = last ? dtmax : alpha_astem / tmpa;
  s.tcur = last ? dtfull : s.tcur + dtcur;
  // step 1: explicit estimate of condensing modes' a_soa and sat
  double beta[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    beta[n] = dtcur * s.uk[n];
    const double del = g_soa - gstar[n];
    const double at = a_soa[n] + beta[n] * del;
    const double sat1 = s.g0 / kmax(s.aopoa[n] + at, a_min1);
    sat[n] = del > 0.0 ? sat1 : sat[n];
  }
  // step 2: implicit in g, semi-implicit in a
  double sa = 0.0, sb = 0.0, den[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    den[n] = 1.0 + beta[n] * sat[n];
    sa = s.skip[n] ? sa : sa + a_soa[n] / den[n];
    sb = s.skip[n] ? sb : sb + beta[n] / den[n];
  }
  g_soa = (tot - sa) / (1.0 + sb);
  g_soa = kmax(0.0, g_soa);
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    // skipped modes: a_soa is never set, so their SOA becomes 0 (quirk of
    // gasaerexch_soaexch.hpp:327-331)
    s.a[n] = s.skip[n] ? 0.0 : (a_soa[n] + beta[n] * g_soa) / den[n];
  }
  s.g = g_soa;
  s.avg += dtcur * (gprv + 0.5 * (g_soa - gprv));
  s.dtsum += dtcur;
  return true;
}

// qgas_avg = max(0, sum(q dt) / sum(dt)) (gasaerexch_soaexch.hpp:337-340)
__device__ __forceinline__ double soa_gavg(const SoaState &s) {
  return kmax(0.0, s.avg / s.dtsum);
}

// ------------------------------------------------------------------ H2SO4
// Analytic non-volatile update (gasaerexch.hpp:479-537).
__device__ __forceinline__ void h2so4_update(double dt, double netprod,
                                             const double uptk[4], double &qgas,
                                             double &qgas_avg, double qso4[4]) {
  double tmpa = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) tmpa += uptk[n];
  const double kxt = tmpa * dt;
  const double pxt = netprod * dt;
  const double q1 = qgas;
  if (kxt >= 1.0e-20) {
    double q3, q4;
    if (kxt > 0.001) {
      const double pok = pxt / kxt;
      const double e = exp(-kxt);
      q3 = (q1 - pok) * e + pok;
      q4 = (q1 - pok) * (1.0 - e) / kxt + pok;
    } else {
      const double k2 = kxt * kxt;
      q3 = q1 * (1.0 - kxt + k2 * 0.5) + pxt * (1.0 - kxt * 0.5 + k2 / 6.0);
      q4 = q1 * (1.0 - kxt * 0.5 + k2 / 6.0) + pxt * (0.5 - kxt / 6.0 + k2 / 24.0);
    }
    qgas = q3;
    qgas_avg = q4;
    const double del = (q1 + pxt) - q3;
#pragma unroll
    for (int n = 0; n < 4; ++n)
      if (!(uptk[n] <= 0.0)) qso4[n] = qso4[n] + del * (uptk[n] / tmpa);
  } else {
    qgas = q1 + pxt;
    qgas_avg = q1 + pxt * 0.5;
  }
}

// ------------------------------------------------------------------ rename
// Aitken -> accumulation renaming of one subarea (rename.hpp:22-389,
// 710-786).  gi/gc: Aitken species growth (AeroId order) of the interstitial
// and cloud-borne aerosol.
__device__ __forceinline__ void rename_ait_to_acc(const ChainConsts &K, bool cloudy,
                                                  Aer &ai, const double gi[7], Aer &ac,
                                                  const double gc[7]) {
  const double smallest = 1.0e-25;
  double vi = 0.0, dvi = 0.0, vc = 0.0, dvc = 0.0;
#pragma unroll
  for (int a = 0; a < 7; ++a) {
    if (!live(a, NAIT)) continue;
    vi += ai.q[a][NAIT] * mass_2_vol(a);
    dvi += gi[a] * mass_2_vol(a);
  }
  if (cloudy) {
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      if (!live(a, NAIT)) continue;
      vc += ac.q[a][NAIT] * mass_2_vol(a);
      dvc += gc[a] * mass_2_vol(a);
    }
  }
  double b4_dryvol = vi - dvi;
  if (cloudy) b4_dryvol = b4_dryvol + (vc - dvc);
  const double b4_bnd = kmax(b4_dryvol, smallest);
  double b4_qnum = ai.n[NAIT];
  if (cloudy) b4_qnum = b4_qnum + ac.n[NAIT];
  b4_qnum = kmax(0.0, b4_qnum);
  const double b4_qnum_bnd = mmb(b4_bnd * K.rn_v2n_hi_rlx, b4_bnd * K.rn_v2n_lo_rlx, b4_qnum);
  double dv = dvi;
  if (cloudy) dv = dv + dvc;
  const double aft_dryvol = b4_dryvol + dv;
  if (aft_dryvol <= smallest) return;
  const double pio6 = PI / 6;
  double b4_diam = cbrt((b4_bnd / b4_qnum_bnd) / pio6) * K.dgn_from_vol[NAIT];
  if (b4_diam > K.rn_threshold) {
    b4_dryvol = b4_dryvol * cube(K.rn_threshold / b4_diam);
    b4_diam = K.rn_threshold;
  }
  if ((aft_dryvol - b4_dryvol) <= 1.0e-6 * b4_bnd) return;
  const double aft_diam = cbrt((aft_dryvol / b4_qnum_bnd) / pio6) * K.dgn_from_vol[NAIT];
  if (aft_diam <= MODE_DNOM[NAIT]) return;
  const double lb4 = log(b4_diam), laft = log(aft_diam);
  const double b4_tail_num = 0.5 * erfc((K.rn_ln_cutoff - lb4) * K.rn_tail_fac);
  const double b4_tail_vol =
      0.5 * erfc((K.rn_ln_cutoff - (lb4 + K.rn_ln_tail_fac)) * K.rn_tail_fac);
  const double aft_tail_num = 0.5 * erfc((K.rn_ln_cutoff - laft) * K.rn_tail_fac);
  const double aft_tail_vol =
      0.5 * erfc((K.rn_ln_cutoff - (laft + K.rn_ln_tail_fac)) * K.rn_tail_fac);
  const double vfrac = aft_tail_vol * aft_dryvol - b4_tail_vol * b4_dryvol;
  if (vfrac <= 0.0) return;
  double xv = kmin(vfrac, aft_dryvol) / aft_dryvol;
  xv = kmin(xv, K.xferfrac_max);
  double xn = aft_tail_num - b4_tail_num;
  xn = kmax(0.0, kmin(xn, xv));
  {
    const double nt = ai.n[NAIT] * xn;
    ai.n[NAIT] -= nt;
    ai.n[NACC] += nt;
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      if (!live(a, NAIT)) continue;
      const double vt = ai.q[a][NAIT] * xv;
      ai.q[a][NAIT] -= vt;
      ai.q[a][NACC] += vt;
    }
  }
  if (cloudy) {
    const double nt = ac.n[NAIT] * xn;
    ac.n[NAIT] -= nt;
    ac.n[NACC] += nt;
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      if (!live(a, NAIT)) continue;
      const double vt = ac.q[a][NAIT] * xv;
      ac.q[a][NAIT] -= vt;
      ac.q[a][NACC] += vt;
    }
  }
}

// ------------------------------------------------------------------ newnuc
// Vehkamaki et al. (2002) eqs 11-14 (vehkamaki2002.hpp:53-179).
__device__ __forceinline__ void vehk2002(double c, double t, double rh, double &rateloge,
                                         double &cnum_tot, double &xcrit) {
  const double lnN = log(c), lnrh = log(rh);
  const double t2 = sq(t), t3 = cube(t);
  const double r2 = sq(lnrh), r3 = cube(lnrh), N2 = sq(lnN), N3 = cube(lnN);
  const double x = 0.740997 - 0.00266379 * t - 0.00349998 * lnN + 0.0000504022 * t * lnN +
                   0.00201048 * lnrh - 0.000183289 * t * lnrh + 0.00157407 * r2 -
                   0.0000179059 * t * r2 + 0.000184403 * r3 - 1.50345e-6 * t * r3;
  const double a = 0.14309 + 2.21956 * t - 0.0273911 * t2 + 0.0000722811 * t3 + 5.91822 / x;
  const double b = 0.117489 + 0.462532 * t - 0.0118059 * t2 + 0.0000404196 * t3 + 15.7963 / x;
  const double cc = -0.215554 - 0.0810269 * t + 0.00143581 * t2 - 4.7758e-6 * t3 - 2.91297 / x;
  const double d = -3.58856 + 0.049508 * t - 0.00021382 * t2 + 3.10801e-7 * t3 - 0.0293333 / x;
  const double e = 1.14598 - 0.600796 * t + 0.00864245 * t2 - 0.0000228947 * t3 - 8.44985 / x;
  const double f = 2.15855 + 0.0808121 * t - 0.000407382 * t2 - 4.01957e-7 * t3 + 0.721326 / x;
  const double g = 1.6241 - 0.0160106 * t + 0.0000377124 * t2 + 3.21794e-8 * t3 - 0.0113255 / x;
  const double h = 9.71682 - 0.115048 * t + 0.000157098 * t2 + 4.00914e-7 * t3 + 0.71186 / x;
  const double i = -1.05611 + 0.00903378 * t - 0.0000198417 * t2 + 2.46048e-8 * t3 - 0.0579087 / x;
  const double j = -0.148712 + 0.00283508 * t - 9.24619e-6 * t2 + 5.00427e-9 * t3 - 0.0127081 / x;
  rateloge = a + b * lnrh + cc * r2 + d * r3 + e * lnN + f * lnrh * lnN + g * r2 * lnN +
             h * N2 + i * lnrh * N2 + j * N3;
  const double A = -0.00295413 - 0.0976834 * t + 0.00102485 * t2 - 2.18646e-6 * t3 - 0.101717 / x;
  const double B = -0.00205064 - 0.00758504 * t + 0.000192654 * t2 - 6.7043e-7 * t3 - 0.255774 / x;
  const double C = +0.00322308 + 0.000852637 * t - 0.0000154757 * t2 + 5.66661e-8 * t3 + 0.0338444 / x;
  const double D = +0.0474323 - 0.000625104 * t + 2.65066e-6 * t2 - 3.67471e-9 * t3 - 0.000267251 / x;
  const double E = -0.0125211 + 0.00580655 * t - 0.000101674 * t2 + 2.88195e-7 * t3 + 0.0942243 / x;
  const double F = -0.038546 - 0.000672316 * t + 2.60288e-6 * t2 + 1.19416e-8 * t3 - 0.00851515 / x;
  const double G = -0.0183749 + 0.000172072 * t - 3.71766e-7 * t2 - 5.14875e-10 * t3 + 0.00026866 / x;
  const double H = -0.0619974 + 0.000906958 * t - 9.11728e-7 * t2 - 5.36796e-9 * t3 - 0.00774234 / x;
  const double I = +0.0121827 - 0.00010665 * t + 2.5346e-7 * t2 - 3.63519e-10 * t3 + 0.000610065 / x;
  const double J = +0.000320184 - 0.0000174762 * t + 6.06504e-8 * t2 - 1.4177e-11 * t3 + 0.000135751 / x;
  cnum_tot = exp(A + B * lnrh + C * r2 + D * r3 + E * lnN + F * lnrh * lnN + G * r2 * lnN +
                 H * N2 + I * lnrh * N2 + J * N3);
  xcrit = x;
}

// mam_newnuc_1subarea (mam4_amicphys.hpp:916-1095) with
// mer07_veh02_nuc_mosaic_1box (nucleation.hpp:346-809) inlined for the path
// the chain takes: method 11, no NH3, one size bin [dplom, dphim].
__device__ __forceinline__ void newnuc(const ChainConsts &K, double deltat, double temp,
                                       double pmid, double zmid, double pblh, double relhum,
                                       double uptkrate_h2so4, double &h2so4,
                                       double h2so4_avg, double &num_ait, double &so4_ait) {
  if (h2so4_avg <= 4.0e-16) return;
  const double relhumnn = kmax(0.01, kmin(0.99, relhum));
  const double pi = PI;
  const double rgas = 8.31446759100000, avogad = 6.022140000000000E+023;
  const double mw_so4a = 96, mw_so4a_host = 115;
  const double qh2so4_cur = h2so4;
  // --- mer07_veh02_nuc_mosaic_1box ---
  const double cair = pmid / (temp * rgas);
  const double so4vol_in = h2so4_avg * cair * avogad * 1.0e-6;
  double rateloge = K.ln_1em38;
  double cnum_h2so4 = 0.0, cnum_tot = 0.0, radius_cluster = 0.0;
  if (so4vol_in >= 1.0e4) {
    const double temp_bb = kmax(230.15, kmin(305.15, temp));
    const double rh_bb = kmax(1.0e-4, kmin(1.0, relhumnn));
    const double so4vol_bb = kmax(1.0e4, kmin(1.0e11, so4vol_in));
    double x;
    vehk2002(so4vol_bb, temp_bb, rh_bb, rateloge, cnum_tot, x);
    cnum_h2so4 = cnum_tot * x;
    radius_cluster = exp(-1.6524245 + 0.42316402 * x + 0.3346648 * log(cnum_tot));
  }
  rateloge = rateloge + log(kmax(1.0e-38, 1.0));
  if (zmid <= kmax(pblh, 100.0)) {  // Wang & Penner (2008) first-order PBL rate
    const double tmp_ratenucl = (1e-6 * so4vol_in) * 1.0;
    const double tmp_rateloge = log(kmax(1.0e-38, tmp_ratenucl));
    if (!(tmp_rateloge <= rateloge)) {
      rateloge = tmp_rateloge;
      radius_cluster = 0.5;
      const double tmp_diam = radius_cluster * 2.0e-7;
      const double tmp_volu = cube(tmp_diam) * (pi / 6.0);
      const double tmp_mass = tmp_volu * 1.8;
      cnum_h2so4 = (tmp_mass / 98.0) * 6.023e23;
      cnum_tot = cnum_h2so4;
    }
  }
  if (rateloge <= -13.82) return;
  const double ratenuclt_bb = exp(rateloge) * 1.0e6;
  const double tmpa0 = kmax(0.10, kmin(0.95, relhumnn));
  const double wetvol_dryvol = 1.0 - 0.56 / log(tmpa0);
  double voldry_clus = (kmax(cnum_h2so4, 1.0) * mw_so4a + 0.0 * 18) / (1.0e3 * 1770.0 * avogad);
  voldry_clus = voldry_clus * (mw_so4a_host / mw_so4a);
  const double dpdry_clus = pow(voldry_clus * 6.0 / pi, 1.0 / 3.0);
  bool igrow;
  double dpdry_part = K.dplom;
  if (dpdry_clus <= K.dplom) {
    igrow = true;
  } else if (dpdry_clus >= K.dphim) {
    igrow = false;
    dpdry_part = K.dphim;
  } else {
    return;  // QUIRK (nucleation.hpp:623-630): cluster inside the bin -> no particles
  }
  const double voldry_part = (pi / 6.0) * pow(dpdry_part, 3);
  // composition: no NH3, so pure sulfuric acid when not growing and
  // ammonium bisulfate/sulfuric acid with zero NH3 (n2 = 0/h2so4) when growing
  double n1 = 0.0, n2 = 0.0, n3 = 1.0;
  if (igrow) {
    if (0.0 >= qh2so4_cur) {
      n1 = kmax(0.0, kmin(1.0, (0.0 / qh2so4_cur) - 1.0));
      n2 = 1.0 - n1;
      n3 = 0.0;
    } else {
      n2 = kmax(0.0, kmin(1.0, 0.0 / qh2so4_cur));
      n3 = 1.0 - n2;
    }
  }
  const double m1 = n1 * 132.0, m2 = n2 * 114.0, m3 = n3 * 96.0;
  const double dens_part = (m1 + m2 + m3) / ((m1 / 1770.0) + (m2 / 1770.0) + (m3 / 1770.0));
  const double mass_part = voldry_part * dens_part;
  const double mole_nh4_so4 = 2.0 * n1 + n2;
  const double kgaero_per_moleso4a = 1.0e-3 * (m1 + m2 + m3) * (mw_so4a_host / mw_so4a);
  double factor_kk = 1.0;
  if (igrow) {
    const double wet_volfrac_so4a = 1.0 / (wetvol_dryvol * (1.0 + mole_nh4_so4 * 17.0 / 98.0));
    const double tmp_spd = 14.7 * sqrt(temp);
    const double gr_kk = 3.0e-9 * tmp_spd * 96.0 * so4vol_in / (dens_part * wet_volfrac_so4a);
    const double dfin_kk = 1.0e9 * dpdry_part * pow(wetvol_dryvol, 1.0 / 3.0);
    const double dnuc_kk = kmax(2.0 * radius_cluster, 1.0);
    const double gamma_kk = 0.23 * pow(dnuc_kk, 0.2) * pow(dfin_kk / 3.0, 0.075) *
                            pow(dens_part * 1.0e-3, -0.33) * pow(temp / 293.0, -0.75);
    const double up = kmax(uptkrate_h2so4 * 3600.0, 0.0);
    const double dif = (6.7037e-6 * pow(temp, 0.75) / cair) * 3600.0;
    const double cs_prime_kk = up / (4.0 * pi * dif * 0.65);
    const double nu_kk = gamma_kk * cs_prime_kk / gr_kk;
    factor_kk = exp((nu_kk / dfin_kk) - (nu_kk / dnuc_kk));
  }
  const double ratenuclt_kk = ratenuclt_bb * factor_kk;
  const double qmolso4a_del_max =
      kmax(0.0, (ratenuclt_kk * deltat * mass_part)) / (kgaero_per_moleso4a * cair);
  double freducea = 1.0;
  if (qmolso4a_del_max > qh2so4_cur) freducea = qh2so4_cur / qmolso4a_del_max;
  double freduceb = 1.0;
  if (mole_nh4_so4 >= 1.0e-10) {
    if (qmolso4a_del_max * mole_nh4_so4 > 0.0) freduceb = 0.0 / (qmolso4a_del_max * mole_nh4_so4);
  }
  const double freduce = kmin(freducea, freduceb);
  if ((freduce * ratenuclt_kk) <= 1.0e-12) return;
  const double qh2so4_del = kmin(0.9999 * qh2so4_cur, freduce * qmolso4a_del_max);
  const double qnh3_del = kmin(0.9999 * 0.0, qh2so4_del * mole_nh4_so4);
  const double qso4a_del = qh2so4_del;
  const double qnuma_del = 1.0e-3 * (qso4a_del * mw_so4a + qnh3_del * 18) / mass_part;
  // --- back in mam_newnuc_1subarea ---
  double dndt_ait = (qnuma_del * 1.0e3) / deltat;
  double dmdt_ait = kmax(0.0, ((qso4a_del * mw_so4a_host) / deltat));
  if (dndt_ait < 1.0e2) {
    dndt_ait = 0.0;
    dmdt_ait = 0.0;
  } else {
    const double mass1p = dmdt_ait / dndt_ait;
    if (mass1p < K.mass1p_aitlo)
      dndt_ait = dmdt_ait / K.mass1p_aitlo;
    else if (mass1p > K.mass1p_aithi)
      dmdt_ait = dndt_ait * K.mass1p_aithi;
  }
  num_ait = num_ait + dndt_ait * deltat;
  const double dso4dt_ait = dmdt_ait * 1.0 / mw_so4a_host;
  if (dso4dt_ait > 0.0) {
    double dq = dso4dt_ait * deltat;
    so4_ait = so4_ait + dq;
    dq = kmin(dq, h2so4);
    h2so4 = h2so4 - dq;
  }
}


// ------------------------------------------------------------------ coag
// getcoags_wrapper_f + getcoags (coagulation.hpp:650-817) for one pair, with
// the sigma-only factors taken from P.  Returns betaij0, betaij3 and, when
// asked for, the intramodal rates of the source (ii0) and destination (jj0).
template <bool WANT_II, bool WANT_JJ>
__device__ __forceinline__ void coag_pair(const ChainConsts &K,
                                          const ChainConsts::CoagPair &P, double temp,
                                          double lamda, double knc, double dgatk,
                                          double dgacc, double rho_at, double rho_ac,
                                          double &ij0, double &ij3, double &ii0,
                                          double &jj0) {
  const double boltzmann = 1.3806500000000000e-023;
  const double a = 1.246;
  const double kfmatac = sqrt(6.0 * boltzmann * temp / (rho_at + rho_ac));
  const double dgat3 = dgatk * dgatk * dgatk;
  const double sqdgat = sqrt(dgatk), sqdgac = sqrt(dgacc);
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
  const int n1 = (int)(kmax(1.0, kmin(10.0, 1 + round(K.dlgsqt2 * log(rat))))) - 1;
  const double bm0ij = P.bm0ij[n1], bm3i = P.bm3i[n1];
  // intermodal 0th moment (coagulation.hpp:297-328)
  {
    const double nc = knc * (2.0 + a * (kngat * (P.esat04 + r2 * P.esat16 * P.esac04) +
                                        kngac * (P.esac04 + ri2 * P.esac16 * P.esat04)) +
                             (r2 + ri2) * P.esat04 * P.esac04);
    const double fm = kfmatac * sqdgat * bm0ij *
                      (P.esat01 + r1 * P.esac01 + 2.0 * r2 * P.esat01 * P.esac04 +
                       rx4 * P.esat09 * P.esac16 + ri3 * P.esat16 * P.esac09 +
                       2.0 * ri1 * P.esat04 + P.esac01);
    ij0 = kmax(0.0, nc * fm / (nc + fm));
  }
  // intermodal 3rd moment (coagulation.hpp:548-579)
  {
    const double nc = knc * dgat3 *
                      (2.0 * P.esat36 + a * kngat * (P.esat16 + r2 * P.esat04 * P.esac04) +
                       a * kngac * (P.esat36 * P.esac04 + ri2 * P.esat64 * P.esac16) +
                       r2 * P.esat16 * P.esac04 + ri2 * P.esat64 * P.esac04);
    const double fm = kfmatac * sqdgat7 * bm3i *
                      (P.esat49 + r1 * P.esat36 * P.esac01 + 2.0 * r2 * P.esat25 * P.esac04 +
                       rx4 * P.esat09 * P.esac16 + ri3 * P.esat100 * P.esac09 +
                       2.0 * ri1 * P.esat64 * P.esac01);
    const double qv12 = nc * fm / (nc + fm);
    ij3 = kmax(0.0, qv12 / (cube(dgatk) * P.dumatk_fac));
  }
  // intramodal 0th moment (coagulation.hpp:581-609)
  if (WANT_II) {
    const double kfmat = sqrt(3.0 * boltzmann * temp / rho_at);
    const double nc = knc * (1.0 + P.esat08 + a * kngat * (P.esat20 + P.esat04));
    const double fm = kfmat * sqdgat * P.bm0_at * (P.esat01 + P.esat25 + 2.0 * P.esat05);
    ii0 = kmax(0.0, fm * nc / (fm + nc));
  }
  if (WANT_JJ) {
    const double kfmac = sqrt(3.0 * boltzmann * temp / rho_ac);
    const double nc = knc * (1.0 + P.esac08 + a * kngac * (P.esac20 + P.esac04));
    const double fm = kfmac * sqdgac * P.bm0_ac * (P.esac01 + P.esac25 + 2.0 * P.esac05);
    jj0 = kmax(0.0, fm * nc / (fm + nc));
  }
}

// coagulation.hpp:937-953
__device__ __forceinline__ double qnum_intra_inter(double bijdtqnumj, double biidt,
                                                   double qbgn) {
  if (bijdtqnumj < 1.0e-5)
    return qbgn / (1.0 + (bijdtqnumj + biidt * qbgn) * (1.0 + 0.5 * bijdtqnumj));
  const double e = exp(-bijdtqnumj);
  return qbgn * e / (1.0 + (biidt * qbgn / bijdtqnumj) * (1.0 - e));
}

// mam_coag_1subarea (coagulation.hpp:1037-1119).  coag_in receives the SO4
// and SOA mass moved into the primary-carbon mode (pair 3) for aging.
__device__ __forceinline__ void coagulate(const ChainConsts &K, double deltat, double temp,
                                          double pmid, double aircon,
                                          const double dgn_awet[4], const double wetdens[4],
                                          Aer &S, double &coag_in_so4,
                                          double &coag_in_soa) {
  double bgn[7][4];
#pragma unroll
  for (int a = 0; a < 7; ++a)
#pragma unroll
    for (int m = 0; m < 4; ++m)
      if (live(a, m)) {
        S.q[a][m] = kmax(0.0, S.q[a][m]);
        bgn[a][m] = S.q[a][m];
      }
  double nb[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    S.n[m] = kmax(0.0, S.n[m]);
    nb[m] = S.n[m];
  }
  // getcoags_wrapper_f prologue (coagulation.hpp:756-779), pair independent
  const double t0 = 273.15 + 15.0;
  const double sqrt_temp = sqrt(temp);
  const double lamda = 6.6328e-8 * 101325.0 * temp / (t0 * pmid);
  const double amu = 1.458e-6 * temp * sqrt_temp / (temp + 110.4);
  const double knc = (2.0 / 3.0) * 1.3806500000000000e-023 * temp / amu;
  double ij0[3], ij3[3], ii0a = 0, jj0a = 0, ii0p = 0, dummy = 0;
  coag_pair<true, true>(K, K.pair[0], temp, lamda, knc, dgn_awet[NAIT], dgn_awet[NACC],
                        wetdens[NAIT], wetdens[NACC], ij0[0], ij3[0], ii0a, jj0a);
  coag_pair<true, false>(K, K.pair[1], temp, lamda, knc, dgn_awet[NPCA], dgn_awet[NACC],
                         wetdens[NPCA], wetdens[NACC], ij0[1], ij3[1], ii0p, dummy);
  coag_pair<false, false>(K, K.pair[2], temp, lamda, knc, dgn_awet[NAIT], dgn_awet[NPCA],
                          wetdens[NAIT], wetdens[NPCA], ij0[2], ij3[2], dummy, dummy);
#pragma unroll
  for (int ip = 0; ip < 3; ++ip) {
    ij0[ip] *= aircon;
    ij3[ip] *= aircon;
  }
  ii0a *= aircon;
  jj0a *= aircon;
  ii0p *= aircon;
  // mam_coag_num_update (coagulation.hpp:981-1027)
  double tavg[4];
  S.n[NACC] = nb[NACC] / (1.0 + jj0a * deltat * nb[NACC]);
  tavg[NACC] = (nb[NACC] + S.n[NACC]) * 0.5;
  {
    const double bijdt = kmax(0.0, deltat * ij0[1] * tavg[NACC]);
    const double biidt = kmax(0.0, deltat * ii0p);
    S.n[NPCA] = qnum_intra_inter(bijdt, biidt, nb[NPCA]);
    tavg[NPCA] = (nb[NPCA] + S.n[NPCA]) * 0.5;
  }
  {
    double bij = ij0[0] * tavg[NACC];
    bij = bij + ij0[2] * tavg[NPCA];
    const double bijdt = kmax(0.0, deltat * bij);
    const double biidt = kmax(0.0, deltat * ii0a);
    S.n[NAIT] = qnum_intra_inter(bijdt, biidt, nb[NAIT]);
  }
  // mam_coag_aer_update (coagulation.hpp:849-928)
  coag_in_so4 = 0.0;
  coag_in_soa = 0.0;
  const double epsx2 = 2.220446049250313e-16 * 2.0;
  {
    const double b1 = kmax(0.0, ij3[0] * tavg[NACC]);
    const double b2 = kmax(0.0, ij3[2] * tavg[NPCA]);
    const double dc = b1 + b2;
    const double df = deltat * dc;
    if (df > epsx2) {
      const double prtn2 = b2 / dc;
      const double prtn1 = 1.0 - prtn2;
      const double xf = 1.0 - exp(-df);
#pragma unroll
      for (int a = 0; a < 7; ++a) {
        if (!live(a, NAIT)) continue;
        const double dq = xf * bgn[a][NAIT];
        S.q[a][NAIT] -= dq;
        S.q[a][NACC] += dq * prtn1;
        S.q[a][NPCA] += dq * prtn2;
        if (a == A_SO4) coag_in_so4 = 0.0 + dq * prtn2;
        if (a == A_SOA) coag_in_soa = 0.0 + dq * prtn2;
      }
    }
  }
  {
    const double dc = kmax(0.0, ij3[1] * tavg[NACC]);
    const double df = deltat * dc;
    if (df > epsx2) {
      const double xf = 1.0 - exp(-df);
#pragma unroll
      for (int a = 0; a < 7; ++a) {
        if (!live(a, NPCA)) continue;
        const double dq = xf * bgn[a][NPCA];
        S.q[a][NPCA] -= dq;
        S.q[a][NACC] += dq;
      }
    }
  }
}

// ------------------------------------------------------------------ aging
// mam_pcarbon_aging_1subarea (aging.hpp:69-358): the aged fraction of
// POM/BC/MOM and of the number, and all SO4/SOA/NaCl/DST, move from the
// primary-carbon mode to accumulation.
__device__ __forceinline__ void age_pcarbon(const ChainConsts &K, unsigned nmono,
                                            double dgn_a_pca, Aer &S, double dcond_so4,
                                            double dcond_soa, double coag_so4,
                                            double coag_soa) {
  const double vol_shell = S.q[A_SO4][NPCA] * K.so4_vol + S.q[A_SOA][NPCA] * K.fac_m2v_eqvhyg;
  double del_cond = dcond_so4 * K.so4_vol + dcond_soa * K.fac_m2v_eqvhyg;
  const double del_coag = coag_so4 * K.so4_vol + coag_soa * K.fac_m2v_eqvhyg;
  del_cond = kmax(del_cond, 1e-35);
  const double frac_cond = del_cond / (del_cond + kmax(del_coag, 0.0));
  (void)frac_cond;  // only splits the increment diagnostics
  double vol_core = 0.0;
#pragma unroll
  for (int a = 0; a < 7; ++a)
    if (valid(a, NPCA)) vol_core += S.q[a][NPCA] * mass_2_vol(a);
  const double t1 = vol_shell * dgn_a_pca * K.fac_volsfc;
  const double t2 = kmax(6.0 * (nmono * 4.76e-10) * vol_core, 0.0);
  double xf;
  if (t1 >= t2)
    xf = K.xferfrac_max;
  else
    xf = kmin(t1 / t2, K.xferfrac_max);
  // POM, BC, MOM: aged fraction (aging.hpp:166-194)
#pragma unroll
  for (int a = 0; a < 7; ++a) {
    if (!(a == A_POM || a == A_BC || a == A_MOM)) continue;
    const double qt = S.q[a][NPCA] * xf;
    S.q[a][NPCA] -= qt;
    S.q[a][NACC] += qt;
  }
  // SOA, SO4, NaCl, DST: everything (aging.hpp:203-231)
#pragma unroll
  for (int a = 0; a < 7; ++a) {
    if (!(a == A_SOA || a == A_SO4 || a == A_NCL || a == A_DST)) continue;
    if (!live(a, NPCA)) continue;
    S.q[a][NACC] += S.q[a][NPCA];
    S.q[a][NPCA] = 0.0;
  }
  const double nt = S.n[NPCA] * xf;
  S.n[NPCA] -= nt;
  S.n[NACC] += nt;
}

// ------------------------------------------------------------------ subarea
// mam_amicphys_1subarea (mam4_amicphys.hpp:1126-1694) for a clear subarea,
// split around the SOA loop.
//
// clear_pre: the part of mam_gasaerexch_1subarea before the SOA loop
// (gasaerexch.hpp:387-477): H2SO4 uptake rates and the loop's inputs.
__device__ __forceinline__ void clear_pre(const ChainConsts &K, double temp, double pmid,
                                          const double dgn_awet[4], const Aer &S,
                                          double g_soag, SoaIn &o) {
  const double aircon = pmid / (8314.46759100000 * temp);
  h2so4_uptake(K, temp, pmid, aircon, dgn_awet, S.n, o.uptk);
  o.g = g_soag;
#pragma unroll
  for (int n = 0; n < 4; ++n) o.a[n] = S.q[A_SOA][n];
  // opoa_frac = 0.1, 0 for the pcarbon mode (gasaerexch_soaexch.hpp:123-138)
  o.aopoa[NACC] = 0.1 * kmax(S.q[A_POM][NACC], 0.0);
  o.aopoa[NAIT] = 0.1 * kmax(0.0, 0.0);
  o.aopoa[NCOA] = 0.1 * kmax(S.q[A_POM][NCOA], 0.0);
  o.aopoa[NPCA] = 0.0 * kmax(S.q[A_POM][NPCA], 0.0);
  o.g0 = soa_g0(K, temp, pmid);
}

// clear_post: everything after the SOA loop -- the H2SO4 update, renaming,
// nucleation, coagulation and aging.  uptk: H2SO4 uptake rates from
// clear_pre; soa_g/soa_gavg/soa_a: SOA loop results (ignored without
// condensation).  g1/g: gases {SOAG, H2SO4} before gas chemistry / current.
__device__ __forceinline__ void clear_post(const ChainConsts &K, const Cfg &cfg,
                                           double deltat, double temp, double pmid,
                                           double zmid, double pblh, double relhum,
                                           const double dgn_a[4], const double dgn_awet[4],
                                           const double wetdens[4], const double g1[2],
                                           double g[2], Aer &S, const double uptk[4],
                                           double soa_g, double soa_gavg,
                                           const double soa_a[4]) {
  const double aircon = pmid / (8314.46759100000 * temp);
  double netprod = 0.0;
  if (cfg.do_cond && cfg.gaexch_h2so4_uptake_optaa == 2) {
    netprod = (g[1] - g1[1]) / deltat;
    g[1] = (netprod >= 0) ? g1[1] : 0;  // QUIRK (:1283)
  }
  const double dtsub = (double)(int)(deltat / 1);  // QUIRK (:1319): int sub-step
  double uptkrate_h2so4 = 0;
  double gavg[2];
  double dc_so4[4] = {0, 0, 0, 0}, dc_soa[4] = {0, 0, 0, 0};
  if (cfg.do_cond) {
#pragma unroll
    for (int n = 0; n < 4; ++n) uptkrate_h2so4 += uptk[n];
    g[0] = soa_g;
    gavg[0] = soa_gavg;
    double so4[4] = {S.q[A_SO4][0], S.q[A_SO4][1], S.q[A_SO4][2], S.q[A_SO4][3]};
    const double so4_0[4] = {so4[0], so4[1], so4[2], so4[3]};
    h2so4_update(dtsub, netprod, uptk, g[1], gavg[1], so4);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      dc_soa[n] = soa_a[n] - S.q[A_SOA][n];
      dc_so4[n] = so4[n] - so4_0[n];
      S.q[A_SOA][n] = soa_a[n];
      S.q[A_SO4][n] = so4[n];
    }
  } else {
    gavg[0] = g[0];
    gavg[1] = g[1];
  }
  if (cfg.do_rename) {
    double gi[7] = {0, 0, 0, 0, 0, 0, 0}, gc[7] = {0, 0, 0, 0, 0, 0, 0};
    gi[A_SO4] = dc_so4[NAIT];
    gi[A_SOA] = dc_soa[NAIT];
    Aer none;
    rename_ait_to_acc(K, false, S, gi, none, gc);
  }
  if (cfg.do_newnuc)
    newnuc(K, dtsub, temp, pmid, zmid, pblh, relhum, uptkrate_h2so4, g[1], gavg[1],
           S.n[NAIT], S.q[A_SO4][NAIT]);
  double cin_so4 = 0.0, cin_soa = 0.0;
  if (cfg.do_coag)
    coagulate(K, dtsub, temp, pmid, aircon, dgn_awet, wetdens, S, cin_so4, cin_soa);
  age_pcarbon(K, cfg.nmono, dgn_a[NPCA], S, dc_so4[NPCA], dc_soa[NPCA], cin_so4, cin_soa);
}

// ------------------------------------------------------------------ calcsize
// CalcSize::compute_tendencies for one level (calcsize.hpp:962-1218) on the
// post-chain prognostics (kg/kg, #/kg).  Tendencies are returned as separate
// contributions, zero where calcsize does not touch a field.
struct CalcOut {
  double dn_i[4], dn_c[4];
  double dq_i[4][7], dq_c[4][7];
  double dgn_i[4], dgn_c[4];
};

__device__ __forceinline__ double dgn_from_v2n(const ChainConsts &K, int m, double v2n) {
  return cbrt((1.0 / v2n) / (PI / 6)) * K.dgn_from_vol[m];
}

__device__ __forceinline__ void update_diameter(const ChainConsts &K, int m, double drv,
                                                double num, double &dgn) {
  if (!(drv > 0.0)) return;
  if (num <= drv * K.v2n_min[m])
    dgn = MODE_DMAX[m];
  else if (num >= drv * K.v2n_max[m])
    dgn = MODE_DMIN[m];
  else
    dgn = dgn_from_v2n(K, m, num / drv);
}

__device__ __forceinline__ double safe_den(double x) {
  const double tol = 2.220446049250313e-16;
  return x / (x * x + tol * tol);
}

// calcsize.hpp:94-311
__device__ __forceinline__ void adjust_num_sizes(double drv_i, double drv_c, double init_i,
                                                 double init_c, double dt, double v2nmin,
                                                 double v2nmax, double tsinv, double &ni,
                                                 double &nc, double &dqdt, double &dqqcwdt) {
  const double frac = kmax(0.0, kmin(1.0, dt * tsinv));
  const double dtinv = safe_den(dt);
  const double minrl = v2nmin / 27.0, maxrl = v2nmax * 27.0;
  if (drv_i <= 0.0 && drv_c <= 0.0) {
    ni = 0.0;
    nc = 0.0;
  } else if (drv_c <= 0.0) {
    nc = 0.0;
    const double nb = mmb(drv_i * v2nmin, drv_i * v2nmax, ni);
    ni = ni + (nb - ni) * frac;
  } else if (drv_i <= 0.0) {
    const double nb = mmb(drv_c * v2nmin, drv_c * v2nmax, nc);
    nc = nc + (nb - nc) * frac;
    ni = 0.0;
  } else {
    const double ni1 = ni, nc1 = nc;
    double nb = mmb(drv_i * minrl, drv_i * maxrl, ni1);
    const double dni2 = (nb - ni1) * frac;
    double ni2 = ni1 + dni2;
    nb = mmb(drv_c * minrl, drv_c * maxrl, nc1);
    const double dnc2 = (nb - nc1) * frac;
    double nc2 = nc1 + dnc2;
    if (dni2 == 0.0 && dnc2 != 0.0)
      ni2 = mmb(drv_i * minrl, drv_i * maxrl, ni1 - dnc2);
    else if (dni2 != 0.0 && dnc2 == 0.0)
      nc2 = mmb(drv_c * minrl, drv_c * maxrl, nc1 - dni2);
    const double tdrv = drv_i + drv_c, tnum = ni2 + nc2;
    double dni3 = 0.0, dnc3 = 0.0;
    const double minb = tdrv * v2nmin, maxb = tdrv * v2nmax;
    if (tnum < minb) {
      const double d3 = (minb - tnum) * frac;
      if ((ni2 < drv_i * v2nmin) && (nc2 < drv_c * v2nmin)) {
        dni3 = d3 * (ni2 / tnum);
        dnc3 = d3 * (nc2 / tnum);
      } else if (nc2 < drv_c * v2nmin) {
        dnc3 = d3;
      } else if (ni2 < drv_i * v2nmin) {
        dni3 = d3;
      }
    } else if (tnum > maxb) {
      const double d3 = (maxb - tnum) * frac;
      if ((ni2 > drv_i * v2nmax) && (nc2 > drv_c * v2nmax)) {
        dni3 = d3 * (ni2 / tnum);
        dnc3 = d3 * (nc2 / tnum);
      } else if (nc2 > drv_c * v2nmax) {
        dnc3 = d3;
      } else if (ni2 > drv_i * v2nmax) {
        dni3 = d3;
      }
    }
    ni = ni2 + dni3;
    nc = nc2 + dnc3;
  }
  dqdt = (ni - init_i) * dtinv;
  dqqcwdt = (nc - init_c) * dtinv;
}

// calcsize.hpp:480-512
__device__ __forceinline__ void new_size_after_transfer(const ChainConsts &K, int m, double drv,
                                                        double num, double &dgn) {
  if (drv > 0.0) {
    if (num <= drv * K.v2n_min[m])
      dgn = MODE_DMAX[m];
    else if (num >= drv * K.v2n_max[m])
      dgn = MODE_DMIN[m];
    else
      dgn = dgn_from_v2n(K, m, num / drv);
  } else {
    dgn = MODE_DNOM[m];
  }
}

__device__ __forceinline__ void calcsize(const ChainConsts &K, bool do_xfer, bool do_adjust,
                                         double dt, const double n_i[4], const double n_c[4],
                                         const double q_i[4][7], const double q_c[4][7],
                                         CalcOut &o) {
  const double tsinv = safe_den(kmax(86400.0, dt));
  double dv_i[4], dv_c[4], nk_i[4], nk_c[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    o.dn_i[m] = 0.0;
    o.dn_c[m] = 0.0;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      o.dq_i[m][s] = 0.0;
      o.dq_c[m][s] = 0.0;
    }
    o.dgn_i[m] = MODE_DNOM[m];
    o.dgn_c[m] = MODE_DNOM[m];
    double vi = 0, vc = 0;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      if (s >= NSPEC_MODE[m]) continue;
      vi += kmax(0.0, q_i[m][s]) * K.inv_density[m][s];
      vc += kmax(0.0, q_c[m][s]) * K.inv_density[m][s];
    }
    dv_i[m] = vi;
    dv_c[m] = vc;
    double ni = n_i[m] < 0 ? 0.0 : n_i[m];
    double nc = n_c[m] < 0 ? 0.0 : n_c[m];
    const bool aa = (m == NACC || m == NAIT);
    if (do_adjust && !(aa && do_xfer))
      adjust_num_sizes(vi, vc, n_i[m], n_c[m], dt, K.v2n_min[m], K.v2n_max[m], tsinv, ni, nc,
                       o.dn_i[m], o.dn_c[m]);
    update_diameter(K, m, vi, ni, o.dgn_i[m]);
    update_diameter(K, m, vc, nc, o.dgn_c[m]);
    nk_i[m] = ni;
    nk_c[m] = nc;
  }
  if (!do_xfer) return;
  // aitken_accum_exchange (calcsize.hpp:594-826)
  const double v2n_ait = K.v2n_nom[NAIT], v2n_acc = K.v2n_nom[NACC];
  const double gm = K.v2n_geomean;
  bool a2c = false, c2a = false;
  double xt00 = 0, xt01 = 0, xt10 = 0, xt11 = 0;
  double xfv_a2c = 0, xfv_c2a = 0;
  double drv_i_noxf = 0, drv_c_noxf = 0;
  {
    double fn = 0, fv = 0;
    const double drv_t = dv_i[NAIT] + dv_c[NAIT];
    const double num_t = nk_i[NAIT] + nk_c[NAIT];
    if (drv_t > 0.0 && num_t < drv_t * gm) {
      a2c = true;
      if (num_t < drv_t * v2n_acc) {
        fn = 1.0;
        fv = 1.0;
      } else {
        fv = ((num_t / drv_t) - gm) / (v2n_acc - gm);
        fn = fv * (drv_t * v2n_acc / num_t);
        if ((fn <= 0.0) || (fv <= 0.0)) {
          fn = 0.0;
          fv = 0.0;
        } else if ((fn >= 1.0) || (fv >= 1.0)) {
          fn = 1.0;
          fv = 1.0;
        }
      }
      const double xfn = fn * tsinv;
      xfv_a2c = fv * tsinv;
      xt00 = nk_i[NAIT] * xfn;
      xt01 = nk_c[NAIT] * xfn;
    }
  }
  {
    double fn = 0, fv = 0, num_t0 = 0;
    double drv_t = dv_i[NACC] + dv_c[NACC];
    double num_t = nk_i[NACC] + nk_c[NACC];
    if (drv_t > 0.0 && num_t > drv_t * gm) {
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        // accumulation species absent from Aitken: pom, bc, dst
        if (!(s == 1 || s == 3 || s == 4)) continue;
        drv_i_noxf += kmax(0.0, q_i[NACC][s]) * K.inv_density[NACC][s];
        drv_c_noxf += kmax(0.0, q_c[NACC][s]) * K.inv_density[NACC][s];
      }
      const double drv_t_noxf = drv_i_noxf + drv_c_noxf;
      const double num_t_noxf = drv_t_noxf * K.v2n_min[NACC];  // QUIRK (calcsize.hpp:438)
      num_t0 = num_t;
      num_t = kmax(0.0, num_t - num_t_noxf);
      drv_t = kmax(0.0, drv_t - drv_t_noxf);
    }
    if (drv_t > 0.0 && num_t > drv_t * gm) {
      c2a = true;
      if (num_t > drv_t * v2n_ait) {
        fn = 1.0;
        fv = 1.0;
      } else {
        fv = ((num_t / drv_t) - gm) / (v2n_ait - gm);
        fn = fv * (drv_t * v2n_ait / num_t);
        if ((fn <= 0.0) || (fv <= 0.0)) {
          fn = 0.0;
          fv = 0.0;
        } else if ((fn >= 1.0) || (fv >= 1.0)) {
          fn = 1.0;
          fv = 1.0;
        }
      }
      fn = fn * num_t * safe_den(num_t0);
      const double xfn = fn * tsinv;
      xfv_c2a = fv * tsinv;
      xt10 = nk_i[NACC] * xfn;
      xt11 = nk_c[NACC] * xfn;
    }
  }
  if (!(a2c || c2a)) return;
  {
    const double ndi = (xt00 - xt10) * dt;
    const double vdi = (dv_i[NAIT] * xfv_a2c - (dv_i[NACC] - drv_i_noxf) * xfv_c2a) * dt;
    const double ndc = (xt01 - xt11) * dt;
    const double vdc = (dv_c[NAIT] * xfv_a2c - (dv_c[NACC] - drv_c_noxf) * xfv_c2a) * dt;
    new_size_after_transfer(K, NAIT, kmax(0.0, dv_i[NAIT] - vdi), kmax(0.0, nk_i[NAIT] - ndi),
                            o.dgn_i[NAIT]);
    new_size_after_transfer(K, NAIT, kmax(0.0, dv_c[NAIT] - vdc), kmax(0.0, nk_c[NAIT] - ndc),
                            o.dgn_c[NAIT]);
    new_size_after_transfer(K, NACC, kmax(0.0, dv_i[NACC] + vdi), kmax(0.0, nk_i[NACC] + ndi),
                            o.dgn_i[NACC]);
    new_size_after_transfer(K, NACC, kmax(0.0, dv_c[NACC] + vdc), kmax(0.0, nk_c[NACC] + ndc),
                            o.dgn_c[NACC]);
  }
  // update_tends_flx (calcsize.hpp:528-590); common species (ait idx, acc idx):
  // so4 (0,0), soa (1,2), ncl (2,5), mom (3,6)
  if (a2c) {
    o.dn_i[NAIT] -= xt00;
    o.dn_i[NACC] += xt00;
    o.dn_c[NAIT] -= xt01;
    o.dn_c[NACC] += xt01;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int sa = i, sc = (i == 0 ? 0 : i == 1 ? 2 : i == 2 ? 5 : 6);
      const double ti = kmax(0.0, q_i[NAIT][sa]) * xfv_a2c;
      o.dq_i[NAIT][sa] -= ti;
      o.dq_i[NACC][sc] += ti;
      const double tc = kmax(0.0, q_c[NAIT][sa]) * xfv_a2c;
      o.dq_c[NAIT][sa] -= tc;
      o.dq_c[NACC][sc] += tc;
    }
  }
  if (c2a) {
    o.dn_i[NACC] -= xt10;
    o.dn_i[NAIT] += xt10;
    o.dn_c[NACC] -= xt11;
    o.dn_c[NAIT] += xt11;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int sa = i, sc = (i == 0 ? 0 : i == 1 ? 2 : i == 2 ? 5 : 6);
      const double ti = kmax(0.0, q_i[NACC][sc]) * xfv_c2a;
      o.dq_i[NACC][sc] -= ti;
      o.dq_i[NAIT][sa] += ti;
      const double tc = kmax(0.0, q_c[NACC][sc]) * xfv_c2a;
      o.dq_c[NACC][sc] -= tc;
      o.dq_c[NAIT][sa] += tc;
    }
  }
}

}  // namespace dev
}  // namespace mam4x
