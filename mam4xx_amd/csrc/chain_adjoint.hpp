// Reverse sweep (adjoint) of the per-cell microphysics step, FP64.
//
// Each routine f_rev recomputes the forward values of its routine f from f's
// inputs -- with the same operations in the same order, so every branch
// takes the primal's side -- and then propagates the adjoints of f's outputs
// back to its inputs, accumulating (+=) into the input adjoints.  The
// derivative is that of the branch each cell takes, as in the forward-mode
// tangent (scalar.hpp, Dual): clips, regime switches and the SOA sub-step
// count are held at the primal's choice; Kokkos' min/max ternaries pass the
// adjoint to the operand they selected.
//
// The SOA loop is reversed sub-step by sub-step from the states saved at the
// start of each sub-step (SoaCkpt); within a sub-step everything is
// recomputed.
#pragma once

#include "chain_cell.hpp"

namespace mam4x {
namespace dev {
namespace rev {

// ------------------------------------------------------------------ helpers
// y = kmax(a, b) = (a < b) ? b : a
__host__ __device__ __forceinline__ void max_rev(double a, double b, double yb, double &ab,
                                                 double &bb) {
  if (a < b)
    bb += yb;
  else
    ab += yb;
}
// y = kmin(a, b) = (b < a) ? b : a
__host__ __device__ __forceinline__ void min_rev(double a, double b, double yb, double &ab,
                                                 double &bb) {
  if (b < a)
    bb += yb;
  else
    ab += yb;
}
// y = kmax(c, x) with a constant c: the adjoint reaches x when x was selected
__host__ __device__ __forceinline__ double cmax_pass(double c, double x, double yb) {
  return (c < x) ? yb : 0.0;
}
// y = kmax(x, c)
__host__ __device__ __forceinline__ double maxc_pass(double x, double c, double yb) {
  return (x < c) ? 0.0 : yb;
}
// y = kmin(c, x)
__host__ __device__ __forceinline__ double cmin_pass(double c, double x, double yb) {
  return (x < c) ? yb : 0.0;
}
// y = kmin(x, c)
__host__ __device__ __forceinline__ double minc_pass(double x, double c, double yb) {
  return (c < x) ? 0.0 : yb;
}
// y = mmb(lo, hi, x) = kmax(lo, kmin(hi, x)) with constant bounds
__host__ __device__ __forceinline__ double mmb_pass(double lo, double hi, double x, double yb) {
  const double m = kmin(hi, x);
  return (lo < m) ? cmin_pass(hi, x, yb) : 0.0;
}

// ------------------------------------------------------------------ uptake
// h2so4_uptake (dgn_awet held constant): adjoints of temp, pmid, aircon and
// the numbers from those of the four rates.
__host__ __device__ __forceinline__ void h2so4_uptake_rev(const ChainConsts &K, double temp,
                                                          double pmid, double aircon,
                                                          const double dgn_awet[4],
                                                          const double qnum[4],
                                                          const double ub[4], double &temp_b,
                                                          double &pmid_b, double &aircon_b,
                                                          double qnum_b[4]) {
  const double p_in_atm = pmid / 1.013e5;
  const double dgas = (1.0e-3 * pow(temp, 1.75) * K.h2so4_sqrt_mw) /
                      (p_in_atm * K.h2so4_vd_term);
  const double gasdiffus = dgas * 1.0e-4;
  const double speed = 145.5 * sqrt(temp / 9.807840000000000E+01);
  const double freepath = 3.0 * gasdiffus / speed;
  const double a283 = 6.5000E-01 * 0.283, a75 = 6.5000E-01 * 0.75;
  const double xghq0 = 7.071067690849304E-01, wghq = 8.862269520759583E-01;
  double gd_b = 0.0, fp_b = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const double lnsg = K.alnsg[n];
    const double lndpgn = log(dgn_awet[n]);
    const double kn0 = 2.0 * freepath / dgn_awet[n];
    const double dk = kn0 * (kn0 + 1.0 + a283) + a75;
    const double tmpa = 1.0 / (1.0 + kn0) - (2.0 * kn0 + 1.0 + a283) / dk;
    const double beta0 = 1.0 - kn0 * tmpa;
    const double beta = kmax(1.0, kmin(2.0, beta0));
    const double constant = K.tworootpi * exp(beta * lndpgn + 0.5 * sq(beta * lnsg));
    double sumghq = 0.0, dp[2], hh[2], pw[2], lndp[2];
#pragma unroll
    for (int iq = 0; iq < 2; ++iq) {
      const double x = iq == 0 ? xghq0 : -xghq0;
      lndp[iq] = lndpgn + beta * lnsg * lnsg + K.root2 * lnsg * x;
      dp[iq] = exp(lndp[iq]);
      const double kn = 2.0 * freepath / dp[iq];
      hh[iq] = (a75 * (1.0 + kn)) / (kn * (kn + 1.0 + a283) + a75);
      pw[iq] = pow(dp[iq], beta);
      sumghq += wghq * dp[iq] * hh[iq] / pw[iq];
    }
    const double rate = constant * gasdiffus * sumghq;
    const double nair = qnum[n] * aircon;
    // uptk = rate * (qnum * aircon)
    const double rate_b = ub[n] * nair;
    qnum_b[n] += ub[n] * rate * aircon;
    aircon_b += ub[n] * rate * qnum[n];
    const double const_b = rate_b * gasdiffus * sumghq;
    gd_b += rate_b * constant * sumghq;
    const double sum_b = rate_b * constant * gasdiffus;
    // constant = tworootpi exp(beta lndpgn + 0.5 (beta lnsg)^2)
    double beta_b = const_b * constant * (lndpgn + beta * lnsg * lnsg);
#pragma unroll
    for (int iq = 0; iq < 2; ++iq) {
      // t = wghq dp hh / pw
      const double t = wghq * dp[iq] * hh[iq] / pw[iq];
      double dp_b = sum_b * wghq * hh[iq] / pw[iq];
      const double hh_b = sum_b * wghq * dp[iq] / pw[iq];
      const double pw_b = -sum_b * t / pw[iq];
      // pw = pow(dp, beta)
      dp_b += pw_b * pw[iq] * beta / dp[iq];
      beta_b += pw_b * pw[iq] * log(dp[iq]);
      // hh = a75 (1 + kn) / (kn (kn + 1 + a283) + a75)
      const double kn = 2.0 * freepath / dp[iq];
      const double den = kn * (kn + 1.0 + a283) + a75;
      const double kn_b = hh_b * (a75 / den - hh[iq] * (2.0 * kn + 1.0 + a283) / den);
      fp_b += kn_b * 2.0 / dp[iq];
      dp_b -= kn_b * kn / dp[iq];
      // dp = exp(lndp), lndp = lndpgn + beta lnsg^2 + ...
      beta_b += dp_b * dp[iq] * lnsg * lnsg;
    }
    // beta = clamp(1 - kn0 tmpa, 1, 2)
    const double b0_b = cmax_pass(1.0, kmin(2.0, beta0), cmin_pass(2.0, beta0, 1.0)) * beta_b;
    if (b0_b != 0.0) {
      const double dtmpa = -1.0 / sq(1.0 + kn0) -
                           (2.0 * dk - sq(2.0 * kn0 + 1.0 + a283)) / sq(dk);
      const double kn0_b = -b0_b * (tmpa + kn0 * dtmpa);
      fp_b += kn0_b * 2.0 / dgn_awet[n];
    }
  }
  // freepath = 3 gasdiffus / speed
  gd_b += fp_b * 3.0 / speed;
  const double speed_b = -fp_b * freepath / speed;
  temp_b += speed_b * speed / (2.0 * temp);
  // gasdiffus = dgas 1e-4, dgas = 1e-3 T^1.75 c / (p_atm vd)
  const double dgas_b = gd_b * 1.0e-4;
  temp_b += dgas_b * dgas * 1.75 / temp;
  pmid_b += -dgas_b * dgas / p_in_atm / 1.013e5;
}

// soa_g0 (temp, pmid)
__host__ __device__ __forceinline__ void soa_g0_rev(const ChainConsts &K, double temp,
                                                    double pmid, double g0_b, double &temp_b,
                                                    double &pmid_b) {
  const double p0 = 1.0e-10 * exp(-(K.soa_c1) * ((1.0 / temp) - K.soa_inv298));
  const double g0 = 1.01325e5 * p0 / pmid;
  pmid_b += -g0_b * g0 / pmid;
  const double p0_b = g0_b * 1.01325e5 / pmid;
  temp_b += p0_b * p0 * K.soa_c1 / (temp * temp);
}

// ------------------------------------------------------------------ SOA
// The state a sub-step starts from: all that soa_tick reads besides the
// loop's constant inputs (uk, aopoa, g0, skip).
struct SoaCkpt {
  double g, a[4], tcur;
};

// Adjoints of the loop state and of its constant inputs.
struct SoaBar {
  double g, a[4], tcur, avg, dtsum;  // state
  double uk[4], aopoa[4], g0;         // constants of the loop
};

// Reverse of one soa_tick from the state c it started at, given the
// adjoints b of the state after it; on return b holds the adjoints of the
// state before it (constants' adjoints accumulated).  The forward part
// reproduces soa_tick<FAST = true>'s values bit for bit (num::div_r with the
// shared reciprocal is num::div_fast); every divisor is inverted once and its
// reciprocal serves the forward quotient and all reverse uses.
__host__ __device__ __forceinline__ void soa_tick_rev(const SoaState<double> &k,
                                                      const SoaCkpt &c, double dtfull,
                                                      SoaBar &b) {
  const double alpha_astem = 0.05, a_min1 = 1.0e-20, g_min1 = 1.0e-20;
  // ---- forward, as soa_tick ----
  const double gprv = c.g;
  const double g_soa0 = kmax(gprv, 0.0);
  double a_soa[4], sat0[4], ryy[4], phi[4], gst[4], mden[4], rm[4];
  double tot = g_soa0, tmpb = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    a_soa[n] = k.skip[n] ? 0.0 : kmax(c.a[n], 0.0);
    const double yy = kmax(k.aopoa[n] + a_soa[n], a_min1);
    ryy[n] = num::rcp(yy);
    sat0[n] = num::div_r(k.g0, yy, ryy[n]);
    gst[n] = sat0[n] * a_soa[n];
    mden[n] = kmax(g_soa0, kmax(gst[n], g_min1));
    rm[n] = num::rcp(mden[n]);
    phi[n] = num::div_r(g_soa0 - gst[n], mden[n], rm[n]);
  }
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    tot = tot + soa_term(k.skip[n], a_soa[n]);
    tmpb = tmpb + soa_term(k.skip[n], k.uk[n] * fabs(phi[n]));
  }
  const double tmpa = kmax(0.0, tmpb);
  const double dtmax = dtfull - c.tcur;
  const bool last = dtmax * tmpa <= alpha_astem;
  double rtmpa = 0.0, dtcur = dtmax;
  if (!last) {  // then tmpa > alpha / dtmax > 0
    rtmpa = num::rcp(tmpa);
    dtcur = num::div_r(alpha_astem, tmpa, rtmpa);
  }
  double beta[4], sat[4], del[4], at[4], rxx[4], rden[4], qa[4];
  double sa = 0.0, sb = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    beta[n] = dtcur * k.uk[n];
    del[n] = g_soa0 - gst[n];
    at[n] = a_soa[n] + beta[n] * del[n];
    const double xx = kmax(k.aopoa[n] + at[n], a_min1);
    rxx[n] = num::rcp(xx);
    const double sat1 = num::div_r(k.g0, xx, rxx[n]);
    sat[n] = del[n] > 0.0 ? sat1 : sat0[n];
    const double den = 1.0 + beta[n] * sat[n];
    rden[n] = num::rcp(den);
    qa[n] = num::div_r(a_soa[n], den, rden[n]);
    sa = sa + soa_term(k.skip[n], qa[n]);
    sb = sb + soa_term(k.skip[n], num::div_r(beta[n], den, rden[n]));
  }
  const double r1sb = num::rcp(1.0 + sb);
  const double gq = num::div_r(tot - sa, 1.0 + sb, r1sb);
  const double g_new = kmax(0.0, gq);

  // ---- reverse ----
  double dtcur_b = 0.0, gprv_b = 0.0, gnew_b = b.g;
  // avg += dtcur (gprv + 0.5 (g_new - gprv)); dtsum += dtcur
  gprv_b += b.avg * 0.5 * dtcur;
  gnew_b += b.avg * 0.5 * dtcur;
  dtcur_b += b.avg * (gprv + 0.5 * (g_new - gprv)) + b.dtsum;
  double a_soa_b[4], beta_b[4], den_b[4];
  // a = (a_soa + beta g_new) / den
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    a_soa_b[n] = beta_b[n] = den_b[n] = 0.0;
    if (k.skip[n]) continue;
    const double anew = (a_soa[n] + beta[n] * g_new) * rden[n];
    const double ab = b.a[n] * rden[n];
    a_soa_b[n] += ab;
    beta_b[n] += ab * g_new;
    gnew_b += ab * beta[n];
    den_b[n] -= ab * anew;
  }
  // g_new = kmax(0, (tot - sa) / (1 + sb))
  const double gq_b = cmax_pass(0.0, gq, gnew_b);
  const double tot_b = gq_b * r1sb;
  const double sa_b = -tot_b;
  const double sb_b = -tot_b * gq;
  double g_soa_b = 0.0, sat0_b[4] = {0, 0, 0, 0};
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    if (k.skip[n]) continue;
    // sa += a_soa / den, sb += beta / den
    a_soa_b[n] += sa_b * rden[n];
    den_b[n] -= sa_b * qa[n] * rden[n];
    beta_b[n] += sb_b * rden[n];
    den_b[n] -= sb_b * (beta[n] * rden[n]) * rden[n];
    // den = 1 + beta sat
    beta_b[n] += den_b[n] * sat[n];
    const double sat_b = den_b[n] * beta[n];
    // sat = del > 0 ? g0 / kmax(aopoa + at, a_min1) : sat0
    double at_b = 0.0, del_b = 0.0;
    if (del[n] > 0.0) {
      b.g0 += sat_b * rxx[n];
      if (!(k.aopoa[n] + at[n] < a_min1)) {
        const double xx_b = -sat_b * sat[n] * rxx[n];
        b.aopoa[n] += xx_b;
        at_b = xx_b;
      }
    } else {
      sat0_b[n] += sat_b;
    }
    // at = a_soa + beta del; del = g_soa - sat0 a_soa
    a_soa_b[n] += at_b;
    beta_b[n] += at_b * del[n];
    del_b += at_b * beta[n];
    g_soa_b += del_b;
    sat0_b[n] -= del_b * a_soa[n];
    a_soa_b[n] -= del_b * sat0[n];
    // beta = dtcur uk
    dtcur_b += beta_b[n] * k.uk[n];
    b.uk[n] += beta_b[n] * dtcur;
  }
  // tcur = last ? dtfull : tcur + dtcur; dtcur = last ? dtfull - tcur : alpha / tmpa
  double tcur_b = 0.0, tmpa_b = 0.0;
  if (last) {
    tcur_b -= dtcur_b;
  } else {
    tcur_b += b.tcur;
    dtcur_b += b.tcur;
    tmpa_b -= dtcur_b * dtcur * rtmpa;
  }
  const double tmpb_b = cmax_pass(0.0, tmpb, tmpa_b);
  // tot = g_soa + sum a_soa
  g_soa_b += tot_b;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    if (k.skip[n]) continue;
    a_soa_b[n] += tot_b;
    // term = uk |phi|
    b.uk[n] += tmpb_b * fabs(phi[n]);
    const double phi_b = tmpb_b * k.uk[n] * (phi[n] < 0 ? -1.0 : 1.0);
    // phi = (g_soa - gstar) / kmax(g_soa, kmax(gstar, g_min1))
    const double nb = phi_b * rm[n];
    const double md_b = -nb * phi[n];
    g_soa_b += nb;
    double gst_b = -nb;
    const double m2 = kmax(gst[n], g_min1);
    if (g_soa0 < m2) {
      gst_b += maxc_pass(gst[n], g_min1, md_b);
    } else {
      g_soa_b += md_b;
    }
    // gstar = sat0 a_soa
    sat0_b[n] += gst_b * a_soa[n];
    a_soa_b[n] += gst_b * sat0[n];
    // sat0 = g0 / kmax(aopoa + a_soa, a_min1)
    b.g0 += sat0_b[n] * ryy[n];
    if (!(k.aopoa[n] + a_soa[n] < a_min1)) {
      const double yy_b = -sat0_b[n] * sat0[n] * ryy[n];
      b.aopoa[n] += yy_b;
      a_soa_b[n] += yy_b;
    }
  }
  // a_soa = kmax(a, 0); g_soa = kmax(gprv, 0)
#pragma unroll
  for (int n = 0; n < 4; ++n) b.a[n] = k.skip[n] ? 0.0 : maxc_pass(c.a[n], 0.0, a_soa_b[n]);
  gprv_b += maxc_pass(gprv, 0.0, g_soa_b);
  b.g = gprv_b;
  b.tcur = tcur_b;
  // avg and dtsum enter linearly: their adjoints pass unchanged
}

// The loop's outputs -> the adjoint state after its last tick: from the
// adjoints of the final g, a[4] and qgas_avg = kmax(0, avg / dtsum) (avg and
// dtsum of the final state).
__host__ __device__ __forceinline__ void soa_rev_seed(double fin_avg, double fin_dtsum,
                                                      double g_b, const double a_b[4],
                                                      double gavg_b, SoaBar &b) {
  b.g = g_b;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    b.a[n] = a_b[n];
    b.uk[n] = 0.0;
    b.aopoa[n] = 0.0;
  }
  b.g0 = 0.0;
  b.tcur = 0.0;
  const double q = fin_avg / fin_dtsum;
  const double q_b = cmax_pass(0.0, q, gavg_b);
  b.avg = q_b != 0.0 ? q_b / fin_dtsum : 0.0;
  b.dtsum = q_b != 0.0 ? -q_b * q / fin_dtsum : 0.0;
}

// soa_init: the adjoint state before the first tick -> the adjoints of the
// loop's inputs, accumulated into in_b (uk = 0.81 uptk unless skipped)
__host__ __device__ __forceinline__ void soa_rev_init(const SoaState<double> &k,
                                                      const SoaBar &b, SoaIn<double> &in_b) {
  in_b.g += b.g;
  in_b.g0 += b.g0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    in_b.a[n] += b.a[n];
    in_b.aopoa[n] += b.aopoa[n];
    if (!k.skip[n]) in_b.uptk[n] += b.uk[n] * 0.81;
  }
}

// soa_init + the loop's outputs: from the adjoints of the final g, a[4] and
// qgas_avg, the adjoints of the loop's inputs.  ck: the states the ticks
// started from, in order; nt: the number of ticks.
template <class Ckpts>
__host__ __device__ __forceinline__ void soa_loop_rev(const SoaIn<double> &in, double dtfull,
                                                      const SoaState<double> &fin,
                                                      const Ckpts &ck, int nt, double g_b,
                                                      const double a_b[4], double gavg_b,
                                                      SoaIn<double> &in_b) {
  SoaState<double> k;
  soa_init(k, in);
  SoaBar b;
  soa_rev_seed(fin.avg, fin.dtsum, g_b, a_b, gavg_b, b);
  for (int i = nt - 1; i >= 0; --i) soa_tick_rev(k, ck(i), dtfull, b);
  soa_rev_init(k, b, in_b);
}

// ------------------------------------------------------------------ H2SO4
// h2so4_update: outputs qgas (q3), qgas_avg (q4), qso4[4]; inputs netprod,
// uptk[4], qgas (q1), qso4[4].  qso4_b holds the outputs' adjoints on entry
// and the inputs' on return.
__host__ __device__ __forceinline__ void h2so4_update_rev(double dt, double netprod,
                                                          const double uptk[4], double q1,
                                                          double q3_b, double q4_b,
                                                          double qso4_b[4], double &netprod_b,
                                                          double uptk_b[4], double &q1_b) {
  double tmpa = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) tmpa += uptk[n];
  const double kxt = tmpa * dt;
  const double pxt = netprod * dt;
  double kxt_b = 0.0, pxt_b = 0.0, tmpa_b = 0.0;
  if (kxt >= 1.0e-20) {
    double q3;
    if (kxt > 0.001) {
      const double pok = pxt / kxt;
      const double e = exp(-kxt);
      q3 = (q1 - pok) * e + pok;
    } else {
      const double k2 = kxt * kxt;
      q3 = q1 * (1.0 - kxt + k2 * 0.5) + pxt * (1.0 - kxt * 0.5 + k2 / 6.0);
    }
    const double del = (q1 + pxt) - q3;
    double del_b = 0.0;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      if (uptk[n] <= 0.0) continue;
      del_b += qso4_b[n] * (uptk[n] / tmpa);
      uptk_b[n] += qso4_b[n] * del / tmpa;
      tmpa_b -= qso4_b[n] * del * uptk[n] / (tmpa * tmpa);
    }
    q1_b += del_b;
    pxt_b += del_b;
    q3_b -= del_b;
    if (kxt > 0.001) {
      const double pok = pxt / kxt;
      const double e = exp(-kxt);
      const double aa = (q1 - pok) * (1.0 - e);
      double pok_b = 0.0, e_b = 0.0;
      q1_b += q3_b * e;
      pok_b += q3_b * (1.0 - e);
      e_b += q3_b * (q1 - pok);
      const double aa_b = q4_b / kxt;
      kxt_b -= q4_b * (aa / kxt) / kxt;
      pok_b += q4_b;
      q1_b += aa_b * (1.0 - e);
      pok_b -= aa_b * (1.0 - e);
      e_b -= aa_b * (q1 - pok);
      kxt_b -= e_b * e;
      pxt_b += pok_b / kxt;
      kxt_b -= pok_b * pok / kxt;
    } else {
      const double k2 = kxt * kxt;
      q1_b += q3_b * (1.0 - kxt + k2 * 0.5);
      pxt_b += q3_b * (1.0 - kxt * 0.5 + k2 / 6.0);
      kxt_b += q3_b * (q1 * (-1.0 + kxt) + pxt * (-0.5 + kxt / 3.0));
      q1_b += q4_b * (1.0 - kxt * 0.5 + k2 / 6.0);
      pxt_b += q4_b * (0.5 - kxt / 6.0 + k2 / 24.0);
      kxt_b += q4_b * (q1 * (-0.5 + kxt / 3.0) + pxt * (-1.0 / 6.0 + kxt / 12.0));
    }
  } else {
    q1_b += q3_b + q4_b;
    pxt_b += q3_b + q4_b * 0.5;
  }
  tmpa_b += kxt_b * dt;
  netprod_b += pxt_b * dt;
#pragma unroll
  for (int n = 0; n < 4; ++n) uptk_b[n] += tmpa_b;
}

// ------------------------------------------------------------------ rename
// rename_ait_to_acc: ai/ac (and gi/gc) are the forward inputs; ai_b/ac_b
// hold the outputs' adjoints on entry and the inputs' on return; gi_b/gc_b
// accumulate.
__host__ __device__ __forceinline__ void rename_rev(const ChainConsts &K, bool cloudy,
                                                    const Aer<double> &ai, const double gi[7],
                                                    const Aer<double> &ac, const double gc[7],
                                                    Aer<double> &ai_b, double gi_b[7],
                                                    Aer<double> &ac_b, double gc_b[7]) {
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
  const double b4_dv0 = b4_dryvol;
  const double b4_bnd = kmax(b4_dryvol, smallest);
  double nsum = ai.n[NAIT];
  if (cloudy) nsum = nsum + ac.n[NAIT];
  const double b4_qnum = kmax(0.0, nsum);
  const double lo = b4_bnd * K.rn_v2n_hi_rlx, hi = b4_bnd * K.rn_v2n_lo_rlx;
  const double b4_qnum_bnd = mmb(lo, hi, b4_qnum);
  double dv = dvi;
  if (cloudy) dv = dv + dvc;
  const double aft_dryvol = b4_dryvol + dv;
  if (aft_dryvol <= smallest) return;
  const double pio6 = PI / 6;
  const double v0 = (b4_bnd / b4_qnum_bnd) / pio6;
  const double d0 = cbrt(v0) * K.dgn_from_vol[NAIT];
  double b4_diam = d0;
  const bool thr = b4_diam > K.rn_threshold;
  if (thr) {
    b4_dryvol = b4_dryvol * cube(K.rn_threshold / b4_diam);
    b4_diam = K.rn_threshold;
  }
  if ((aft_dryvol - b4_dryvol) <= 1.0e-6 * b4_bnd) return;
  const double v1 = (aft_dryvol / b4_qnum_bnd) / pio6;
  const double aft_diam = cbrt(v1) * K.dgn_from_vol[NAIT];
  if (aft_diam <= MODE_DNOM[NAIT]) return;
  const double lb4 = log(b4_diam), laft = log(aft_diam);
  const double u_bn = (K.rn_ln_cutoff - lb4) * K.rn_tail_fac;
  const double u_bv = (K.rn_ln_cutoff - (lb4 + K.rn_ln_tail_fac)) * K.rn_tail_fac;
  const double u_an = (K.rn_ln_cutoff - laft) * K.rn_tail_fac;
  const double u_av = (K.rn_ln_cutoff - (laft + K.rn_ln_tail_fac)) * K.rn_tail_fac;
  const double b4_tail_num = 0.5 * erfc(u_bn);
  const double b4_tail_vol = 0.5 * erfc(u_bv);
  const double aft_tail_num = 0.5 * erfc(u_an);
  const double aft_tail_vol = 0.5 * erfc(u_av);
  const double vfrac = aft_tail_vol * aft_dryvol - b4_tail_vol * b4_dryvol;
  if (vfrac <= 0.0) return;
  const double xv1 = kmin(vfrac, aft_dryvol) / aft_dryvol;
  const double xv = kmin(xv1, K.xferfrac_max);
  const double x0 = aft_tail_num - b4_tail_num;
  const double xm = kmin(x0, xv);
  const double xn = kmax(0.0, xm);

  // ---- reverse of the transfers ----
  double xv_b = 0.0, xn_b = 0.0;
  {
    const double nt_b = ai_b.n[NACC] - ai_b.n[NAIT];
    xn_b += nt_b * ai.n[NAIT];
    ai_b.n[NAIT] += nt_b * xn;
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      if (!live(a, NAIT)) continue;
      const double vt_b = ai_b.q[a][NACC] - ai_b.q[a][NAIT];
      xv_b += vt_b * ai.q[a][NAIT];
      ai_b.q[a][NAIT] += vt_b * xv;
    }
  }
  if (cloudy) {
    const double nt_b = ac_b.n[NACC] - ac_b.n[NAIT];
    xn_b += nt_b * ac.n[NAIT];
    ac_b.n[NAIT] += nt_b * xn;
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      if (!live(a, NAIT)) continue;
      const double vt_b = ac_b.q[a][NACC] - ac_b.q[a][NAIT];
      xv_b += vt_b * ac.q[a][NAIT];
      ac_b.q[a][NAIT] += vt_b * xv;
    }
  }
  // xn = kmax(0, kmin(x0, xv))
  double x0_b = 0.0;
  const double xm_b = cmax_pass(0.0, xm, xn_b);
  min_rev(x0, xv, xm_b, x0_b, xv_b);
  // xv = kmin(kmin(vfrac, aft) / aft, xferfrac_max)
  const double xv1_b = minc_pass(xv1, K.xferfrac_max, xv_b);
  double aft_b = 0.0, vfrac_b = 0.0;
  {
    const double mm_b = xv1_b / aft_dryvol;
    aft_b -= xv1_b * xv1 / aft_dryvol;
    min_rev(vfrac, aft_dryvol, mm_b, vfrac_b, aft_b);
  }
  // vfrac = atv aft - btv b4dv; x0 = atn - btn
  const double atv_b = vfrac_b * aft_dryvol;
  aft_b += vfrac_b * aft_tail_vol;
  const double btv_b = -vfrac_b * b4_dryvol;
  double b4dv_b = -vfrac_b * b4_tail_vol;
  const double atn_b = x0_b, btn_b = -x0_b;
  // tails: 0.5 erfc(u), du/dl = -tail_fac
  const double c = K.rn_tail_fac * 0.5641895835477563;  // 1/sqrt(pi)
  const double laft_b = (atn_b * exp(-u_an * u_an) + atv_b * exp(-u_av * u_av)) * c;
  const double lb4_b = (btn_b * exp(-u_bn * u_bn) + btv_b * exp(-u_bv * u_bv)) * c;
  const double adiam_b = laft_b / aft_diam;
  double bdiam_b = lb4_b / b4_diam;
  // aft_diam = cbrt(v1) c, v1 = aft / qbnd / pio6
  double qbnd_b = 0.0, bnd_b = 0.0, dv0_b = 0.0;
  {
    const double v1_b = adiam_b * aft_diam / (3.0 * v1);
    aft_b += v1_b / (b4_qnum_bnd * pio6);
    qbnd_b -= v1_b * v1 / b4_qnum_bnd;
  }
  // threshold: b4_dryvol = b4_dv0 (thr / d0)^3, b4_diam = thr
  double d0_b = 0.0;
  if (thr) {
    dv0_b += b4dv_b * cube(K.rn_threshold / d0);
    d0_b -= 3.0 * b4dv_b * b4_dryvol / d0;
    bdiam_b = 0.0;
  } else {
    dv0_b += b4dv_b;
    d0_b += bdiam_b;
  }
  {
    const double v0_b = d0_b * d0 / (3.0 * v0);
    bnd_b += v0_b / (b4_qnum_bnd * pio6);
    qbnd_b -= v0_b * v0 / b4_qnum_bnd;
  }
  // aft_dryvol = b4_dv0 + dv
  dv0_b += aft_b;
  const double dv_b = aft_b;
  // qbnd = mmb(b4_bnd hi_rlx, b4_bnd lo_rlx, b4_qnum)
  double qn_b = 0.0;
  {
    const double m = kmin(hi, b4_qnum);
    if (lo < m) {
      if (b4_qnum < hi)
        qn_b += qbnd_b;
      else
        bnd_b += qbnd_b * K.rn_v2n_lo_rlx;
    } else {
      bnd_b += qbnd_b * K.rn_v2n_hi_rlx;
    }
  }
  const double nsum_b = cmax_pass(0.0, nsum, qn_b);
  ai_b.n[NAIT] += nsum_b;
  if (cloudy) ac_b.n[NAIT] += nsum_b;
  dv0_b += maxc_pass(b4_dv0, smallest, bnd_b);
  // b4_dv0 = (vi - dvi) + (vc - dvc); dv = dvi + dvc
#pragma unroll
  for (int a = 0; a < 7; ++a) {
    if (!live(a, NAIT)) continue;
    ai_b.q[a][NAIT] += dv0_b * mass_2_vol(a);
    gi_b[a] += (dv_b - dv0_b) * mass_2_vol(a);
    if (cloudy) {
      ac_b.q[a][NAIT] += dv0_b * mass_2_vol(a);
      gc_b[a] += (dv_b - dv0_b) * mass_2_vol(a);
    }
  }
}

// ------------------------------------------------------------------ newnuc
// Coefficients of the Vehkamaki (2002) fits as p0 + p1 t + p2 t^2 + p3 t^3
// + p4 / x, in the order of vehk2002 (rate a..j, then cluster size A..J).
#define MAM4X_VEHK_COEF                                                       \
  { \
        {0.14309, 2.21956, -0.0273911, 0.0000722811, 5.91822}, \
        {0.117489, 0.462532, -0.0118059, 0.0000404196, 15.7963}, \
        {-0.215554, -0.0810269, 0.00143581, -4.7758e-6, -2.91297}, \
        {-3.58856, 0.049508, -0.00021382, 3.10801e-7, -0.0293333}, \
        {1.14598, -0.600796, 0.00864245, -0.0000228947, -8.44985}, \
        {2.15855, 0.0808121, -0.000407382, -4.01957e-7, 0.721326}, \
        {1.6241, -0.0160106, 0.0000377124, 3.21794e-8, -0.0113255}, \
        {9.71682, -0.115048, 0.000157098, 4.00914e-7, 0.71186}, \
        {-1.05611, 0.00903378, -0.0000198417, 2.46048e-8, -0.0579087}, \
        {-0.148712, 0.00283508, -9.24619e-6, 5.00427e-9, -0.0127081}, \
        {-0.00295413, -0.0976834, 0.00102485, -2.18646e-6, -0.101717}, \
        {-0.00205064, -0.00758504, 0.000192654, -6.7043e-7, -0.255774}, \
        {0.00322308, 0.000852637, -0.0000154757, 5.66661e-8, 0.0338444}, \
        {0.0474323, -0.000625104, 2.65066e-6, -3.67471e-9, -0.000267251}, \
        {-0.0125211, 0.00580655, -0.000101674, 2.88195e-7, 0.0942243}, \
        {-0.038546, -0.000672316, 2.60288e-6, 1.19416e-8, -0.00851515}, \
        {-0.0183749, 0.000172072, -3.71766e-7, -5.14875e-10, 0.00026866}, \
        {-0.0619974, 0.000906958, -9.11728e-7, -5.36796e-9, -0.00774234}, \
        {0.0121827, -0.00010665, 2.5346e-7, -3.63519e-10, 0.000610065}, \
        {0.000320184, -0.0000174762, 6.06504e-8, -1.4177e-11, 0.000135751}}

// vehk2002: from the adjoints of rateloge, cnum_tot and x, those of c
// (so4vol), t and rh.
__host__ __device__ __forceinline__ void vehk2002_rev(double cc, double t, double rh,
                                                      double rl_b, double cn_b, double x_b,
                                                      double &c_b, double &t_b, double &rh_b) {
  double rateloge, cnum_tot, x;
  vehk2002(cc, t, rh, rateloge, cnum_tot, x);
  const double V[20][5] = MAM4X_VEHK_COEF;
  const double lnN = log(cc), lnrh = log(rh);
  const double r2 = sq(lnrh), r3 = cube(lnrh), N2 = sq(lnN), N3 = cube(lnN);
  // monomials m_k(lnrh, lnN) and their partials
  const double m[10] = {1.0, lnrh, r2, r3, lnN, lnrh * lnN, r2 * lnN, N2, lnrh * N2, N3};
  const double mr[10] = {0.0, 1.0, 2.0 * lnrh, 3.0 * r2, 0.0, lnN, 2.0 * lnrh * lnN, 0.0, N2,
                         0.0};
  const double mn[10] = {0.0, 0.0, 0.0, 0.0, 1.0, lnrh, r2, 2.0 * lnN, 2.0 * lnrh * lnN,
                         3.0 * N2};
  double lnrh_b = 0.0, lnN_b = 0.0, xx_b = x_b;
  const double s_b = cn_b * cnum_tot;  // cnum_tot = exp(sum L_k m_k)
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    const double *p = V[k], *q = V[k + 10];
    const double Kk = p[0] + p[1] * t + p[2] * t * t + p[3] * t * t * t + p[4] / x;
    const double Lk = q[0] + q[1] * t + q[2] * t * t + q[3] * t * t * t + q[4] / x;
    const double kb = rl_b * m[k], lb = s_b * m[k];
    t_b += kb * (p[1] + 2.0 * p[2] * t + 3.0 * p[3] * t * t) +
           lb * (q[1] + 2.0 * q[2] * t + 3.0 * q[3] * t * t);
    xx_b -= (kb * p[4] + lb * q[4]) / (x * x);
    lnrh_b += (rl_b * Kk + s_b * Lk) * mr[k];
    lnN_b += (rl_b * Kk + s_b * Lk) * mn[k];
  }
  // x = 0.740997 - 0.00266379 t - 0.00349998 lnN + 0.0000504022 t lnN + ...
  t_b += xx_b * (-0.00266379 + 0.0000504022 * lnN - 0.000183289 * lnrh - 0.0000179059 * r2 -
                 1.50345e-6 * r3);
  lnN_b += xx_b * (-0.00349998 + 0.0000504022 * t);
  lnrh_b += xx_b * (0.00201048 - 0.000183289 * t + 2.0 * lnrh * (0.00157407 - 0.0000179059 * t) +
                    3.0 * r2 * (0.000184403 - 1.50345e-6 * t));
  c_b += lnN_b / cc;
  rh_b += lnrh_b / rh;
}

// newnuc: h2so4, num_ait and so4_ait are in-out; their adjoints hold the
// outputs' on entry and the inputs' on return.
__host__ __device__ __forceinline__ void newnuc_rev(
    const ChainConsts &K, double deltat, double temp, double pmid, double zmid, double pblh,
    double relhum, double uptkrate_h2so4, double h2so4, double h2so4_avg, double num_ait,
    double so4_ait, double &temp_b, double &pmid_b, double &relhum_b, double &uptk_b,
    double &h2so4_b, double &h2so4_avg_b, double &num_ait_b, double &so4_ait_b) {
  (void)num_ait;
  (void)so4_ait;
  if (h2so4_avg <= 4.0e-16) return;
  const double relhumnn = kmax(0.01, kmin(0.99, relhum));
  const double pi = PI;
  const double rgas = 8.31446759100000, avogad = 6.022140000000000E+023;
  const double mw_so4a = 96, mw_so4a_host = 115;
  const double qh2so4_cur = h2so4;
  const double cair = pmid / (temp * rgas);
  const double so4vol_in = h2so4_avg * cair * avogad * 1.0e-6;
  double rateloge = K.ln_1em38;
  double cnum_h2so4 = 0.0, cnum_tot = 0.0, radius_cluster = 0.0, xv = 0.0;
  double temp_bb = 0.0, rh_bb = 0.0, so4vol_bb = 0.0;
  const bool vehk = so4vol_in >= 1.0e4;
  if (vehk) {
    temp_bb = kmax(230.15, kmin(305.15, temp));
    rh_bb = kmax(1.0e-4, kmin(1.0, relhumnn));
    so4vol_bb = kmax(1.0e4, kmin(1.0e11, so4vol_in));
    vehk2002(so4vol_bb, temp_bb, rh_bb, rateloge, cnum_tot, xv);
    cnum_h2so4 = cnum_tot * xv;
    radius_cluster = exp(-1.6524245 + 0.42316402 * xv + 0.3346648 * log(cnum_tot));
  }
  rateloge = rateloge + ::log(kmax(1.0e-38, 1.0));
  bool pbl = false;
  double tmp_ratenucl = 0.0;
  if (zmid <= kmax(pblh, 100.0)) {
    tmp_ratenucl = (1e-6 * so4vol_in) * 1.0;
    const double tmp_rateloge = log(kmax(1.0e-38, tmp_ratenucl));
    if (!(tmp_rateloge <= rateloge)) {
      pbl = true;
      rateloge = tmp_rateloge;
      radius_cluster = 0.5;
      const double tmp_diam = 0.5 * 2.0e-7;
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
    return;
  }
  const double voldry_part = (pi / 6.0) * ::pow(dpdry_part, 3);
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
  double wet_volfrac = 0.0, tmp_spd = 0.0, gr_kk = 0.0, dfin_kk = 0.0, dnuc_kk = 0.0;
  double gamma_kk = 0.0, up = 0.0, dif = 0.0, cs_prime_kk = 0.0, nu_kk = 0.0;
  if (igrow) {
    const double tw = temp, wvd = wetvol_dryvol;
    wet_volfrac = 1.0 / (wvd * (1.0 + mole_nh4_so4 * 17.0 / 98.0));
    tmp_spd = 14.7 * sqrt(tw);
    gr_kk = 3.0e-9 * tmp_spd * 96.0 * so4vol_in / (dens_part * wet_volfrac);
    dfin_kk = 1.0e9 * dpdry_part * pow(wvd, 1.0 / 3.0);
    dnuc_kk = kmax(2.0 * radius_cluster, 1.0);
    gamma_kk = 0.23 * pow(dnuc_kk, 0.2) * pow(dfin_kk / 3.0, 0.075) *
               pow(dens_part * 1.0e-3, -0.33) * pow(tw / 293.0, -0.75);
    up = kmax(uptkrate_h2so4 * 3600.0, 0.0);
    dif = (6.7037e-6 * pow(tw, 0.75) / cair) * 3600.0;
    cs_prime_kk = up / (4.0 * pi * dif * 0.65);
    nu_kk = gamma_kk * cs_prime_kk / gr_kk;
    factor_kk = exp((nu_kk / dfin_kk) - (nu_kk / dnuc_kk));
  }
  const double ratenuclt_kk = ratenuclt_bb * factor_kk;
  const double qm_num = ratenuclt_kk * deltat * mass_part;
  const double qmolso4a_del_max = kmax(0.0, qm_num) / (kgaero_per_moleso4a * cair);
  double freducea = 1.0;
  const bool fa_var = qmolso4a_del_max > qh2so4_cur;
  if (fa_var) freducea = qh2so4_cur / qmolso4a_del_max;
  double freduceb = 1.0;
  if (mole_nh4_so4 >= 1.0e-10) {
    if (qmolso4a_del_max * mole_nh4_so4 > 0.0)
      freduceb = 0.0 / (qmolso4a_del_max * mole_nh4_so4);
  }
  const double freduce = kmin(freducea, freduceb);
  if ((freduce * ratenuclt_kk) <= 1.0e-12) return;
  const double fq = freduce * qmolso4a_del_max;
  const double qh2so4_del = kmin(0.9999 * qh2so4_cur, fq);
  const double qnh3_del = kmin(0.9999 * 0.0, qh2so4_del * mole_nh4_so4);
  const double qso4a_del = qh2so4_del;
  const double qnuma_del = 1.0e-3 * (qso4a_del * mw_so4a + qnh3_del * 18) / mass_part;
  const double dndt0 = (qnuma_del * 1.0e3) / deltat;
  const double dmdt0 = kmax(0.0, ((qso4a_del * mw_so4a_host) / deltat));
  double dndt = dndt0, dmdt = dmdt0;
  int clamp = 0;  // 0 none, 1 below the detection limit, 2 mass1p low, 3 high
  if (dndt < 1.0e2) {
    dndt = 0.0;
    dmdt = 0.0;
    clamp = 1;
  } else {
    const double mass1p = dmdt / dndt;
    if (mass1p < K.mass1p_aitlo) {
      dndt = dmdt / K.mass1p_aitlo;
      clamp = 2;
    } else if (mass1p > K.mass1p_aithi) {
      dmdt = dndt * K.mass1p_aithi;
      clamp = 3;
    }
  }
  const double dso4dt = dmdt * 1.0 / mw_so4a_host;

  // ---- reverse ----
  double qcur_b = 0.0, dmdt_b = 0.0, dndt_b = 0.0;
  if (dso4dt > 0.0) {
    const double dq = dso4dt * deltat;
    double dq_b = so4_ait_b;
    // h2so4 -= kmin(dq, h2so4)
    const double dq2_b = -h2so4_b;
    if (h2so4 < dq)
      h2so4_b += dq2_b;
    else
      dq_b += dq2_b;
    dmdt_b += dq_b * deltat / mw_so4a_host;
  }
  dndt_b += num_ait_b * deltat;
  double dndt0_b = 0.0, dmdt0_b = 0.0;
  if (clamp == 0) {
    dndt0_b = dndt_b;
    dmdt0_b = dmdt_b;
  } else if (clamp == 2) {
    dmdt0_b = dmdt_b + dndt_b / K.mass1p_aitlo;
  } else if (clamp == 3) {
    dndt0_b = dndt_b + dmdt_b * K.mass1p_aithi;
  }
  double qso4a_b = cmax_pass(0.0, (qso4a_del * mw_so4a_host) / deltat, dmdt0_b) *
                   mw_so4a_host / deltat;
  const double qnuma_b = dndt0_b * 1.0e3 / deltat;
  qso4a_b += qnuma_b * 1.0e-3 * mw_so4a / mass_part;
  const double qnh3_b = qnuma_b * 1.0e-3 * 18 / mass_part;
  double qdel_b = qso4a_b;
  if (qh2so4_del * mole_nh4_so4 < 0.0) qdel_b += qnh3_b * mole_nh4_so4;
  // qh2so4_del = kmin(0.9999 q, freduce qmax)
  double freduce_b = 0.0, qmax_b = 0.0;
  if (fq < 0.9999 * qh2so4_cur) {
    freduce_b += qdel_b * qmolso4a_del_max;
    qmax_b += qdel_b * freduce;
  } else {
    qcur_b += qdel_b * 0.9999;
  }
  // freduce = kmin(freducea, freduceb); freduceb is a constant
  if (!(freduceb < freducea) && fa_var) {
    qcur_b += freduce_b / qmolso4a_del_max;
    qmax_b -= freduce_b * freducea / qmolso4a_del_max;
  }
  // qmax = kmax(0, rk dt mass) / (kg cair)
  double cair_b = -qmax_b * qmolso4a_del_max / cair;
  const double rk_b = cmax_pass(0.0, qm_num, qmax_b / (kgaero_per_moleso4a * cair)) * deltat *
                      mass_part;
  double rbb_b = rk_b * factor_kk;
  const double fac_b = rk_b * ratenuclt_bb;
  double rc_b = 0.0, so4vol_b = 0.0, tw_b = 0.0, wvd_b = 0.0;
  if (igrow) {
    const double e_b = fac_b * factor_kk;
    const double nu_b = e_b * (1.0 / dfin_kk - 1.0 / dnuc_kk);
    double dfin_b = -e_b * nu_kk / (dfin_kk * dfin_kk);
    double dnuc_b = e_b * nu_kk / (dnuc_kk * dnuc_kk);
    const double gamma_b = nu_b * cs_prime_kk / gr_kk;
    const double cs_b = nu_b * gamma_kk / gr_kk;
    const double gr_b = -nu_b * nu_kk / gr_kk;
    const double up_b = cs_b / (4.0 * pi * dif * 0.65);
    const double dif_b = -cs_b * cs_prime_kk / dif;
    tw_b += dif_b * dif * 0.75 / temp;
    cair_b -= dif_b * dif / cair;
    uptk_b += maxc_pass(uptkrate_h2so4 * 3600.0, 0.0, up_b) * 3600.0;
    dnuc_b += gamma_b * gamma_kk * 0.2 / dnuc_kk;
    dfin_b += gamma_b * gamma_kk * 0.075 / dfin_kk;
    tw_b += gamma_b * gamma_kk * (-0.75) / temp;
    rc_b += maxc_pass(2.0 * radius_cluster, 1.0, dnuc_b) * 2.0;
    wvd_b += dfin_b * dfin_kk / (3.0 * wetvol_dryvol);
    const double spd_b = gr_b * gr_kk / tmp_spd;
    so4vol_b += gr_b * gr_kk / so4vol_in;
    const double wvf_b = -gr_b * gr_kk / wet_volfrac;
    tw_b += spd_b * tmp_spd / (2.0 * temp);
    wvd_b -= wvf_b * wet_volfrac / wetvol_dryvol;
  }
  temp_b += tw_b;
  double rhnn_b = 0.0;
  {
    const double lt = log(tmpa0);
    const double t0_b = wvd_b * 0.56 / (lt * lt * tmpa0);
    rhnn_b += mmb_pass(0.10, 0.95, relhumnn, t0_b);
  }
  double rl_b = rbb_b * ratenuclt_bb;
  if (pbl) {
    so4vol_b += cmax_pass(1.0e-38, tmp_ratenucl, rl_b / tmp_ratenucl) * 1e-6;
  } else if (vehk) {
    double cn_b = 0.0, x_b = 0.0;
    {
      const double e = radius_cluster;
      const double ex_b = rc_b * e;
      x_b += ex_b * 0.42316402;
      cn_b += ex_b * 0.3346648 / cnum_tot;
    }
    double c_b = 0.0, tt_b = 0.0, r_b = 0.0;
    vehk2002_rev(so4vol_bb, temp_bb, rh_bb, rl_b, cn_b, x_b, c_b, tt_b, r_b);
    so4vol_b += mmb_pass(1.0e4, 1.0e11, so4vol_in, c_b);
    temp_b += mmb_pass(230.15, 305.15, temp, tt_b);
    rhnn_b += mmb_pass(1.0e-4, 1.0, relhumnn, r_b);
  }
  // so4vol = h2so4_avg cair avogad 1e-6; cair = pmid / (temp rgas)
  h2so4_avg_b += so4vol_b * cair * avogad * 1.0e-6;
  cair_b += so4vol_b * h2so4_avg * avogad * 1.0e-6;
  pmid_b += cair_b / (temp * rgas);
  temp_b -= cair_b * cair / temp;
  relhum_b += mmb_pass(0.01, 0.99, relhum, rhnn_b);
  h2so4_b += qcur_b;
}

// ------------------------------------------------------------------ coag
// Partials of h = nc fm / (nc + fm) (both regimes' coefficients positive).
__host__ __device__ __forceinline__ void hmean_partials(double nc, double fm, double &h_nc,
                                                        double &h_fm) {
  const double s = nc + fm;
  h_nc = fm * fm / (s * s);
  h_fm = nc * nc / (s * s);
}

// coag_pair: adjoints of temp, lamda and knc from those of ij0, ij3, ii0, jj0
// (diameters and densities held constant).
template <bool WANT_II, bool WANT_JJ>
__host__ __device__ __forceinline__ void coag_pair_rev(const ChainConsts &K,
                                                       const ChainConsts::CoagPair &P,
                                                       double temp, double lamda, double knc,
                                                       double dgatk, double dgacc, double rho_at,
                                                       double rho_ac, double ij0_b, double ij3_b,
                                                       double ii0_b, double jj0_b, double &temp_b,
                                                       double &lamda_b, double &knc_b) {
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
  const int n1 = (int)num::val(kmax(1.0, kmin(10.0, 1 + round(K.dlgsqt2 * log(rat))))) - 1;
  const double bm0ij = P.bm0ij[n1], bm3i = P.bm3i[n1];
  {
    const double ca = P.esat04 + r2 * P.esat16 * P.esac04;
    const double cb = P.esac04 + ri2 * P.esac16 * P.esat04;
    const double nc = knc * (2.0 + a * (kngat * ca + kngac * cb) + (r2 + ri2) * P.esat04 * P.esac04);
    const double fm = kfmatac * sqdgat * bm0ij *
                      (P.esat01 + r1 * P.esac01 + 2.0 * r2 * P.esat01 * P.esac04 +
                       rx4 * P.esat09 * P.esac16 + ri3 * P.esat16 * P.esac09 +
                       2.0 * ri1 * P.esat04 + P.esac01);
    const double h = nc * fm / (nc + fm);
    const double hb = cmax_pass(0.0, h, ij0_b);
    double h_nc, h_fm;
    hmean_partials(nc, fm, h_nc, h_fm);
    const double nc_b = hb * h_nc, fm_b = hb * h_fm;
    knc_b += nc_b * nc / knc;
    lamda_b += nc_b * knc * a * (2.0 / dgatk * ca + 2.0 / dgacc * cb);
    temp_b += fm_b * fm / (2.0 * temp);
  }
  {
    const double ca = P.esat16 + r2 * P.esat04 * P.esac04;
    const double cb = P.esat36 * P.esac04 + ri2 * P.esat64 * P.esac16;
    const double nc = knc * dgat3 *
                      (2.0 * P.esat36 + a * kngat * ca + a * kngac * cb +
                       r2 * P.esat16 * P.esac04 + ri2 * P.esat64 * P.esac04);
    const double fm = kfmatac * sqdgat7 * bm3i *
                      (P.esat49 + r1 * P.esat36 * P.esac01 + 2.0 * r2 * P.esat25 * P.esac04 +
                       rx4 * P.esat09 * P.esac16 + ri3 * P.esat100 * P.esac09 +
                       2.0 * ri1 * P.esat64 * P.esac01);
    const double qv12 = nc * fm / (nc + fm);
    const double den = cube(dgatk) * P.dumatk_fac;
    const double hb = cmax_pass(0.0, qv12 / den, ij3_b) / den;
    double h_nc, h_fm;
    hmean_partials(nc, fm, h_nc, h_fm);
    const double nc_b = hb * h_nc, fm_b = hb * h_fm;
    knc_b += nc_b * nc / knc;
    lamda_b += nc_b * knc * dgat3 * a * (2.0 / dgatk * ca + 2.0 / dgacc * cb);
    temp_b += fm_b * fm / (2.0 * temp);
  }
  if (WANT_II) {
    const double kfmat = sqrt(3.0 * boltzmann * temp / rho_at);
    const double cs = P.esat20 + P.esat04;
    const double nc = knc * (1.0 + P.esat08 + a * kngat * cs);
    const double fm = kfmat * sqdgat * P.bm0_at * (P.esat01 + P.esat25 + 2.0 * P.esat05);
    const double h = fm * nc / (fm + nc);
    const double hb = cmax_pass(0.0, h, ii0_b);
    double h_nc, h_fm;
    hmean_partials(nc, fm, h_nc, h_fm);
    const double nc_b = hb * h_nc, fm_b = hb * h_fm;
    knc_b += nc_b * nc / knc;
    lamda_b += nc_b * knc * a * 2.0 / dgatk * cs;
    temp_b += fm_b * fm / (2.0 * temp);
  }
  if (WANT_JJ) {
    const double kfmac = sqrt(3.0 * boltzmann * temp / rho_ac);
    const double cs = P.esac20 + P.esac04;
    const double nc = knc * (1.0 + P.esac08 + a * kngac * cs);
    const double fm = kfmac * sqdgac * P.bm0_ac * (P.esac01 + P.esac25 + 2.0 * P.esac05);
    const double h = fm * nc / (fm + nc);
    const double hb = cmax_pass(0.0, h, jj0_b);
    double h_nc, h_fm;
    hmean_partials(nc, fm, h_nc, h_fm);
    const double nc_b = hb * h_nc, fm_b = hb * h_fm;
    knc_b += nc_b * nc / knc;
    lamda_b += nc_b * knc * a * 2.0 / dgacc * cs;
    temp_b += fm_b * fm / (2.0 * temp);
  }
}

// qnum_intra_inter: adjoints of bijdtqnumj (x), biidt (y) and qbgn (q)
__host__ __device__ __forceinline__ void qnum_intra_inter_rev(double x, double y, double q,
                                                              double r_b, double &x_b,
                                                              double &y_b, double &q_b) {
  if (x < 1.0e-5) {
    const double f = 1.0 + 0.5 * x;
    const double dn = 1.0 + (x + y * q) * f;
    const double r = q / dn;
    const double dn_b = -r_b * r / dn;
    q_b += r_b / dn + dn_b * y * f;
    x_b += dn_b * (f + 0.5 * (x + y * q));
    y_b += dn_b * q * f;
    return;
  }
  const double e = exp(-x);
  const double u = (y * q / x) * (1.0 - e);
  const double dn = 1.0 + u;
  const double r = q * e / dn;
  const double dn_b = -r_b * r / dn;
  const double num_b = r_b / dn;
  q_b += num_b * e;
  double e_b = num_b * q;
  // u = (y q / x)(1 - e)
  const double yq = y * q / x;
  const double yq_b = dn_b * (1.0 - e);
  e_b -= dn_b * yq;
  y_b += yq_b * q / x;
  q_b += yq_b * y / x;
  x_b -= yq_b * yq / x;
  x_b -= e_b * e;
}

// coagulate: S is the forward input; S_b holds the outputs' adjoints on
// entry and the inputs' on return.  cin_so4_b / cin_soa_b: adjoints of the
// SO4 / SOA moved into the primary-carbon mode.
__host__ __device__ __forceinline__ void coagulate_rev(
    const ChainConsts &K, double deltat, double temp, double pmid, double aircon,
    const double dgn_awet[4], const double wetdens[4], const Aer<double> &S0, double cin_so4_b,
    double cin_soa_b, Aer<double> &S_b, double &temp_b, double &pmid_b, double &aircon_b) {
  double bgn[7][4], nb[4];
#pragma unroll
  for (int a = 0; a < 7; ++a)
#pragma unroll
    for (int m = 0; m < 4; ++m) bgn[a][m] = live(a, m) ? kmax(0.0, S0.q[a][m]) : 0.0;
#pragma unroll
  for (int m = 0; m < 4; ++m) nb[m] = kmax(0.0, S0.n[m]);
  const double t0 = 273.15 + 15.0;
  const double sqrt_temp = sqrt(temp);
  const double lamda = 6.6328e-8 * 101325.0 * temp / (t0 * pmid);
  const double amu = 1.458e-6 * temp * sqrt_temp / (temp + 110.4);
  const double knc = (2.0 / 3.0) * 1.3806500000000000e-023 * temp / amu;
  double wij0[3], wij3[3], wii0a = 0, wjj0a = 0, wii0p = 0, dummy = 0;
  coag_pair<double, true, true>(K, K.pair[0], temp, lamda, knc, dgn_awet[NAIT],
                                dgn_awet[NACC], wetdens[NAIT], wetdens[NACC], wij0[0], wij3[0],
                                wii0a, wjj0a);
  coag_pair<double, true, false>(K, K.pair[1], temp, lamda, knc, dgn_awet[NPCA],
                                 dgn_awet[NACC], wetdens[NPCA], wetdens[NACC], wij0[1], wij3[1],
                                 wii0p, dummy);
  coag_pair<double, false, false>(K, K.pair[2], temp, lamda, knc, dgn_awet[NAIT],
                                  dgn_awet[NPCA], wetdens[NAIT], wetdens[NPCA], wij0[2],
                                  wij3[2], dummy, dummy);
  double ij0[3], ij3[3];
#pragma unroll
  for (int ip = 0; ip < 3; ++ip) {
    ij0[ip] = wij0[ip] * aircon;
    ij3[ip] = wij3[ip] * aircon;
  }
  const double ii0a = wii0a * aircon, jj0a = wjj0a * aircon, ii0p = wii0p * aircon;
  // number update
  const double dacc = 1.0 + jj0a * deltat * nb[NACC];
  const double n_acc = nb[NACC] / dacc;
  const double tavg_acc = (nb[NACC] + n_acc) * 0.5;
  const double bijdt_p0 = deltat * ij0[1] * tavg_acc, biidt_p0 = deltat * ii0p;
  const double bijdt_p = kmax(0.0, bijdt_p0), biidt_p = kmax(0.0, biidt_p0);
  const double n_pca = qnum_intra_inter(bijdt_p, biidt_p, nb[NPCA]);
  const double tavg_pca = (nb[NPCA] + n_pca) * 0.5;
  double bij = ij0[0] * tavg_acc;
  bij = bij + ij0[2] * tavg_pca;
  const double bijdt_a0 = deltat * bij, biidt_a0 = deltat * ii0a;
  const double bijdt_a = kmax(0.0, bijdt_a0), biidt_a = kmax(0.0, biidt_a0);
  // mass update, block 1 (Aitken -> accumulation and primary carbon)
  const double epsx2 = 2.220446049250313e-16 * 2.0;
  const double b1 = kmax(0.0, ij3[0] * tavg_acc), b2 = kmax(0.0, ij3[2] * tavg_pca);
  const double dc1 = b1 + b2, df1 = deltat * dc1;
  const bool blk1 = df1 > epsx2;
  const double dc2 = kmax(0.0, ij3[1] * tavg_acc), df2 = deltat * dc2;
  const bool blk2 = df2 > epsx2;

  // ---- reverse ----
  double ij0_b[3] = {0, 0, 0}, ij3_b[3] = {0, 0, 0};
  double ii0a_b = 0, jj0a_b = 0, ii0p_b = 0, tacc_b = 0, tpca_b = 0;
  double bgn_b[7][4], nb_b[4];
#pragma unroll
  for (int a = 0; a < 7; ++a)
#pragma unroll
    for (int m = 0; m < 4; ++m) bgn_b[a][m] = live(a, m) ? S_b.q[a][m] : 0.0;
  if (blk2) {
    const double xf = 1.0 - exp(-df2);
    double xf_b = 0.0;
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      if (!live(a, NPCA)) continue;
      const double dq_b = S_b.q[a][NACC] - S_b.q[a][NPCA];
      xf_b += dq_b * bgn[a][NPCA];
      bgn_b[a][NPCA] += dq_b * xf;
    }
    const double dc_b = xf_b * exp(-df2) * deltat;
    const double x_b = cmax_pass(0.0, ij3[1] * tavg_acc, dc_b);
    ij3_b[1] += x_b * tavg_acc;
    tacc_b += x_b * ij3[1];
  }
  if (blk1) {
    const double prtn2 = b2 / dc1, prtn1 = 1.0 - prtn2;
    const double xf = 1.0 - exp(-df1);
    double xf_b = 0.0, p1_b = 0.0, p2_b = 0.0;
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      if (!live(a, NAIT)) continue;
      const double dq = xf * bgn[a][NAIT];
      double cin_b = 0.0;
      if (a == A_SO4) cin_b = cin_so4_b;
      if (a == A_SOA) cin_b = cin_soa_b;
      const double dq_b = -S_b.q[a][NAIT] + S_b.q[a][NACC] * prtn1 +
                          (S_b.q[a][NPCA] + cin_b) * prtn2;
      p1_b += S_b.q[a][NACC] * dq;
      p2_b += (S_b.q[a][NPCA] + cin_b) * dq;
      xf_b += dq_b * bgn[a][NAIT];
      bgn_b[a][NAIT] += dq_b * xf;
    }
    p2_b -= p1_b;
    double b2_b = p2_b / dc1;
    double dc_b = -p2_b * prtn2 / dc1;
    dc_b += xf_b * exp(-df1) * deltat;
    double b1_b = dc_b;
    b2_b += dc_b;
    const double x1_b = cmax_pass(0.0, ij3[0] * tavg_acc, b1_b);
    ij3_b[0] += x1_b * tavg_acc;
    tacc_b += x1_b * ij3[0];
    const double x2_b = cmax_pass(0.0, ij3[2] * tavg_pca, b2_b);
    ij3_b[2] += x2_b * tavg_pca;
    tpca_b += x2_b * ij3[2];
  }
  // number update, in reverse: Aitken, primary carbon, accumulation
#pragma unroll
  for (int m = 0; m < 4; ++m) nb_b[m] = 0.0;
  {
    double x_b = 0, y_b = 0;
    qnum_intra_inter_rev(bijdt_a, biidt_a, nb[NAIT], S_b.n[NAIT], x_b, y_b, nb_b[NAIT]);
    const double bij_b = cmax_pass(0.0, bijdt_a0, x_b) * deltat;
    ii0a_b += cmax_pass(0.0, biidt_a0, y_b) * deltat;
    ij0_b[0] += bij_b * tavg_acc;
    tacc_b += bij_b * ij0[0];
    ij0_b[2] += bij_b * tavg_pca;
    tpca_b += bij_b * ij0[2];
  }
  {
    // tavg_pca = (nb + n_pca) 0.5
    nb_b[NPCA] += tpca_b * 0.5;
    const double npca_b = S_b.n[NPCA] + tpca_b * 0.5;
    double x_b = 0, y_b = 0;
    qnum_intra_inter_rev(bijdt_p, biidt_p, nb[NPCA], npca_b, x_b, y_b, nb_b[NPCA]);
    const double bx_b = cmax_pass(0.0, bijdt_p0, x_b);
    ij0_b[1] += bx_b * deltat * tavg_acc;
    tacc_b += bx_b * deltat * ij0[1];
    ii0p_b += cmax_pass(0.0, biidt_p0, y_b) * deltat;
  }
  {
    nb_b[NACC] += tacc_b * 0.5;
    const double nacc_b = S_b.n[NACC] + tacc_b * 0.5;
    // n_acc = nb / (1 + jj0a dt nb)
    nb_b[NACC] += nacc_b / dacc - nacc_b * n_acc * jj0a * deltat / dacc;
    jj0a_b -= nacc_b * n_acc * deltat * nb[NACC] / dacc;
  }
  nb_b[NCOA] += S_b.n[NCOA];
  // coefficients times aircon
  double wij0_b[3], wij3_b[3];
#pragma unroll
  for (int ip = 0; ip < 3; ++ip) {
    wij0_b[ip] = ij0_b[ip] * aircon;
    wij3_b[ip] = ij3_b[ip] * aircon;
    aircon_b += ij0_b[ip] * wij0[ip] + ij3_b[ip] * wij3[ip];
  }
  aircon_b += ii0a_b * wii0a + jj0a_b * wjj0a + ii0p_b * wii0p;
  double lamda_b = 0.0, knc_b = 0.0;
  coag_pair_rev<true, true>(K, K.pair[0], temp, lamda, knc, dgn_awet[NAIT], dgn_awet[NACC],
                            wetdens[NAIT], wetdens[NACC], wij0_b[0], wij3_b[0], ii0a_b * aircon,
                            jj0a_b * aircon, temp_b, lamda_b, knc_b);
  coag_pair_rev<true, false>(K, K.pair[1], temp, lamda, knc, dgn_awet[NPCA], dgn_awet[NACC],
                             wetdens[NPCA], wetdens[NACC], wij0_b[1], wij3_b[1],
                             ii0p_b * aircon, 0.0, temp_b, lamda_b, knc_b);
  coag_pair_rev<false, false>(K, K.pair[2], temp, lamda, knc, dgn_awet[NAIT], dgn_awet[NPCA],
                              wetdens[NAIT], wetdens[NPCA], wij0_b[2], wij3_b[2], 0.0, 0.0,
                              temp_b, lamda_b, knc_b);
  // knc = (2/3) kB T / amu; amu = 1.458e-6 T sqrt(T) / (T + 110.4);
  // lamda = c T / (t0 p)
  temp_b += knc_b * knc / temp;
  const double amu_b = -knc_b * knc / amu;
  temp_b += amu_b * amu * (1.5 / temp - 1.0 / (temp + 110.4));
  temp_b += lamda_b * lamda / temp;
  pmid_b -= lamda_b * lamda / pmid;
  // the clips at the start
#pragma unroll
  for (int a = 0; a < 7; ++a)
#pragma unroll
    for (int m = 0; m < 4; ++m)
      if (live(a, m)) S_b.q[a][m] = cmax_pass(0.0, S0.q[a][m], bgn_b[a][m]);
#pragma unroll
  for (int m = 0; m < 4; ++m) S_b.n[m] = cmax_pass(0.0, S0.n[m], nb_b[m]);
}

// ------------------------------------------------------------------ aging
// age_pcarbon: S is the forward input; S_b holds the outputs' adjoints on
// entry and the inputs' on return.  The condensation / coagulation
// increments only split the diagnostics, so they get no adjoint.
__host__ __device__ __forceinline__ void age_pcarbon_rev(const ChainConsts &K, unsigned nmono,
                                                         double dgn_a_pca, const Aer<double> &S,
                                                         Aer<double> &S_b) {
  const double vol_shell = S.q[A_SO4][NPCA] * K.so4_vol + S.q[A_SOA][NPCA] * K.fac_m2v_eqvhyg;
  double vol_core = 0.0;
#pragma unroll
  for (int a = 0; a < 7; ++a)
    if (valid(a, NPCA)) vol_core += S.q[a][NPCA] * mass_2_vol(a);
  const double t1 = vol_shell * dgn_a_pca * K.fac_volsfc;
  const double t2x = 6.0 * (nmono * 4.76e-10) * vol_core;
  const double t2 = kmax(t2x, 0.0);
  double xf;
  const bool full = t1 >= t2;
  if (full)
    xf = K.xferfrac_max;
  else
    xf = kmin(t1 / t2, K.xferfrac_max);
  double xf_b = 0.0;
  {
    const double nt_b = S_b.n[NACC] - S_b.n[NPCA];
    xf_b += nt_b * S.n[NPCA];
    S_b.n[NPCA] += nt_b * xf;
  }
#pragma unroll
  for (int a = 0; a < 7; ++a) {
    if (a == A_SOA || a == A_SO4 || a == A_NCL || a == A_DST) {
      if (!live(a, NPCA)) continue;
      S_b.q[a][NPCA] = S_b.q[a][NACC];  // all of it moved; the new value is 0
    } else if (a == A_POM || a == A_BC || a == A_MOM) {
      const double qt_b = S_b.q[a][NACC] - S_b.q[a][NPCA];
      xf_b += qt_b * S.q[a][NPCA];
      S_b.q[a][NPCA] += qt_b * xf;
    }
  }
  if (!full) {
    const double r = t1 / t2;
    const double r_b = minc_pass(r, K.xferfrac_max, xf_b);
    const double t1_b = r_b / t2;
    const double t2_b = -r_b * r / t2;
    const double vc_b = maxc_pass(t2x, 0.0, t2_b) * 6.0 * (nmono * 4.76e-10);
    const double vs_b = t1_b * dgn_a_pca * K.fac_volsfc;
#pragma unroll
    for (int a = 0; a < 7; ++a)
      if (valid(a, NPCA)) S_b.q[a][NPCA] += vc_b * mass_2_vol(a);
    S_b.q[A_SO4][NPCA] += vs_b * K.so4_vol;
    S_b.q[A_SOA][NPCA] += vs_b * K.fac_m2v_eqvhyg;
  }
}

// ------------------------------------------------------------------ subarea
// One sub-step of clear_post (clear_post_sub, no diagnostics) over dtsub:
// g_in / S_in are the gases and aerosol it starts from, netprod the H2SO4
// net production from clear_post_begin.  S_b / g_b hold the outputs'
// adjoints on entry and the inputs' on return; netprod_b, uptk_b, the SOA
// loop results' adjoints and temp_b, pmid_b, relhum_b accumulate.
__host__ __device__ __forceinline__ void clear_post_sub_rev(
    const ChainConsts &K, const Cfg &cfg, double dtsub, double netprod, double temp,
    double pmid, double zmid, double pblh, double relhum, const double dgn_a[4],
    const double dgn_awet[4], const double wetdens[4], const double g_in[2],
    const Aer<double> &S_in, const double uptk[4], double soa_g, double soa_gavg,
    const double soa_a[4], Aer<double> &S_b, double g_b[2], double uptk_b[4],
    double &netprod_b, double &soa_g_b, double &soa_gavg_b, double soa_a_b[4],
    double &temp_b, double &pmid_b, double &relhum_b) {
  // ---- forward, keeping the state before each process ----
  const double aircon = pmid / (8314.46759100000 * temp);
  double g[2] = {g_in[0], g_in[1]};
  const double g_np1 = g[1];
  double uptkrate = 0.0, gavg[2];
  Aer<double> S = S_in;
  double dc_so4[4] = {0, 0, 0, 0}, dc_soa[4] = {0, 0, 0, 0}, so4_0[4] = {0, 0, 0, 0};
  if (cfg.do_cond) {
#pragma unroll
    for (int n = 0; n < 4; ++n) uptkrate += uptk[n];
    g[0] = soa_g;
    gavg[0] = soa_gavg;
    double so4[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) so4[n] = so4_0[n] = S.q[A_SO4][n];
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
  const Aer<double> S1 = S;  // after condensation
  double gi[7] = {0, 0, 0, 0, 0, 0, 0}, gc[7] = {0, 0, 0, 0, 0, 0, 0};
  gi[A_SO4] = dc_so4[NAIT];
  gi[A_SOA] = dc_soa[NAIT];
  Aer<double> none;
  if (cfg.do_rename) rename_ait_to_acc<Fp64>(K, false, S, gi, none, gc);
  const Aer<double> S2 = S;  // after renaming
  const double h_nn = g[1];
  if (cfg.do_newnuc)
    newnuc<Fp64>(K, dtsub, temp, pmid, zmid, pblh, relhum, uptkrate, g[1], gavg[1], S.n[NAIT],
                 S.q[A_SO4][NAIT]);
  const Aer<double> S3 = S;  // after nucleation
  double cin_so4 = 0.0, cin_soa = 0.0;
  if (cfg.do_coag)
    coagulate<Fp64>(K, dtsub, temp, pmid, aircon, dgn_awet, wetdens, S, cin_so4, cin_soa);
  // S: before aging

  // ---- reverse ----
  age_pcarbon_rev(K, cfg.nmono, dgn_a[NPCA], S, S_b);
  double aircon_b = 0.0;
  if (cfg.do_coag)
    coagulate_rev(K, dtsub, temp, pmid, aircon, dgn_awet, wetdens, S3, 0.0, 0.0, S_b, temp_b,
                  pmid_b, aircon_b);
  double gavg_b[2] = {0.0, 0.0}, uptkrate_b = 0.0;
  if (cfg.do_newnuc)
    newnuc_rev(K, dtsub, temp, pmid, zmid, pblh, relhum, uptkrate, h_nn, gavg[1], S2.n[NAIT],
               S2.q[A_SO4][NAIT], temp_b, pmid_b, relhum_b, uptkrate_b, g_b[1], gavg_b[1],
               S_b.n[NAIT], S_b.q[A_SO4][NAIT]);
  double gi_b[7] = {0, 0, 0, 0, 0, 0, 0}, gc_b[7] = {0, 0, 0, 0, 0, 0, 0};
  if (cfg.do_rename) {
    Aer<double> none_b;
    rename_rev(K, false, S1, gi, none, gc, S_b, gi_b, none_b, gc_b);
  }
  double dc_so4_b[4] = {0, 0, 0, 0}, dc_soa_b[4] = {0, 0, 0, 0};
  dc_so4_b[NAIT] += gi_b[A_SO4];
  dc_soa_b[NAIT] += gi_b[A_SOA];
  if (cfg.do_cond) {
    double so4_b[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      soa_a_b[n] += S_b.q[A_SOA][n] + dc_soa_b[n];
      S_b.q[A_SOA][n] = -dc_soa_b[n];
      so4_b[n] = S_b.q[A_SO4][n] + dc_so4_b[n];
    }
    double q1_b = 0.0;
    h2so4_update_rev(dtsub, netprod, uptk, g_np1, g_b[1], gavg_b[1], so4_b, netprod_b, uptk_b,
                     q1_b);
#pragma unroll
    for (int n = 0; n < 4; ++n) S_b.q[A_SO4][n] = so4_b[n] - dc_so4_b[n];
    soa_g_b += g_b[0];
    soa_gavg_b += gavg_b[0];
    g_b[0] = 0.0;
    g_b[1] = q1_b;
#pragma unroll
    for (int n = 0; n < 4; ++n) uptk_b[n] += uptkrate_b;
  } else {
    g_b[0] += gavg_b[0];
    g_b[1] += gavg_b[1];
  }
  pmid_b += aircon_b / (8314.46759100000 * temp);
  temp_b -= aircon_b * aircon / temp;
}

// clear_post_begin: g_b (H2SO4) and netprod_b -> the adjoints of the gases
// before it (g_b) and before gas chemistry (g1_b).
__host__ __device__ __forceinline__ void clear_post_begin_rev(const Cfg &cfg, double deltat,
                                                              double netprod, double netprod_b,
                                                              double g_b[2], double g1_b[2]) {
  if (cfg.do_cond && cfg.gaexch_h2so4_uptake_optaa == 2) {
    if (netprod >= 0) g1_b[1] += g_b[1];
    g_b[1] = netprod_b / deltat;
    g1_b[1] -= netprod_b / deltat;
  }
}

// clear_post (no diagnostics, one sub-step): S_in / g_in are its inputs;
// S_b / g_b hold the outputs' adjoints on entry and the inputs' on return.
__host__ __device__ __forceinline__ void clear_post_rev(
    const ChainConsts &K, const Cfg &cfg, double deltat, double temp, double pmid, double zmid,
    double pblh, double relhum, const double dgn_a[4], const double dgn_awet[4],
    const double wetdens[4], const double g1[2], const double g_in[2], const Aer<double> &S_in,
    const double uptk[4], double soa_g, double soa_gavg, const double soa_a[4],
    Aer<double> &S_b, double g_b[2], double g1_b[2], double uptk_b[4], double &soa_g_b,
    double &soa_gavg_b, double soa_a_b[4], double &temp_b, double &pmid_b, double &relhum_b) {
  double g[2] = {g_in[0], g_in[1]};
  const double netprod = clear_post_begin(cfg, deltat, g1, g);
  double netprod_b = 0.0;
  clear_post_sub_rev(K, cfg, (double)(int)(deltat / 1), netprod, temp, pmid, zmid, pblh, relhum,
                     dgn_a, dgn_awet, wetdens, g, S_in, uptk, soa_g, soa_gavg, soa_a, S_b, g_b,
                     uptk_b, netprod_b, soa_g_b, soa_gavg_b, soa_a_b, temp_b, pmid_b, relhum_b);
  clear_post_begin_rev(cfg, deltat, netprod, netprod_b, g_b, g1_b);
}

// ------------------------------------------------------------------ calcsize
// y = mmb(lo, hi, x) with variable bounds
__host__ __device__ __forceinline__ void mmb_rev(double lo, double hi, double x, double yb,
                                                 double &lo_b, double &hi_b, double &x_b) {
  const double m = kmin(hi, x);
  if (lo < m) {
    if (x < hi)
      x_b += yb;
    else
      hi_b += yb;
  } else {
    lo_b += yb;
  }
}

// adjust_num_sizes: adjoints of the clipped numbers ni/nc it starts from
// and of the dry volumes, from those of the numbers it ends with.
__host__ __device__ __forceinline__ void adjust_num_sizes_rev(double drv_i, double drv_c,
                                                              double ni0, double nc0, double dt,
                                                              double v2nmin, double v2nmax,
                                                              double tsinv, double ni_b,
                                                              double nc_b, double &ni0_b,
                                                              double &nc0_b, double &drvi_b,
                                                              double &drvc_b) {
  const double frac = kmax(0.0, kmin(1.0, dt * tsinv));
  const double minrl = v2nmin / 27.0, maxrl = v2nmax * 27.0;
  if (drv_i <= 0.0 && drv_c <= 0.0) return;
  if (drv_c <= 0.0) {
    ni0_b += ni_b * (1.0 - frac);
    double lo_b = 0, hi_b = 0;
    mmb_rev(drv_i * v2nmin, drv_i * v2nmax, ni0, ni_b * frac, lo_b, hi_b, ni0_b);
    drvi_b += lo_b * v2nmin + hi_b * v2nmax;
    return;
  }
  if (drv_i <= 0.0) {
    nc0_b += nc_b * (1.0 - frac);
    double lo_b = 0, hi_b = 0;
    mmb_rev(drv_c * v2nmin, drv_c * v2nmax, nc0, nc_b * frac, lo_b, hi_b, nc0_b);
    drvc_b += lo_b * v2nmin + hi_b * v2nmax;
    return;
  }
  // forward
  const double ni1 = ni0, nc1 = nc0;
  const double nb1 = mmb(drv_i * minrl, drv_i * maxrl, ni1);
  const double dni2 = (nb1 - ni1) * frac;
  double ni2 = ni1 + dni2;
  const double nb2 = mmb(drv_c * minrl, drv_c * maxrl, nc1);
  const double dnc2 = (nb2 - nc1) * frac;
  double nc2 = nc1 + dnc2;
  const bool ovr_i = dni2 == 0.0 && dnc2 != 0.0;
  const bool ovr_c = !ovr_i && dni2 != 0.0 && dnc2 == 0.0;
  if (ovr_i)
    ni2 = mmb(drv_i * minrl, drv_i * maxrl, ni1 - dnc2);
  else if (ovr_c)
    nc2 = mmb(drv_c * minrl, drv_c * maxrl, nc1 - dni2);
  const double tdrv = drv_i + drv_c, tnum = ni2 + nc2;
  const double minb = tdrv * v2nmin, maxb = tdrv * v2nmax;
  int kind = 0;  // 1 both, 2 cloud-borne, 3 interstitial
  int side = 0;  // 1 below the lower bound, 2 above the upper
  double d3 = 0.0;
  if (tnum < minb) {
    side = 1;
    d3 = (minb - tnum) * frac;
    if ((ni2 < drv_i * v2nmin) && (nc2 < drv_c * v2nmin))
      kind = 1;
    else if (nc2 < drv_c * v2nmin)
      kind = 2;
    else if (ni2 < drv_i * v2nmin)
      kind = 3;
  } else if (tnum > maxb) {
    side = 2;
    d3 = (maxb - tnum) * frac;
    if ((ni2 > drv_i * v2nmax) && (nc2 > drv_c * v2nmax))
      kind = 1;
    else if (nc2 > drv_c * v2nmax)
      kind = 2;
    else if (ni2 > drv_i * v2nmax)
      kind = 3;
  }
  // reverse
  double ni2_b = ni_b, nc2_b = nc_b, d3_b = 0.0, tnum_b = 0.0, tdrv_b = 0.0;
  if (kind == 1) {
    d3_b += ni_b * (ni2 / tnum) + nc_b * (nc2 / tnum);
    ni2_b += ni_b * d3 / tnum;
    nc2_b += nc_b * d3 / tnum;
    tnum_b -= (ni_b * d3 * ni2 + nc_b * d3 * nc2) / (tnum * tnum);
  } else if (kind == 2) {
    d3_b += nc_b;
  } else if (kind == 3) {
    d3_b += ni_b;
  }
  if (kind != 0) {
    const double bound_b = d3_b * frac;
    tnum_b -= d3_b * frac;
    tdrv_b += bound_b * (side == 1 ? v2nmin : v2nmax);
  }
  ni2_b += tnum_b;
  nc2_b += tnum_b;
  drvi_b += tdrv_b;
  drvc_b += tdrv_b;
  double ni1_b = 0, nc1_b = 0, dni2_b = 0, dnc2_b = 0;
  if (ovr_i) {
    double lo_b = 0, hi_b = 0, x_b = 0;
    mmb_rev(drv_i * minrl, drv_i * maxrl, ni1 - dnc2, ni2_b, lo_b, hi_b, x_b);
    drvi_b += lo_b * minrl + hi_b * maxrl;
    ni1_b += x_b;
    dnc2_b -= x_b;
  } else {
    ni1_b += ni2_b;
    dni2_b += ni2_b;
  }
  if (ovr_c) {
    double lo_b = 0, hi_b = 0, x_b = 0;
    mmb_rev(drv_c * minrl, drv_c * maxrl, nc1 - dni2, nc2_b, lo_b, hi_b, x_b);
    drvc_b += lo_b * minrl + hi_b * maxrl;
    nc1_b += x_b;
    dni2_b -= x_b;
  } else {
    nc1_b += nc2_b;
    dnc2_b += nc2_b;
  }
  {
    double lo_b = 0, hi_b = 0;
    mmb_rev(drv_i * minrl, drv_i * maxrl, ni1, dni2_b * frac, lo_b, hi_b, ni1_b);
    ni1_b -= dni2_b * frac;
    drvi_b += lo_b * minrl + hi_b * maxrl;
  }
  {
    double lo_b = 0, hi_b = 0;
    mmb_rev(drv_c * minrl, drv_c * maxrl, nc1, dnc2_b * frac, lo_b, hi_b, nc1_b);
    nc1_b -= dnc2_b * frac;
    drvc_b += lo_b * minrl + hi_b * maxrl;
  }
  ni0_b += ni1_b;
  nc0_b += nc1_b;
}

// fv / fn of one direction of the Aitken <-> accumulation exchange at
// ratio num_t / drv_t: adjoints of num_t and drv_t from those of fn and fv
// (zero where the fractions are clipped to 0 or 1).
__host__ __device__ __forceinline__ void xfer_frac_rev(double num_t, double drv_t, double v2n,
                                                       double gm, bool to_acc, double fn_b,
                                                       double fv_b, double &num_b,
                                                       double &drv_b) {
  const bool full = to_acc ? (num_t < drv_t * v2n) : (num_t > drv_t * v2n);
  if (full) return;
  const double fv = ((num_t / drv_t) - gm) / (v2n - gm);
  const double rr = drv_t * v2n / num_t;
  const double fn = fv * rr;
  if ((fn <= 0.0) || (fv <= 0.0) || (fn >= 1.0) || (fv >= 1.0)) return;
  const double fvt_b = fv_b + fn_b * rr;
  const double rr_b = fn_b * fv;
  drv_b += rr_b * v2n / num_t;
  num_b -= rr_b * rr / num_t;
  const double ratio_b = fvt_b / (v2n - gm);
  num_b += ratio_b / drv_t;
  drv_b -= ratio_b * (num_t / drv_t) / drv_t;
}

// calcsize (CalcSize, not the E3SM variant): adjoints of n_i, n_c, q_i, q_c
// (accumulated) from those of the tendencies o_b (diameters held constant).
__host__ __device__ __forceinline__ void calcsize_rev(const ChainConsts &K, bool do_xfer,
                                                      bool do_adjust, double dt,
                                                      const double n_i[4], const double n_c[4],
                                                      const double q_i[4][7],
                                                      const double q_c[4][7],
                                                      const CalcOut<double> &o_b,
                                                      double n_i_b[4], double n_c_b[4],
                                                      double q_i_b[4][7], double q_c_b[4][7]) {
  const double tsinv = safe_den(kmax(86400.0, dt));
  const double dtinv = safe_den(dt);
  double dv_i[4], dv_c[4], ni0[4], nc0[4], nk_i[4], nk_c[4];
  bool adj[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    double vi = 0, vc = 0;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      if (s >= NSPEC_MODE[m]) continue;
      vi += kmax(0.0, q_i[m][s]) * K.inv_density[m][s];
      vc += kmax(0.0, q_c[m][s]) * K.inv_density[m][s];
    }
    dv_i[m] = vi;
    dv_c[m] = vc;
    ni0[m] = n_i[m] < 0 ? 0.0 : n_i[m];
    nc0[m] = n_c[m] < 0 ? 0.0 : n_c[m];
    double ni = ni0[m], nc = nc0[m], d1, d2;
    const bool aa = (m == NACC || m == NAIT);
    adj[m] = do_adjust && !(aa && do_xfer);
    if (adj[m])
      adjust_num_sizes(vi, vc, n_i[m], n_c[m], dt, K.v2n_min[m], K.v2n_max[m], tsinv, ni, nc, d1,
                       d2);
    nk_i[m] = ni;
    nk_c[m] = nc;
  }
  double dv_i_b[4] = {0, 0, 0, 0}, dv_c_b[4] = {0, 0, 0, 0};
  double nk_i_b[4] = {0, 0, 0, 0}, nk_c_b[4] = {0, 0, 0, 0};
  if (do_xfer) {
    const double v2n_ait = K.v2n_nom[NAIT], v2n_acc = K.v2n_nom[NACC];
    const double gm = K.v2n_geomean;
    // Aitken -> accumulation
    bool a2c = false;
    double fn_a = 0, fv_a = 0;
    const double drv_ta = dv_i[NAIT] + dv_c[NAIT], num_ta = nk_i[NAIT] + nk_c[NAIT];
    if (drv_ta > 0.0 && num_ta < drv_ta * gm) {
      a2c = true;
      if (num_ta < drv_ta * v2n_acc) {
        fn_a = fv_a = 1.0;
      } else {
        fv_a = ((num_ta / drv_ta) - gm) / (v2n_acc - gm);
        fn_a = fv_a * (drv_ta * v2n_acc / num_ta);
        if ((fn_a <= 0.0) || (fv_a <= 0.0))
          fn_a = fv_a = 0.0;
        else if ((fn_a >= 1.0) || (fv_a >= 1.0))
          fn_a = fv_a = 1.0;
      }
    }
    // accumulation -> Aitken
    bool c2a = false;
    double fn_c = 0, fv_c = 0, fn_c1 = 0, num_t0 = 0, drv_t0 = 0, num_tc = 0, drv_tc = 0;
    double drv_i_noxf = 0, drv_c_noxf = 0, num_t_noxf = 0;
    drv_t0 = dv_i[NACC] + dv_c[NACC];
    num_tc = nk_i[NACC] + nk_c[NACC];
    drv_tc = drv_t0;
    bool noxf = false;
    if (drv_tc > 0.0 && num_tc > drv_tc * gm) {
      noxf = true;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (!(s == 1 || s == 3 || s == 4)) continue;
        drv_i_noxf += kmax(0.0, q_i[NACC][s]) * K.inv_density[NACC][s];
        drv_c_noxf += kmax(0.0, q_c[NACC][s]) * K.inv_density[NACC][s];
      }
      num_t_noxf = (drv_i_noxf + drv_c_noxf) * K.v2n_min[NACC];
      num_t0 = num_tc;
      num_tc = kmax(0.0, num_tc - num_t_noxf);
      drv_tc = kmax(0.0, drv_tc - (drv_i_noxf + drv_c_noxf));
    }
    if (drv_tc > 0.0 && num_tc > drv_tc * gm) {
      c2a = true;
      if (num_tc > drv_tc * v2n_ait) {
        fn_c1 = fv_c = 1.0;
      } else {
        fv_c = ((num_tc / drv_tc) - gm) / (v2n_ait - gm);
        fn_c1 = fv_c * (drv_tc * v2n_ait / num_tc);
        if ((fn_c1 <= 0.0) || (fv_c <= 0.0))
          fn_c1 = fv_c = 0.0;
        else if ((fn_c1 >= 1.0) || (fv_c >= 1.0))
          fn_c1 = fv_c = 1.0;
      }
      fn_c = fn_c1 * num_tc * safe_den(num_t0);
    }
    if (a2c || c2a) {
      const double xfv_a = fv_a * tsinv, xfn_a = fn_a * tsinv;
      const double xfv_c = fv_c * tsinv, xfn_c = fn_c * tsinv;
      double xfv_a_b = 0, xfn_a_b = 0, xfv_c_b = 0, xfn_c_b = 0;
      if (a2c) {
        const double xt00_b = o_b.dn_i[NACC] - o_b.dn_i[NAIT];
        const double xt01_b = o_b.dn_c[NACC] - o_b.dn_c[NAIT];
        nk_i_b[NAIT] += xt00_b * xfn_a;
        nk_c_b[NAIT] += xt01_b * xfn_a;
        xfn_a_b += xt00_b * nk_i[NAIT] + xt01_b * nk_c[NAIT];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int sa = i, sc = (i == 0 ? 0 : i == 1 ? 2 : i == 2 ? 5 : 6);
          const double ti_b = o_b.dq_i[NACC][sc] - o_b.dq_i[NAIT][sa];
          const double tc_b = o_b.dq_c[NACC][sc] - o_b.dq_c[NAIT][sa];
          q_i_b[NAIT][sa] += cmax_pass(0.0, q_i[NAIT][sa], ti_b * xfv_a);
          q_c_b[NAIT][sa] += cmax_pass(0.0, q_c[NAIT][sa], tc_b * xfv_a);
          xfv_a_b += ti_b * kmax(0.0, q_i[NAIT][sa]) + tc_b * kmax(0.0, q_c[NAIT][sa]);
        }
      }
      if (c2a) {
        const double xt10_b = o_b.dn_i[NAIT] - o_b.dn_i[NACC];
        const double xt11_b = o_b.dn_c[NAIT] - o_b.dn_c[NACC];
        nk_i_b[NACC] += xt10_b * xfn_c;
        nk_c_b[NACC] += xt11_b * xfn_c;
        xfn_c_b += xt10_b * nk_i[NACC] + xt11_b * nk_c[NACC];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int sa = i, sc = (i == 0 ? 0 : i == 1 ? 2 : i == 2 ? 5 : 6);
          const double ti_b = o_b.dq_i[NAIT][sa] - o_b.dq_i[NACC][sc];
          const double tc_b = o_b.dq_c[NAIT][sa] - o_b.dq_c[NACC][sc];
          q_i_b[NACC][sc] += cmax_pass(0.0, q_i[NACC][sc], ti_b * xfv_c);
          q_c_b[NACC][sc] += cmax_pass(0.0, q_c[NACC][sc], tc_b * xfv_c);
          xfv_c_b += ti_b * kmax(0.0, q_i[NACC][sc]) + tc_b * kmax(0.0, q_c[NACC][sc]);
        }
      }
      if (a2c) {
        double num_b = 0, drv_b = 0;
        xfer_frac_rev(num_ta, drv_ta, v2n_acc, gm, true, xfn_a_b * tsinv, xfv_a_b * tsinv, num_b,
                      drv_b);
        nk_i_b[NAIT] += num_b;
        nk_c_b[NAIT] += num_b;
        dv_i_b[NAIT] += drv_b;
        dv_c_b[NAIT] += drv_b;
      }
      if (c2a) {
        // fn = fn1 num_t safe_den(num_t0)
        const double fnf_b = xfn_c_b * tsinv;
        const double sd = safe_den(num_t0);
        const double fn1_b = fnf_b * num_tc * sd;
        double num_b = fnf_b * fn_c1 * sd, drv_b = 0;
        const double tol2 = sq(2.220446049250313e-16);
        const double num0_b =
            fnf_b * fn_c1 * num_tc * (tol2 - num_t0 * num_t0) / sq(num_t0 * num_t0 + tol2);
        xfer_frac_rev(num_tc, drv_tc, v2n_ait, gm, false, fn1_b, xfv_c_b * tsinv, num_b, drv_b);
        // num_t = kmax(0, num_t0 - noxf); drv_t = kmax(0, drv_t0 - drv_noxf)
        double n0_b = num0_b, d0_b = 0.0, dnoxf_b = 0.0;
        if (noxf) {
          const double nb = cmax_pass(0.0, num_t0 - num_t_noxf, num_b);
          n0_b += nb;
          dnoxf_b -= nb * K.v2n_min[NACC];
          const double db = cmax_pass(0.0, drv_t0 - (drv_i_noxf + drv_c_noxf), drv_b);
          d0_b += db;
          dnoxf_b -= db;
#pragma unroll
          for (int s = 0; s < 7; ++s) {
            if (!(s == 1 || s == 3 || s == 4)) continue;
            q_i_b[NACC][s] += cmax_pass(0.0, q_i[NACC][s], dnoxf_b * K.inv_density[NACC][s]);
            q_c_b[NACC][s] += cmax_pass(0.0, q_c[NACC][s], dnoxf_b * K.inv_density[NACC][s]);
          }
        } else {
          n0_b += num_b;
          d0_b += drv_b;
        }
        nk_i_b[NACC] += n0_b;
        nk_c_b[NACC] += n0_b;
        dv_i_b[NACC] += d0_b;
        dv_c_b[NACC] += d0_b;
      }
    }
  }
  // per mode: size adjustment, clips, dry volumes
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    double ni0_b = 0, nc0_b = 0;
    if (adj[m]) {
      // dqdt = (ni - n_i) dtinv, and ni also reaches the exchange as nk_i
      const double ni_b = nk_i_b[m] + o_b.dn_i[m] * dtinv;
      const double nc_b = nk_c_b[m] + o_b.dn_c[m] * dtinv;
      n_i_b[m] -= o_b.dn_i[m] * dtinv;
      n_c_b[m] -= o_b.dn_c[m] * dtinv;
      adjust_num_sizes_rev(dv_i[m], dv_c[m], ni0[m], nc0[m], dt, K.v2n_min[m], K.v2n_max[m],
                           tsinv, ni_b, nc_b, ni0_b, nc0_b, dv_i_b[m], dv_c_b[m]);
    } else {
      ni0_b = nk_i_b[m];
      nc0_b = nk_c_b[m];
    }
    if (!(n_i[m] < 0)) n_i_b[m] += ni0_b;
    if (!(n_c[m] < 0)) n_c_b[m] += nc0_b;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      if (s >= NSPEC_MODE[m]) continue;
      q_i_b[m][s] += cmax_pass(0.0, q_i[m][s], dv_i_b[m] * K.inv_density[m][s]);
      q_c_b[m][s] += cmax_pass(0.0, q_c[m][s], dv_c_b[m] * K.inv_density[m][s]);
    }
  }
}

// ------------------------------------------------------------------ cell
// Adjoints of the per-cell context (CellCtx) fields the step reads.
struct CtxBar {
  double T, p, rh_clear;
  double h1, s1, h2, s2, h3, s3;
  double fa_n[4], fc_n[4], fa_q[4], fc_q[4];
};

__host__ __device__ __forceinline__ void zero(CtxBar &b) {
  b.T = b.p = b.rh_clear = 0.0;
  b.h1 = b.s1 = b.h2 = b.s2 = b.h3 = b.s3 = 0.0;
#pragma unroll
  for (int m = 0; m < 4; ++m) b.fa_n[m] = b.fc_n[m] = b.fa_q[m] = b.fc_q[m] = 0.0;
}
__host__ __device__ __forceinline__ void zero(Aer<double> &b) {
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    b.n[m] = 0.0;
#pragma unroll
    for (int a = 0; a < 7; ++a) b.q[a][m] = 0.0;
  }
}
__host__ __device__ __forceinline__ void zero(ClearGas<double> &b) {
  b.g1[0] = b.g1[1] = b.g[0] = b.g[1] = 0.0;
  b.hcl = b.scl = b.hcd = b.scd = 0.0;
}

// adjoint of a clipped load kmax(0, state[f] c) into st_b[f]
__host__ __device__ __forceinline__ void ld3_rev(const ChainArgs &a, long off, int f, double c,
                                                 double yb, double st_b[NSTATE]) {
  st_b[f] += cmax_pass(0.0, a.state[f][off] * c, yb) * c;
}
// ld2: the pre-cloud-chemistry plane when given (not differentiated), else
// the state
template <bool PRE>
__host__ __device__ __forceinline__ void ld2_rev(const ChainArgs &a, long off, int f, double c,
                                                 double yb, double st_b[NSTATE]) {
  if (PRE && a.precld[f]) return;
  ld3_rev(a, off, f, c, yb, st_b);
}

// partition_factors: adjoints of qi and qc from those of fa and fc
__host__ __device__ __forceinline__ void partition_factors_rev(double qi, double qc,
                                                               double fcldy, double fclea,
                                                               double fa_b, double fc_b,
                                                               double &qi_b, double &qc_b) {
  const double qc_cldy = qc / fcldy;
  const double x1 = (qi + qc) - qc_cldy;
  const double qi_cldy = kmax(0.0, x1);
  const double qi_clea = (qi - fcldy * qi_cldy) / fclea;
  const double nn = kmax(1.e-35, qi_clea * fclea), dd = kmax(1.e-35, qi);
  const double r0 = nn / dd;
  const double r_b = fa_b / fclea - fc_b / fcldy;
  const double r0_b = mmb_pass(0.0, 1.0, r0, r_b);
  const double nn_b = r0_b / dd, dd_b = -r0_b * r0 / dd;
  const double qcl_b = cmax_pass(1.e-35, qi_clea * fclea, nn_b) * fclea;
  qi_b += cmax_pass(1.e-35, qi, dd_b);
  qi_b += qcl_b / fclea;
  const double qcd_b = -qcl_b * fcldy / fclea;
  const double x1_b = cmax_pass(0.0, x1, qcd_b);
  qi_b += x1_b;
  qc_b += x1_b - x1_b / fcldy;
}

// cell_setup: st_b, T_b, p_b, qv_b accumulate the inputs' adjoints
template <bool PRE>
__host__ __device__ __forceinline__ void cell_setup_rev(const ChainArgs &a, long off,
                                                        const CellCtx<Fp64> &X,
                                                        const CtxBar &Xb, double st_b[NSTATE],
                                                        double &T_b, double &p_b,
                                                        double &qv_b) {
  const double T = X.T, p = X.p, qv = a.qv[off];
  T_b += Xb.T;
  p_b += Xb.p;
  // relative humidity (Magnus)
  const double epsqs = MW_H2O / MW_DRY_AIR;
  const double tc = T - 273.15;
  const double ew = 610.94 * exp(17.625 * tc / (234.04 + tc));
  const double fp = 1.00071 * exp(4.5e-8 * p);
  const double ev = fp * ew;
  const double dd = p - (1 - epsqs) * ev;
  const double qvs = epsqs * ev / dd;
  const double r = qv / qvs;
  const double rhgcm = kmax(0.0, kmin(1.0, r));
  double rh_b = 0.0;
  if (X.cat == 2)
    rh_b = mmb_pass(0.0, 1.0, (rhgcm - X.fcldy) / X.fclea, Xb.rh_clear) / X.fclea;
  else
    rh_b = Xb.rh_clear;
  const double r_b = mmb_pass(0.0, 1.0, r, rh_b);
  qv_b += r_b / qvs;
  const double qvs_b = -r_b * r / qvs;
  const double ev_b = qvs_b * (epsqs / dd + qvs * (1 - epsqs) / dd);
  p_b -= qvs_b * qvs / dd;
  p_b += ev_b * ev * 4.5e-8;
  const double ew_b = ev_b * fp;
  T_b += ew_b * ew * 17.625 * 234.04 / sq(234.04 + tc);
  // the gases before gas chemistry / before cloud chemistry / now
  if (!(PRE && a.pre_h2so4)) ld3_rev(a, off, F_H2SO4, TO_VMR_H2SO4, Xb.h1, st_b);
  if (!(PRE && a.pre_soag)) ld3_rev(a, off, F_SOAG, TO_VMR_SOAG, Xb.s1, st_b);
  ld3_rev(a, off, F_H2SO4, TO_VMR_H2SO4, Xb.h3, st_b);
  ld3_rev(a, off, F_SOAG, TO_VMR_SOAG, Xb.s3, st_b);
  ld2_rev<PRE>(a, off, F_H2SO4, TO_VMR_H2SO4, Xb.h2, st_b);
  ld2_rev<PRE>(a, off, F_SOAG, TO_VMR_SOAG, Xb.s2, st_b);
  if (X.cat != 2) return;
  // partition factors
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    {
      const double qi = ld2<PRE, Fp64>(a, off, f_ni(m), TO_VMR_NUM);
      const double qc = ld2<PRE, Fp64>(a, off, f_nc(m), TO_VMR_NUM);
      double qi_b = 0, qc_b = 0;
      partition_factors_rev(qi, qc, X.fcldy, X.fclea, Xb.fa_n[m], Xb.fc_n[m], qi_b, qc_b);
      ld2_rev<PRE>(a, off, f_ni(m), TO_VMR_NUM, qi_b, st_b);
      ld2_rev<PRE>(a, off, f_nc(m), TO_VMR_NUM, qc_b, st_b);
    }
    double qi = 0.0, qc = 0.0;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      if (s >= NSPEC_MODE[m]) continue;
      qi = qi + ld2<PRE, Fp64>(a, off, f_qi(m, s), to_vmr_mass(m, s));
      qc = qc + ld2<PRE, Fp64>(a, off, f_qc(m, s), to_vmr_mass(m, s));
    }
    double qi_b = 0, qc_b = 0;
    partition_factors_rev(qi, qc, X.fcldy, X.fclea, Xb.fa_q[m], Xb.fc_q[m], qi_b, qc_b);
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      if (s >= NSPEC_MODE[m]) continue;
      ld2_rev<PRE>(a, off, f_qi(m, s), to_vmr_mass(m, s), qi_b, st_b);
      ld2_rev<PRE>(a, off, f_qc(m, s), to_vmr_mass(m, s), qc_b, st_b);
    }
  }
}

// clear_build: from the adjoints of the clear subarea S and its gases G,
// those of the context and the state
__host__ __device__ __forceinline__ void clear_build_rev(const ChainArgs &a, long off,
                                                         const CellCtx<Fp64> &X,
                                                         const Aer<double> &S_b,
                                                         const ClearGas<double> &G_b,
                                                         CtxBar &Xb, double st_b[NSTATE]) {
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const double nn = ld3<Fp64>(a, off, f_ni(m), TO_VMR_NUM);
    if (X.cat == 2) {
      Xb.fa_n[m] += S_b.n[m] * nn;
      ld3_rev(a, off, f_ni(m), TO_VMR_NUM, S_b.n[m] * X.fa_n[m], st_b);
    } else {
      ld3_rev(a, off, f_ni(m), TO_VMR_NUM, S_b.n[m], st_b);
    }
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      if (s >= NSPEC_MODE[m]) continue;
      const int sp = mode_spec(m, s);
      const double qq = ld3<Fp64>(a, off, f_qi(m, s), to_vmr_mass(m, s));
      const double yb = S_b.q[sp][m] * fcvt_aer(sp);
      if (X.cat == 2) {
        Xb.fa_q[m] += yb * qq;
        ld3_rev(a, off, f_qi(m, s), to_vmr_mass(m, s), yb * X.fa_q[m], st_b);
      } else {
        ld3_rev(a, off, f_qi(m, s), to_vmr_mass(m, s), yb, st_b);
      }
    }
  }
  double hcl_b = G_b.hcl, scl_b = G_b.scl;
  if (X.map_gas) {
    Xb.s1 += G_b.g1[0] * FCVT_SOAG;
    Xb.h1 += G_b.g1[1];
    scl_b += G_b.g[0] * FCVT_SOAG;
    hcl_b += G_b.g[1];
  }
  if (X.cat == 2) {
    if ((X.h3 - X.fclea * X.h2) / X.fcldy < 0) {
      Xb.h3 += hcl_b / X.fclea;
    } else {
      Xb.h3 += G_b.hcd / X.fcldy;
      Xb.h2 += hcl_b - G_b.hcd * X.fclea / X.fcldy;
    }
    if ((X.s3 - X.fclea * X.s2) / X.fcldy < 0) {
      Xb.s3 += scl_b / X.fclea;
    } else {
      Xb.s3 += G_b.scd / X.fcldy;
      Xb.s2 += scl_b - G_b.scd * X.fclea / X.fcldy;
    }
  } else {
    Xb.h3 += hcl_b;
    Xb.s3 += scl_b;
  }
}

// soa_in_state: from the adjoints of the loop inputs it sets (in_b.g, .a,
// .aopoa), those of the subarea's state S and SOAG
__host__ __device__ __forceinline__ void soa_in_state_rev(const Aer<double> &S,
                                                          const SoaIn<double> &in_b,
                                                          Aer<double> &S_b, double &g_soag_b) {
  g_soag_b += in_b.g;
#pragma unroll
  for (int n = 0; n < 4; ++n) S_b.q[A_SOA][n] += in_b.a[n];
  S_b.q[A_POM][NACC] += maxc_pass(S.q[A_POM][NACC], 0.0, 0.1 * in_b.aopoa[NACC]);
  S_b.q[A_POM][NCOA] += maxc_pass(S.q[A_POM][NCOA], 0.0, 0.1 * in_b.aopoa[NCOA]);
}

// clear_pre: from the adjoints of the SOA loop's inputs (in_b; its uptk
// entry also carries the H2SO4 uptake's other uses), those of S, the SOAG
// it started from, temp and pmid
__host__ __device__ __forceinline__ void clear_pre_rev(const ChainConsts &K, double temp,
                                                       double pmid, const double dgn_awet[4],
                                                       const Aer<double> &S,
                                                       const SoaIn<double> &in_b,
                                                       Aer<double> &S_b, double &g_soag_b,
                                                       double &temp_b, double &pmid_b) {
  const double aircon = pmid / (8314.46759100000 * temp);
  double aircon_b = 0.0;
  h2so4_uptake_rev(K, temp, pmid, aircon, dgn_awet, S.n, in_b.uptk, temp_b, pmid_b, aircon_b,
                   S_b.n);
  soa_in_state_rev(S, in_b, S_b, g_soag_b);
  soa_g0_rev(K, temp, pmid, in_b.g0, temp_b, pmid_b);
  pmid_b += aircon_b / (8314.46759100000 * temp);
  temp_b -= aircon_b * aircon / temp;
}

// cell_finish (no diagnostics): from the adjoints of the tendencies tb, those
// of the clear subarea S and gases G after clear_post, the context and the
// state (the state's direct use in the tendency included)
template <bool PRE>
__host__ __device__ __forceinline__ void cell_finish_rev(
    const ChainArgs &a, long off, const CellCtx<Fp64> &X, const Aer<double> &S,
    const ClearGas<double> &G, const double tb[NSTATE], Aer<double> &S_b,
    ClearGas<double> &G_b, CtxBar &Xb, double st_b[NSTATE]) {
  const ChainConsts &K = *a.K;
  const double dt = a.dt;
  const int ns = nsub(a.cfg);
  // ---- forward ----
  double gi_out[NSTATE];
  if (X.cat != 1) {
    const double hout = X.map_gas ? G.g[1] / 1.0 : G.hcl;
    const double sout = X.map_gas ? G.g[0] / FCVT_SOAG : G.scl;
    gi_out[F_H2SO4] = hout * X.fclea;
    gi_out[F_SOAG] = sout * X.fclea;
    if (X.cat == 2) {
      gi_out[F_H2SO4] += G.hcd * X.fcldy;
      gi_out[F_SOAG] += G.scd * X.fcldy;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      gi_out[f_ni(m)] = (S.n[m] / 1.0) * X.fclea;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        gi_out[f_qi(m, s)] = (S.q[sp][m] / fcvt_aer(sp)) * X.fclea;
      }
    }
  } else {
    gi_out[F_H2SO4] = X.h3 * 1.0;
    gi_out[F_SOAG] = X.s3 * 1.0;
  }
  Aer<double> C, Cc, C0, Cc0;
  double gi[7] = {0, 0, 0, 0, 0, 0, 0}, gc[7] = {0, 0, 0, 0, 0, 0, 0};
  if (X.cat != 0) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double n3 = ld3<Fp64>(a, off, f_ni(m), TO_VMR_NUM);
      const double nc3 = ld3<Fp64>(a, off, f_nc(m), TO_VMR_NUM);
      C.n[m] = (X.cat == 2 ? n3 * X.fc_n[m] : n3) * 1.0;
      Cc.n[m] = (X.cat == 2 ? nc3 / X.fcldy : nc3) * 1.0;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        const double q3 = ld3<Fp64>(a, off, f_qi(m, s), to_vmr_mass(m, s));
        const double qc3 = ld3<Fp64>(a, off, f_qc(m, s), to_vmr_mass(m, s));
        C.q[sp][m] = (X.cat == 2 ? q3 * X.fc_q[m] : q3) * fcvt_aer(sp);
        Cc.q[sp][m] = (X.cat == 2 ? qc3 / X.fcldy : qc3) * fcvt_aer(sp);
        if (m == NAIT) {
          const double q2 = ld2<PRE, Fp64>(a, off, f_qi(m, s), to_vmr_mass(m, s));
          const double qc2 = ld2<PRE, Fp64>(a, off, f_qc(m, s), to_vmr_mass(m, s));
          const double a2 = (X.cat == 2 ? q2 * X.fc_q[m] : q2) * fcvt_aer(sp);
          const double c2 = (X.cat == 2 ? qc2 / X.fcldy : qc2) * fcvt_aer(sp);
          gi[sp] = (C.q[sp][m] - a2) / ns;
          gc[sp] = (Cc.q[sp][m] - c2) / ns;
        }
      }
    }
    C0 = C;
    Cc0 = Cc;
    // one renaming per sub-step, each with 1/ns of the cloud-chemistry growth
    if (a.cfg.do_rename)
      for (int j = 0; j < ns; ++j) rename_ait_to_acc<Fp64>(K, true, C, gi, Cc, gc);
    const double w = X.fcldy;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double ni = (C.n[m] / 1.0) * w;
      gi_out[f_ni(m)] = (X.cat == 2) ? gi_out[f_ni(m)] + ni : ni;
      gi_out[f_nc(m)] = (Cc.n[m] / 1.0) * w;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        const double qi = (C.q[sp][m] / fcvt_aer(sp)) * w;
        gi_out[f_qi(m, s)] = (X.cat == 2) ? gi_out[f_qi(m, s)] + qi : qi;
        gi_out[f_qc(m, s)] = (Cc.q[sp][m] / fcvt_aer(sp)) * w;
      }
    }
  } else {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      gi_out[f_nc(m)] = ld3w<Fp64>(a, off, f_nc(m), TO_VMR_NUM);
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        gi_out[f_qc(m, s)] = ld3w<Fp64>(a, off, f_qc(m, s), to_vmr_mass(m, s));
      }
    }
  }
  double post[NSTATE], tomm[NSTATE];
  bool clip[NSTATE];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    tomm[f_ni(m)] = TO_MMR_NUM;
    tomm[f_nc(m)] = TO_MMR_NUM;
    clip[f_ni(m)] = true;
    clip[f_nc(m)] = false;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      if (s >= NSPEC_MODE[m]) continue;
      tomm[f_qi(m, s)] = to_mmr_mass(m, s);
      tomm[f_qc(m, s)] = to_mmr_mass(m, s);
      clip[f_qi(m, s)] = true;
      clip[f_qc(m, s)] = false;
    }
  }
  tomm[F_H2SO4] = TO_MMR_H2SO4;
  tomm[F_SOAG] = TO_MMR_SOAG;
  clip[F_H2SO4] = clip[F_SOAG] = true;
#pragma unroll
  for (int f = 0; f < NSTATE; ++f)
    post[f] = (clip[f] ? kmax(0.0, gi_out[f]) : gi_out[f]) * tomm[f];

  // ---- reverse ----
  double post_b[NSTATE];
#pragma unroll
  for (int f = 0; f < NSTATE; ++f) {
    post_b[f] = tb[f] / dt;
    st_b[f] -= tb[f] / dt;
  }
  if (a.do_calcsize) {
    double n_i[4], n_c[4], q_i[4][7], q_c[4][7];
    double n_i_b[4] = {0, 0, 0, 0}, n_c_b[4] = {0, 0, 0, 0}, q_i_b[4][7], q_c_b[4][7];
    CalcOut<double> o_b;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      n_i[m] = post[f_ni(m)];
      n_c[m] = post[f_nc(m)];
      o_b.dn_i[m] = tb[f_ni(m)];
      o_b.dn_c[m] = tb[f_nc(m)];
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        const bool ok = s < NSPEC_MODE[m];
        q_i[m][s] = ok ? post[f_qi(m, s)] : 0.0;
        q_c[m][s] = ok ? post[f_qc(m, s)] : 0.0;
        o_b.dq_i[m][s] = ok ? tb[f_qi(m, s)] : 0.0;
        o_b.dq_c[m][s] = ok ? tb[f_qc(m, s)] : 0.0;
        q_i_b[m][s] = q_c_b[m][s] = 0.0;
      }
    }
    calcsize_rev(K, a.do_xfer != 0, a.do_adjust != 0, dt, n_i, n_c, q_i, q_c, o_b, n_i_b, n_c_b,
                 q_i_b, q_c_b);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      post_b[f_ni(m)] += n_i_b[m];
      post_b[f_nc(m)] += n_c_b[m];
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        post_b[f_qi(m, s)] += q_i_b[m][s];
        post_b[f_qc(m, s)] += q_c_b[m][s];
      }
    }
  }
  double go_b[NSTATE];
#pragma unroll
  for (int f = 0; f < NSTATE; ++f)
    go_b[f] = clip[f] ? cmax_pass(0.0, gi_out[f], post_b[f] * tomm[f]) : post_b[f] * tomm[f];
  if (X.cat != 0) {
    const double w = X.fcldy;
    Aer<double> C_b, Cc_b;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      C_b.n[m] = go_b[f_ni(m)] * w;
      Cc_b.n[m] = go_b[f_nc(m)] * w;
#pragma unroll
      for (int aa = 0; aa < 7; ++aa) C_b.q[aa][m] = Cc_b.q[aa][m] = 0.0;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        C_b.q[sp][m] = go_b[f_qi(m, s)] * w / fcvt_aer(sp);
        Cc_b.q[sp][m] = go_b[f_qc(m, s)] * w / fcvt_aer(sp);
      }
    }
    double gi_b[7] = {0, 0, 0, 0, 0, 0, 0}, gc_b[7] = {0, 0, 0, 0, 0, 0, 0};
    if (a.cfg.do_rename) {
      // sub-step j's renaming starts from C0 renamed j times (recomputed)
      for (int j = ns - 1; j >= 0; --j) {
        Aer<double> Cj = C0, Ccj = Cc0;
        for (int i = 0; i < j; ++i) rename_ait_to_acc<Fp64>(K, true, Cj, gi, Ccj, gc);
        rename_rev(K, true, Cj, gi, Ccj, gc, C_b, gi_b, Cc_b, gc_b);
      }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double n3 = ld3<Fp64>(a, off, f_ni(m), TO_VMR_NUM);
      if (X.cat == 2) {
        Xb.fc_n[m] += C_b.n[m] * n3;
        ld3_rev(a, off, f_ni(m), TO_VMR_NUM, C_b.n[m] * X.fc_n[m], st_b);
        ld3_rev(a, off, f_nc(m), TO_VMR_NUM, Cc_b.n[m] / X.fcldy, st_b);
      } else {
        ld3_rev(a, off, f_ni(m), TO_VMR_NUM, C_b.n[m], st_b);
        ld3_rev(a, off, f_nc(m), TO_VMR_NUM, Cc_b.n[m], st_b);
      }
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        double cq_b = C_b.q[sp][m], ccq_b = Cc_b.q[sp][m];
        if (m == NAIT) {
          // gi = C.q - a2, gc = Cc.q - c2
          const double gib = gi_b[sp] / ns, gcb = gc_b[sp] / ns;
          cq_b += gib;
          ccq_b += gcb;
          const double q2 = ld2<PRE, Fp64>(a, off, f_qi(m, s), to_vmr_mass(m, s));
          const double a2_b = -gib * fcvt_aer(sp), c2_b = -gcb * fcvt_aer(sp);
          if (X.cat == 2) {
            Xb.fc_q[m] += a2_b * q2;
            ld2_rev<PRE>(a, off, f_qi(m, s), to_vmr_mass(m, s), a2_b * X.fc_q[m], st_b);
            ld2_rev<PRE>(a, off, f_qc(m, s), to_vmr_mass(m, s), c2_b / X.fcldy, st_b);
          } else {
            ld2_rev<PRE>(a, off, f_qi(m, s), to_vmr_mass(m, s), a2_b, st_b);
            ld2_rev<PRE>(a, off, f_qc(m, s), to_vmr_mass(m, s), c2_b, st_b);
          }
        }
        const double q3 = ld3<Fp64>(a, off, f_qi(m, s), to_vmr_mass(m, s));
        const double yb = cq_b * fcvt_aer(sp), ycb = ccq_b * fcvt_aer(sp);
        if (X.cat == 2) {
          Xb.fc_q[m] += yb * q3;
          ld3_rev(a, off, f_qi(m, s), to_vmr_mass(m, s), yb * X.fc_q[m], st_b);
          ld3_rev(a, off, f_qc(m, s), to_vmr_mass(m, s), ycb / X.fcldy, st_b);
        } else {
          ld3_rev(a, off, f_qi(m, s), to_vmr_mass(m, s), yb, st_b);
          ld3_rev(a, off, f_qc(m, s), to_vmr_mass(m, s), ycb, st_b);
        }
      }
    }
  } else {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      ld3_rev(a, off, f_nc(m), TO_VMR_NUM, go_b[f_nc(m)], st_b);
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        ld3_rev(a, off, f_qc(m, s), to_vmr_mass(m, s), go_b[f_qc(m, s)], st_b);
      }
    }
  }
  if (X.cat != 1) {
    if (X.map_gas) {
      G_b.g[1] += go_b[F_H2SO4] * X.fclea;
      G_b.g[0] += go_b[F_SOAG] * X.fclea / FCVT_SOAG;
    } else {
      G_b.hcl += go_b[F_H2SO4] * X.fclea;
      G_b.scl += go_b[F_SOAG] * X.fclea;
    }
    if (X.cat == 2) {
      G_b.hcd += go_b[F_H2SO4] * X.fcldy;
      G_b.scd += go_b[F_SOAG] * X.fcldy;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      S_b.n[m] += go_b[f_ni(m)] * X.fclea;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        S_b.q[sp][m] += go_b[f_qi(m, s)] * X.fclea / fcvt_aer(sp);
      }
    }
  } else {
    Xb.h3 += go_b[F_H2SO4];
    Xb.s3 += go_b[F_SOAG];
  }
}

// What the phase before the SOA loop's reverse hands on: the adjoints of the
// loop's results (soa_g_b, soa_a_b, soa_gavg_b) and of its uptake-rate input
// from the later processes (uptk_b), and the gradient contributed by every
// path that avoids the loop (g, T_b, p_b, qv_b).  The reverses of clear_pre,
// clear_build and cell_setup are linear in their adjoints, so the second
// phase adds the loop's contribution to this partial gradient.  Between the
// phases of the GPU's split sweep it is held in HBM.
struct PostBar {
  double g[NSTATE], T_b, p_b, qv_b;
  double uptk_b[4];
  double soa_g_b, soa_gavg_b, soa_a_b[4];
};

// Phase "post" of a cell with a clear subarea (X.cat != 1): clear_post run
// forward from S0 / G0 (clear_build's outputs), in (clear_pre's; zero without
// condensation) and the SOA loop's results; the reverse of cell_finish and
// clear_post from the cotangent tb; then the reverse of clear_build and
// cell_setup for every adjoint that does not pass through the SOA loop.
template <bool PRE>
__host__ __device__ __forceinline__ void adjoint_post_phase(
    const ChainArgs &a, long off, const CellCtx<Fp64> &X, const Aer<double> &S0,
    const ClearGas<double> &G0, const SoaIn<double> &in, double sg, double sgavg,
    const double sa[4], const double tb[NSTATE], PostBar &pb) {
  const ChainConsts &K = *a.K;
  Aer<double> S = S0;
  ClearGas<double> G = G0;
  clear_post<Fp64>(K, a.cfg, a.dt, X.T, X.p, X.z, X.pblh, X.rh_clear, X.dga, X.dgw, X.wd, G.g1,
                   G.g, S, in.uptk, sg, sgavg, sa);
  Aer<double> S_b;
  ClearGas<double> G_b;
  CtxBar Xb;
  zero(S_b);
  zero(G_b);
  zero(Xb);
#pragma unroll
  for (int f = 0; f < NSTATE; ++f) pb.g[f] = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    pb.uptk_b[n] = 0.0;
    pb.soa_a_b[n] = 0.0;
  }
  pb.soa_g_b = pb.soa_gavg_b = 0.0;
  pb.T_b = pb.p_b = pb.qv_b = 0.0;
  cell_finish_rev<PRE>(a, off, X, S, G, tb, S_b, G_b, Xb, pb.g);
  clear_post_rev(K, a.cfg, a.dt, X.T, X.p, X.z, X.pblh, X.rh_clear, X.dga, X.dgw, X.wd, G0.g1,
                 G0.g, S0, in.uptk, sg, sgavg, sa, S_b, G_b.g, G_b.g1, pb.uptk_b, pb.soa_g_b,
                 pb.soa_gavg_b, pb.soa_a_b, Xb.T, Xb.p, Xb.rh_clear);
  clear_build_rev(a, off, X, S_b, G_b, Xb, pb.g);
  cell_setup_rev<PRE>(a, off, X, Xb, pb.g, pb.T_b, pb.p_b, pb.qv_b);
}

// Phase "pre" of a cell with a clear subarea and condensation: the gradient
// through the SOA loop's inputs, from their adjoints in_b (its uptk entry
// holding the later processes' part too), by the reverse of clear_pre,
// clear_build and cell_setup.  st_b, T_b, p_b, qv_b receive it (zeroed here);
// the cell's gradient is this plus the post phase's partial one.
template <bool PRE>
__host__ __device__ __forceinline__ void adjoint_pre_phase(
    const ChainArgs &a, long off, const CellCtx<Fp64> &X, const Aer<double> &S0,
    const SoaIn<double> &in_b, double st_b[NSTATE], double &T_b, double &p_b, double &qv_b) {
  Aer<double> S_b;
  ClearGas<double> G_b;
  CtxBar Xb;
  zero(S_b);
  zero(G_b);
  zero(Xb);
#pragma unroll
  for (int f = 0; f < NSTATE; ++f) st_b[f] = 0.0;
  T_b = p_b = qv_b = 0.0;
  clear_pre_rev(*a.K, X.T, X.p, X.dgw, S0, in_b, S_b, G_b.g[0], Xb.T, Xb.p);
  clear_build_rev(a, off, X, S_b, G_b, Xb, st_b);
  cell_setup_rev<PRE>(a, off, X, Xb, st_b, T_b, p_b, qv_b);
}

// The state a SOA sub-step starts from, as the reverse sweep keeps it.
__host__ __device__ __forceinline__ SoaCkpt soa_ckpt(const SoaState<double> &st) {
  SoaCkpt c;
  c.g = st.g;
#pragma unroll
  for (int n = 0; n < 4; ++n) c.a[n] = st.a[n];
  c.tcur = st.tcur;
  return c;
}

// A SOA loop's sub-step states for the reverse sweep in fixed storage
// (2 C records): record() runs the loop keeping up to C states; a longer
// loop (nt > C, at most C^2 sub-steps: C = 32 covers the 1000-sub-step cap)
// is run a second time keeping every s-th state, s = ceil(nt / C), and
// get(i), called for i = nt - 1 down to 0, recomputes each s-long segment
// from its first state once, when it first needs it.
template <int C>
struct SoaTape2 {
  SoaState<double> init;
  SoaCkpt coarse[C], fine[C];
  double dtfull;
  int nt, s, seg;
  __host__ __device__ int record(SoaState<double> &st, double dt) {
    init = st;
    dtfull = dt;
    nt = 0;
    s = 1;
    seg = -1;
    while (true) {
      const SoaCkpt c = soa_ckpt(st);
      if (!soa_tick<true>(st, dtfull)) break;
      if (nt < C) coarse[nt] = c;
      ++nt;
    }
    if (nt > C) {
      s = (nt + C - 1) / C;
      SoaState<double> t = init;
      for (int i = 0; i < nt; ++i) {
        if (i % s == 0) coarse[i / s] = soa_ckpt(t);
        soa_tick<true>(t, dtfull);
      }
    }
    return nt;
  }
  __host__ __device__ SoaCkpt get(int i) {
    if (s == 1) return coarse[i];
    const int sg = i / s;
    if (sg != seg) {
      SoaState<double> t = init;
      t.g = coarse[sg].g;
#pragma unroll
      for (int n = 0; n < 4; ++n) t.a[n] = coarse[sg].a[n];
      t.tcur = coarse[sg].tcur;
      t.niter = sg * s;
      const int lo = sg * s, hi = lo + s < nt ? lo + s : nt;
      for (int k = lo; k < hi; ++k) {
        fine[k - lo] = soa_ckpt(t);
        soa_tick<true>(t, dtfull);
      }
      seg = sg;
    }
    return fine[i - sg * s];
  }
};

// One sub-step's subarea state kept in registers (ntsubstep = 1 only).
struct OneSubTape {
  Aer<double> S;
  double g[2];
  __host__ __device__ void save(int, const Aer<double> &S_, const double g_[2]) {
    S = S_;
    g[0] = g_[0];
    g[1] = g_[1];
  }
  __host__ __device__ void load(int, Aer<double> &S_, double g_[2]) const {
    S_ = S;
    g_[0] = g[0];
    g_[1] = g[1];
  }
};

// The whole step's adjoint for one cell, with ntsubstep sub-steps: the
// forward step recomputed, keeping the subarea's aerosol and gases at the
// start of every sub-step (sub.save / sub.load, 34 doubles each); then the
// reverse sweep, sub-step by sub-step from the last, each sub-step's SOA
// loop recomputed from its start (soa.record) and read back in reverse
// (soa.get(i), i = nt - 1 .. 0).  tb: cotangent of the 52 tendencies;
// st_b, T_b, p_b, qv_b receive the gradient of <tb, tendencies> w.r.t. the
// state, T, p and qv (st_b zeroed by the caller).  Returns the SOA sub-step
// count.  The GPU runs the one-sub-step case as separate kernels, the SOA
// loop's on a work queue.
template <bool PRE, class SoaTape, class SubTape>
__host__ __device__ __forceinline__ int adjoint_cell(const ChainArgs &a, long off, int icol,
                                                     const double tb[NSTATE], SoaTape &soa,
                                                     SubTape &sub, double st_b[NSTATE],
                                                     double &T_b, double &p_b, double &qv_b) {
  const ChainConsts &K = *a.K;
  const Cfg &cfg = a.cfg;
  CellCtx<Fp64> X;
  cell_setup<PRE, Fp64>(a, off, icol, X);
  CtxBar Xb;
  zero(Xb);
  int niter = 0;
  if (X.cat != 1) {
    Aer<double> S0;
    ClearGas<double> G0;
    clear_build<Fp64>(a, off, X, S0, G0);
    SoaIn<double> in0 = {};
    if (cfg.do_cond) clear_pre<Fp64>(K, X.T, X.p, X.dgw, S0, G0.g[0], in0);
    const int ns = nsub(cfg);
    const double dtsub = dt_sub(cfg, a.dt);
    double g[2] = {G0.g[0], G0.g[1]};
    const double netprod = clear_post_begin(cfg, a.dt, G0.g1, g);
    // ---- forward over the sub-steps ----
    Aer<double> S = S0;
    for (int j = 0; j < ns; ++j) {
      sub.save(j, S, g);
      SoaIn<double> in = in0;
      if (j > 0 && cfg.do_cond) soa_in_state<Fp64>(S, g[0], in);
      double sg = 0.0, sgavg = 0.0, sa[4] = {0, 0, 0, 0};
      if (cfg.do_cond) {
        SoaState<double> st;
        soa_init(st, in);
        while (soa_tick<true>(st, dtsub)) {
        }
        niter += st.niter;
        sg = st.g;
        sgavg = soa_gavg(st);
#pragma unroll
        for (int n = 0; n < 4; ++n) sa[n] = st.a[n];
      }
      clear_post_sub<Fp64>(K, cfg, dtsub, netprod, X.T, X.p, X.z, X.pblh, X.rh_clear, X.dga,
                           X.dgw, X.wd, g, S, in0.uptk, sg, sgavg, sa);
    }
    ClearGas<double> G = G0;
    G.g[0] = g[0];
    G.g[1] = g[1];
    // ---- reverse ----
    Aer<double> S_b;
    ClearGas<double> G_b;
    zero(S_b);
    zero(G_b);
    cell_finish_rev<PRE>(a, off, X, S, G, tb, S_b, G_b, Xb, st_b);
    double netprod_b = 0.0;
    SoaIn<double> in0_b = {};
    for (int j = ns - 1; j >= 0; --j) {
      Aer<double> Sj;
      double gj[2];
      sub.load(j, Sj, gj);
      SoaIn<double> in = in0;
      if (j > 0 && cfg.do_cond) soa_in_state<Fp64>(Sj, gj[0], in);
      double sg = 0.0, sgavg = 0.0, sa[4] = {0, 0, 0, 0};
      SoaState<double> st;
      int nt = 0;
      if (cfg.do_cond) {
        soa_init(st, in);
        nt = soa.record(st, dtsub);
        sg = st.g;
        sgavg = soa_gavg(st);
#pragma unroll
        for (int n = 0; n < 4; ++n) sa[n] = st.a[n];
      }
      double sg_b = 0.0, sgavg_b = 0.0, sa_b[4] = {0, 0, 0, 0}, uptk_b[4] = {0, 0, 0, 0};
      clear_post_sub_rev(K, cfg, dtsub, netprod, X.T, X.p, X.z, X.pblh, X.rh_clear, X.dga, X.dgw,
                         X.wd, gj, Sj, in0.uptk, sg, sgavg, sa, S_b, G_b.g, uptk_b, netprod_b,
                         sg_b, sgavg_b, sa_b, Xb.T, Xb.p, Xb.rh_clear);
#pragma unroll
      for (int n = 0; n < 4; ++n) in0_b.uptk[n] += uptk_b[n];
      if (cfg.do_cond) {
        SoaIn<double> in_b = {};
        auto get = [&](int i) { return soa.get(i); };
        soa_loop_rev(in, dtsub, st, get, nt, sg_b, sa_b, sgavg_b, in_b);
        // the rates and g0 are the first sub-step's for all
#pragma unroll
        for (int n = 0; n < 4; ++n) in0_b.uptk[n] += in_b.uptk[n];
        in0_b.g0 += in_b.g0;
        if (j > 0) {
          soa_in_state_rev(Sj, in_b, S_b, G_b.g[0]);
        } else {
          in0_b.g += in_b.g;
#pragma unroll
          for (int n = 0; n < 4; ++n) {
            in0_b.a[n] += in_b.a[n];
            in0_b.aopoa[n] += in_b.aopoa[n];
          }
        }
      }
    }
    clear_post_begin_rev(cfg, a.dt, netprod, netprod_b, G_b.g, G_b.g1);
    if (cfg.do_cond)
      clear_pre_rev(K, X.T, X.p, X.dgw, S0, in0_b, S_b, G_b.g[0], Xb.T, Xb.p);
    clear_build_rev(a, off, X, S_b, G_b, Xb, st_b);
  } else {
    // overcast: no clear subarea, no SOA loop
    Aer<double> S, S_b;
    ClearGas<double> G, G_b;
    zero(S_b);
    zero(G_b);
    cell_finish_rev<PRE>(a, off, X, S, G, tb, S_b, G_b, Xb, st_b);
  }
  cell_setup_rev<PRE>(a, off, X, Xb, st_b, T_b, p_b, qv_b);
  return niter;
}

}  // namespace rev
}  // namespace dev
}  // namespace mam4x
