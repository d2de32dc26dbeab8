// Per-cell MAM4 microphysics chain for gfx950 (one lane = one cell).
//
// The reference (eagles-project/mam4xx, mam4_amicphys.hpp) carries each
// subarea through ~8 KB of [31][3] and [31][4][3] scratch arrays per cell.
// Here a subarea is the 30 values that can actually change -- 2 gases, 4
// numbers and the 24 (species, mode) masses that are prognostic or that
// condensation/coagulation can put into the primary-carbon mode before aging
// empties it (chain_consts.hpp: live()) -- so the whole chain stays in
// registers.  Per-process increment bookkeeping the reference keeps only for
// its optional history diagnostics is kept only by the instantiation that
// forms them (P::diag, ProcSink below); the few increments that feed back
// into the physics (Aitken growth for renaming, primary-carbon
// condensation/coagulation increments for aging) are always kept.
//
// Every routine is a template over a precision policy P (scalar.hpp): R is
// the bulk type, W the type of the FP64-kept parts.  With P = Fp64 both are
// double and the code performs the reference's operations one by one.
//
// min/max are Kokkos' ternaries, not fmin/fmax (NaN semantics differ).
#pragma once

#include "scalar.hpp"

namespace mam4x {
namespace dev {

template <class A, class B>
__host__ __device__ __forceinline__ auto kmax(A a, B b) {
  using T = decltype(a + b);
  return (T(a) < T(b)) ? T(b) : T(a);
}
template <class A, class B>
__host__ __device__ __forceinline__ auto kmin(A a, B b) {
  using T = decltype(a + b);
  return (T(b) < T(a)) ? T(b) : T(a);
}
template <class T>
__host__ __device__ __forceinline__ T sq(T x) { return x * x; }
template <class T>
__host__ __device__ __forceinline__ T cube(T x) { return x * x * x; }
template <class A, class B, class C>
__host__ __device__ __forceinline__ auto mmb(A lo, B hi, C x) {  // utils.hpp:23
  return kmax(lo, kmin(hi, x));
}

// Aerosol of one subarea in chain units: numbers [#/kmol], masses [mol/mol]
// indexed [AeroId][mode]; only live() entries are ever read or written.
template <class R>
struct Aer {
  R n[4];
  R q[7][4];
};

struct Cfg {
  int do_cond, do_rename, do_newnuc, do_coag;
  int gaexch_h2so4_uptake_optaa;
  unsigned nmono;
  int ntsubstep;  // condensation sub-steps per step; 0 or 1: one (the reference's)
};

// Sub-steps of the subarea chain per step (mam4_amicphys.hpp:1288 fixes 1)
// and their length, an int (:1319).
__host__ __device__ __forceinline__ int nsub(const Cfg &c) {
  return c.ntsubstep > 1 ? c.ntsubstep : 1;
}
__host__ __device__ __forceinline__ double dt_sub(const Cfg &c, double deltat) {
  return (double)(int)(deltat / nsub(c));
}

// Per-process tendency diagnostics of one cell: the gas_aero_exchange_*
// arrays of modal_aero_amicphys_intr (mam4_amicphys.hpp:1940-1986,
// 2069-2113, 2365-2374).  Block k (0 condensation with its share of aging,
// 1 renaming, 2 nucleation, 3 coagulation, 4 cloud-borne renaming) holds
// TMR tendencies [mol/mol/s, #/kmol/s] of the 31 gas_pcnst tracers.  A
// subarea's change over a process, divided by fcvt * deltat and weighted by
// the subarea's area fraction, is added to the zeroed planes in subarea
// order (clear, then cloudy) -- the reference's order of operations.
struct ProcSink {
  double *blk[5];  // NULL: block not formed
  long stride;     // between the tracer planes of a block
  long off;        // the cell's plane offset
  double deltat;
  double afrac;    // area fraction of the subarea being added
  __host__ __device__ __forceinline__ void zero() const {
#pragma unroll
    for (int k = 0; k < 5; ++k)
      if (blk[k])
        for (int i = 0; i < 31; ++i) blk[k][i * stride + off] = 0.0;
  }
  __host__ __device__ __forceinline__ void add(int k, int slot, double dq, double fcvt) const {
    if (!blk[k]) return;
    double *p = blk[k] + slot * stride + off;
    *p = *p + (dq / (fcvt * deltat)) * afrac;
  }
  // the change of a subarea's aerosol between two snapshots
  __host__ __device__ __forceinline__ void add_aer(int k, const Aer<double> &now,
                                                   const Aer<double> &before) const {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      add(k, NUMPTR[m], now.n[m] - before.n[m], 1.0);
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        add(k, mass_slot(m, s), now.q[sp][m] - before.q[sp][m], fcvt_aer(sp));
      }
    }
  }
};

// ------------------------------------------------------------------ uptake
// gas_aer_uptkrates_1box1gas for H2SO4 (gasaerexch.hpp:267-350, accom 0.65,
// beta from the Knudsen number) times number*aircon (gasaerexch.hpp:427-435).
// The SOAG rates are 0.81 x these (:458-465), so SOAG's own transport
// properties, which the reference computes and discards, are not evaluated.
template <class P>
__device__ __forceinline__ void h2so4_uptake(const ChainConsts &K, typename P::R temp,
                                             typename P::R pmid, typename P::R aircon,
                                             const typename P::R dgn_awet[4],
                                             const typename P::R qnum[4],
                                             typename P::R uptk[4]) {
  using R = typename P::R;
  const R p_in_atm = pmid / 1.013e5;
  const R dgas = (1.0e-3 * pow(temp, R(1.75)) * K.h2so4_sqrt_mw) /
                 (p_in_atm * K.h2so4_vd_term);
  const R gasdiffus = dgas * 1.0e-4;
  const R speed = 145.5 * sqrt(temp / 9.807840000000000E+01);
  const R freepath = 3.0 * gasdiffus / speed;
  const double a283 = 6.5000E-01 * 0.283, a75 = 6.5000E-01 * 0.75;
  const double xghq0 = 7.071067690849304E-01, wghq = 8.862269520759583E-01;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const double lnsg = K.alnsg[n];
    const R lndpgn = log(dgn_awet[n]);
    const R kn0 = 2.0 * freepath / dgn_awet[n];
    const R tmpa = 1.0 / (1.0 + kn0) -
                   (2.0 * kn0 + 1.0 + a283) / (kn0 * (kn0 + 1.0 + a283) + a75);
    R beta = 1.0 - kn0 * tmpa;
    beta = kmax(R(1.0), kmin(R(2.0), beta));
    const R constant = K.tworootpi * exp(beta * lndpgn + 0.5 * sq(beta * lnsg));
    R sumghq = 0.0;
#pragma unroll
    for (int iq = 0; iq < 2; ++iq) {
      const double x = iq == 0 ? xghq0 : -xghq0;
      const R lndp = lndpgn + beta * lnsg * lnsg + K.root2 * lnsg * x;
      const R D_p = exp(lndp);
      const R kn = 2.0 * freepath / D_p;
      const R hh = (a75 * (1.0 + kn)) / (kn * (kn + 1.0 + a283) + a75);
      sumghq += wghq * D_p * hh / pow(D_p, beta);
    }
    uptk[n] = (constant * gasdiffus * sumghq) * (qnum[n] * aircon);
  }
}

// ------------------------------------------------------------------ SOA
// mam_soaexch_1subarea (gasaerexch_soaexch.hpp:79-341), one SOA species, as
// a resumable state machine: soa_tick advances one adaptive sub-step, so the
// persistent work-queue kernel can hand a lane a new cell as soon as its
// current one has converged.  Always in the wide type: the step size
// 0.05 / sum(k |phi|) divides by a near-cancelling phi, so any rounding change
// alters the step sequence.
template <class W>
struct SoaIn {        // inputs of the SOA loop of one clear subarea
  W g;                // SOAG [mol/mol, fcvt applied]
  W a[4];             // SOA in the four modes
  W uptk[4];          // H2SO4 uptake rates [1/s]; SOAG's are 0.81x
  W aopoa[4];         // oxygenated POM (0.1 x POM, 0 in the pcarbon mode)
  W g0;               // equilibrium SOAG mixing ratio
};

template <class W>
struct SoaState {
  W g, a[4], uk[4], aopoa[4], g0;
  W tcur, avg, dtsum;
  int niter;
  bool skip[4];
};

template <class R>
__device__ __forceinline__ R soa_g0(const ChainConsts &K, R temp, R pmid) {
  const R p0_soa = 1.0e-10 * exp(-(K.soa_c1) * ((1.0 / temp) - K.soa_inv298));
  return 1.01325e5 * p0_soa / pmid;
}

template <class W>
__device__ __forceinline__ void soa_init(SoaState<W> &s, const SoaIn<W> &in) {
  s.g = in.g;
  s.g0 = in.g0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const W uk = in.uptk[n] * 0.81;  // gasaerexch.hpp:458-465
    s.skip[n] = !(uk > 1.0e-15);
    s.uk[n] = s.skip[n] ? W(0.0) : uk;
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
template <class W>
__device__ __forceinline__ double soa_predict(const SoaState<W> &s, double dtfull) {
  const double g_soa = num::val(kmax(s.g, 0.0));
  double tmpb = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    if (s.skip[n]) continue;
    const double a_soa = num::val(kmax(s.a[n], 0.0));
    const double sat = num::val(s.g0) / kmax(num::val(s.aopoa[n]) + a_soa, 1.0e-20);
    const double gstar = sat * a_soa;
    tmpb += num::val(s.uk[n]) * ::fabs((g_soa - gstar) / kmax(g_soa, kmax(gstar, 1.0e-20)));
  }
  return dtfull * tmpb / 0.05;
}

// A skipped mode's term of a sum.  In the plain types a skipped mode's term is
// already exactly +0 (see soa_tick), so it is added unconditionally; the Dual
// type selects, so that a derivative of a skipped mode's discarded quotient
// (0 * inf) cannot reach the sums.
__device__ __forceinline__ double soa_term(bool, double x) { return x; }
__device__ __forceinline__ num::Dual soa_term(bool skip, num::Dual x) {
  return skip ? num::Dual(0.0) : x;
}

// One pass of the while loop (gasaerexch_soaexch.hpp:170-335).  Returns false
// once the loop has ended (time reached or 1000 sub-steps taken).
//
// FAST divides with num::div_fast (reciprocal, Newton steps, Markstein
// correction: IEEE's quotient wherever numerators are 0 or at least 2^-968,
// quotients normal and divisors in range); otherwise each quotient is the
// IEEE a / b.  soa_fast_exact below states when every FAST quotient of a
// cell's loop is the IEEE one; soa_run runs a cell's loop on that condition.
//
// Written without branches on the mode.  A skipped mode (uptake <= 1e-15) has
// uk = 0 and a_soa = 0, so beta = 0, den = 1 and each of its terms
// (a_soa, uk |phi|, a_soa / den, beta / den) is exactly +0: adding it leaves
// a sum bit-for-bit unchanged, and its new SOA, (0 + 0 g) / 1, is +0 -- the
// values the reference's skip produces.  The four modes' divisions are then
// four independent chains the scheduler interleaves, instead of a branch per
// mode that serialises them.
template <bool FAST, class W>
__device__ __forceinline__ W soa_div(W a, W b) {
  if constexpr (FAST)
    return num::div_fast(a, b);
  else
    return a / b;
}

template <bool FAST = false, class W>
__device__ __forceinline__ bool soa_tick(SoaState<W> &s, double dtfull) {
  const double alpha_astem = 0.05, a_min1 = 1.0e-20, g_min1 = 1.0e-20;
  if (!(s.tcur < dtfull - 1.0e-3)) return false;
  s.niter++;
  if (s.niter > 1000) return false;
  const W gprv = s.g;
  W g_soa = kmax(gprv, 0.0);
  W a_soa[4], sat[4], term[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    a_soa[n] = s.skip[n] ? W(0.0) : kmax(s.a[n], 0.0);
    sat[n] = soa_div<FAST, W>(s.g0, kmax(s.aopoa[n] + a_soa[n], a_min1));
    const W gstar = sat[n] * a_soa[n];
    const W phi = soa_div<FAST, W>(g_soa - gstar, kmax(g_soa, kmax(gstar, g_min1)));
    term[n] = soa_term(s.skip[n], s.uk[n] * fabs(phi));
  }
  // the sums in the reference's mode order
  W tot = g_soa, tmpb = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    tot = tot + soa_term(s.skip[n], a_soa[n]);
    tmpb = tmpb + term[n];
  }
  const W tmpa = kmax(0.0, tmpb);
  const W dtmax = dtfull - s.tcur;
  const bool last = dtmax * tmpa <= alpha_astem;
  const W dtcur = last ? dtmax : soa_div<FAST, W>(W(alpha_astem), tmpa);
  s.tcur = last ? W(dtfull) : s.tcur + dtcur;
  // step 1: explicit estimate of condensing modes' a_soa and sat
  W beta[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    beta[n] = dtcur * s.uk[n];
    const W gstar = sat[n] * a_soa[n];
    const W del = g_soa - gstar;
    const W at = a_soa[n] + beta[n] * del;
    const W sat1 = soa_div<FAST, W>(s.g0, kmax(s.aopoa[n] + at, a_min1));
    sat[n] = del > 0.0 ? sat1 : sat[n];
  }
  // step 2: implicit in g, semi-implicit in a
  W sa = 0.0, sb = 0.0, den[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    den[n] = 1.0 + beta[n] * sat[n];
    sa = sa + soa_term(s.skip[n], soa_div<FAST, W>(a_soa[n], den[n]));
    sb = sb + soa_term(s.skip[n], soa_div<FAST, W>(beta[n], den[n]));
  }
  g_soa = soa_div<FAST, W>(tot - sa, 1.0 + sb);
  g_soa = kmax(0.0, g_soa);
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    // skipped modes: a_soa is never set, so their SOA becomes 0 (quirk of
    // gasaerexch_soaexch.hpp:327-331)
    s.a[n] = soa_term(s.skip[n], soa_div<FAST, W>(a_soa[n] + beta[n] * g_soa, den[n]));
  }
  s.g = g_soa;
  s.avg += dtcur * (gprv + 0.5 * (g_soa - gprv));
  s.dtsum += dtcur;
  return true;
}

// qgas_avg = max(0, sum(q dt) / sum(dt)) (gasaerexch_soaexch.hpp:337-340)
template <class W>
__device__ __forceinline__ W soa_gavg(const SoaState<W> &s) {
  return kmax(0.0, s.avg / s.dtsum);
}

// ---- when the FAST sub-step is IEEE's, bit for bit
// soa_fast_exact: with the cell's inputs in the ranges soa_fast_ok checks,
// and SOAG and each mode's SOA at 0 or >= 2^-400 in every state the loop
// passes through (soa_state_exp >= kSoaExpMin, the final state included),
// every kept quotient of a FAST sub-step has a numerator 0 or >= 2^-810, a
// normal result and a divisor in [2^-67, 2^158], so num::div_fast returns
// IEEE's quotient.  The bounds, for a state whose masses lie in
// {0} u [2^-400, 2^11] (the loop conserves g + sum a, which starts below
// 5 x 2^8; skipped modes' quotients are discarded):
//   sat = g0 / max(aopoa + a, 1e-20)    g0 in {0} u [2^-300, 2^10]: [2^-312, 2^77]
//   g* = sat a                          {0} u [2^-712, 2^88]
//   g - g* (phi's numerator)            multiples of 2^-764; divisor [2^-67, 2^88]
//   dtcur                               [2^-67, 2^20]: dtmax > 1e-3, or 0.05 / tmpa
//                                       with tmpa <= sum uk |phi| <= 2^62
//   beta = dtcur uk                     {0} u [2^-117, 2^80]: uk > 1e-15 unless
//                                       skipped, uk <= 2^60
//   sat1 = g0 / max(aopoa + at, 1e-20)  |at| <= 2^169: >= 2^-470
//   den = 1 + beta sat                  [1, 2^158]
//   a / den, beta / den                 >= 2^-558, >= 2^-275
//   tot - sa                            multiples of 2^-610 (every term >= 2^-558)
//   g' = (tot - sa) / (1 + sb)          {0} u [2^-693, 2^11]
//   a + beta g'                         {0} u [2^-810, ...)
//   a' = (a + beta g') / den            {0} u [2^-968, 2^11], normal
// A cell outside these (a mass below 2^-400 at some sub-step, an input out of
// range or not finite) runs the IEEE loop from the start instead.
constexpr int kSoaExpMin = -399;     // frexp exponent of 2^-400
constexpr int kSoaExpFail = -100000;  // a cell whose inputs are out of range

__device__ __forceinline__ bool soa_fast_ok(const SoaState<double> &s, double dtfull) {
  bool ok = dtfull <= 0x1p20 && (s.g0 == 0.0 || (s.g0 >= 0x1p-300 && s.g0 <= 0x1p10)) &&
            ::fabs(s.g) <= 0x1p8;
#pragma unroll
  for (int n = 0; n < 4; ++n)
    ok = ok && s.uk[n] <= 0x1p60 && ::fabs(s.aopoa[n]) <= 0x1p10 && ::fabs(s.a[n]) <= 0x1p8;
  return ok;
}
__device__ __forceinline__ bool soa_fast_ok(const SoaState<num::Dual> &, double) { return true; }

__device__ __forceinline__ int frexp_exp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_frexp_exp(x);
#else
  int e;
  (void)::frexp(x, &e);
  return e;
#endif
}

// The smallest frexp exponent among SOAG and the modes' SOA (0 for a zero).
__device__ __forceinline__ int soa_state_exp(const SoaState<double> &s) {
  int e = frexp_exp(s.g);
#pragma unroll
  for (int n = 0; n < 4; ++n) e = kmin(e, frexp_exp(s.a[n]));
  return e;
}
__device__ __forceinline__ int soa_state_exp(const SoaState<num::Dual> &) { return 0; }

// One cell's SOA loop from its inputs, equal bit for bit to the IEEE loop:
// FAST sub-steps while soa_fast_exact holds, else the IEEE loop from the
// start.
template <class W>
__device__ __forceinline__ void soa_run(SoaState<W> &st, const SoaIn<W> &in, double dtfull) {
  soa_init(st, in);
  int emin = soa_fast_ok(st, dtfull) ? 0 : kSoaExpFail;
  while (true) {
    emin = kmin(emin, soa_state_exp(st));
    if (!soa_tick<true>(st, dtfull)) break;
  }
  if (emin < kSoaExpMin) {
    soa_init(st, in);
    while (soa_tick<false>(st, dtfull)) {
    }
  }
}

// ------------------------------------------------------------------ H2SO4
// Analytic non-volatile update (gasaerexch.hpp:479-537), in the wide type.
template <class W>
__device__ __forceinline__ void h2so4_update(double dt, W netprod, const W uptk[4], W &qgas,
                                             W &qgas_avg, W qso4[4]) {
  W tmpa = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) tmpa += uptk[n];
  const W kxt = tmpa * dt;
  const W pxt = netprod * dt;
  const W q1 = qgas;
  if (kxt >= 1.0e-20) {
    W q3, q4;
    if (kxt > 0.001) {
      const W pok = pxt / kxt;
      const W e = exp(-kxt);
      q3 = (q1 - pok) * e + pok;
      q4 = (q1 - pok) * (1.0 - e) / kxt + pok;
    } else {
      const W k2 = kxt * kxt;
      q3 = q1 * (1.0 - kxt + k2 * 0.5) + pxt * (1.0 - kxt * 0.5 + k2 / 6.0);
      q4 = q1 * (1.0 - kxt * 0.5 + k2 / 6.0) + pxt * (0.5 - kxt / 6.0 + k2 / 24.0);
    }
    qgas = q3;
    qgas_avg = q4;
    const W del = (q1 + pxt) - q3;
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
// and cloud-borne aerosol.  Diameters and erfc tail fractions in the wide
// type.
template <class P>
__device__ __forceinline__ void rename_ait_to_acc(const ChainConsts &K, bool cloudy,
                                                  Aer<typename P::R> &ai,
                                                  const typename P::R gi[7],
                                                  Aer<typename P::R> &ac,
                                                  const typename P::R gc[7]) {
  using R = typename P::R;
  using W = typename P::W;
  const double smallest = 1.0e-25;
  R vi = 0.0, dvi = 0.0, vc = 0.0, dvc = 0.0;
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
  W b4_dryvol = num::to<W>(vi - dvi);
  if (cloudy) b4_dryvol = b4_dryvol + num::to<W>(vc - dvc);
  const W b4_bnd = kmax(b4_dryvol, smallest);
  W b4_qnum = num::to<W>(ai.n[NAIT]);
  if (cloudy) b4_qnum = b4_qnum + num::to<W>(ac.n[NAIT]);
  b4_qnum = kmax(0.0, b4_qnum);
  const W b4_qnum_bnd = mmb(b4_bnd * K.rn_v2n_hi_rlx, b4_bnd * K.rn_v2n_lo_rlx, b4_qnum);
  W dv = num::to<W>(dvi);
  if (cloudy) dv = dv + num::to<W>(dvc);
  const W aft_dryvol = b4_dryvol + dv;
  if (aft_dryvol <= smallest) return;
  const double pio6 = PI / 6;
  W b4_diam = cbrt((b4_bnd / b4_qnum_bnd) / pio6) * K.dgn_from_vol[NAIT];
  if (b4_diam > K.rn_threshold) {
    b4_dryvol = b4_dryvol * cube(K.rn_threshold / b4_diam);
    b4_diam = K.rn_threshold;
  }
  if ((aft_dryvol - b4_dryvol) <= 1.0e-6 * b4_bnd) return;
  const W aft_diam = cbrt((aft_dryvol / b4_qnum_bnd) / pio6) * K.dgn_from_vol[NAIT];
  if (aft_diam <= MODE_DNOM[NAIT]) return;
  const W lb4 = log(b4_diam), laft = log(aft_diam);
  const W b4_tail_num = 0.5 * erfc((K.rn_ln_cutoff - lb4) * K.rn_tail_fac);
  const W b4_tail_vol = 0.5 * erfc((K.rn_ln_cutoff - (lb4 + K.rn_ln_tail_fac)) * K.rn_tail_fac);
  const W aft_tail_num = 0.5 * erfc((K.rn_ln_cutoff - laft) * K.rn_tail_fac);
  const W aft_tail_vol =
      0.5 * erfc((K.rn_ln_cutoff - (laft + K.rn_ln_tail_fac)) * K.rn_tail_fac);
  const W vfrac = aft_tail_vol * aft_dryvol - b4_tail_vol * b4_dryvol;
  if (vfrac <= 0.0) return;
  W xvw = kmin(vfrac, aft_dryvol) / aft_dryvol;
  xvw = kmin(xvw, K.xferfrac_max);
  W xnw = aft_tail_num - b4_tail_num;
  xnw = kmax(0.0, kmin(xnw, xvw));
  const R xv = num::to<R>(xvw), xn = num::to<R>(xnw);
  {
    const R nt = ai.n[NAIT] * xn;
    ai.n[NAIT] -= nt;
    ai.n[NACC] += nt;
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      if (!live(a, NAIT)) continue;
      const R vt = ai.q[a][NAIT] * xv;
      ai.q[a][NAIT] -= vt;
      ai.q[a][NACC] += vt;
    }
  }
  if (cloudy) {
    const R nt = ac.n[NAIT] * xn;
    ac.n[NAIT] -= nt;
    ac.n[NACC] += nt;
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      if (!live(a, NAIT)) continue;
      const R vt = ac.q[a][NAIT] * xv;
      ac.q[a][NAIT] -= vt;
      ac.q[a][NACC] += vt;
    }
  }
}

// ------------------------------------------------------------------ newnuc
// Vehkamaki et al. (2002) eqs 11-14 (vehkamaki2002.hpp:53-179): the ln J
// polynomials, in the wide type.
template <class W>
__device__ __forceinline__ void vehk2002(W c, W t, W rh, W &rateloge, W &cnum_tot, W &xcrit) {
  const W lnN = log(c), lnrh = log(rh);
  const W t2 = sq(t), t3 = cube(t);
  const W r2 = sq(lnrh), r3 = cube(lnrh), N2 = sq(lnN), N3 = cube(lnN);
  const W x = 0.740997 - 0.00266379 * t - 0.00349998 * lnN + 0.0000504022 * t * lnN +
              0.00201048 * lnrh - 0.000183289 * t * lnrh + 0.00157407 * r2 -
              0.0000179059 * t * r2 + 0.000184403 * r3 - 1.50345e-6 * t * r3;
  const W a = 0.14309 + 2.21956 * t - 0.0273911 * t2 + 0.0000722811 * t3 + 5.91822 / x;
  const W b = 0.117489 + 0.462532 * t - 0.0118059 * t2 + 0.0000404196 * t3 + 15.7963 / x;
  const W cc = -0.215554 - 0.0810269 * t + 0.00143581 * t2 - 4.7758e-6 * t3 - 2.91297 / x;
  const W d = -3.58856 + 0.049508 * t - 0.00021382 * t2 + 3.10801e-7 * t3 - 0.0293333 / x;
  const W e = 1.14598 - 0.600796 * t + 0.00864245 * t2 - 0.0000228947 * t3 - 8.44985 / x;
  const W f = 2.15855 + 0.0808121 * t - 0.000407382 * t2 - 4.01957e-7 * t3 + 0.721326 / x;
  const W g = 1.6241 - 0.0160106 * t + 0.0000377124 * t2 + 3.21794e-8 * t3 - 0.0113255 / x;
  const W h = 9.71682 - 0.115048 * t + 0.000157098 * t2 + 4.00914e-7 * t3 + 0.71186 / x;
  const W i = -1.05611 + 0.00903378 * t - 0.0000198417 * t2 + 2.46048e-8 * t3 - 0.0579087 / x;
  const W j = -0.148712 + 0.00283508 * t - 9.24619e-6 * t2 + 5.00427e-9 * t3 - 0.0127081 / x;
  rateloge = a + b * lnrh + cc * r2 + d * r3 + e * lnN + f * lnrh * lnN + g * r2 * lnN +
             h * N2 + i * lnrh * N2 + j * N3;
  const W A = -0.00295413 - 0.0976834 * t + 0.00102485 * t2 - 2.18646e-6 * t3 - 0.101717 / x;
  const W B = -0.00205064 - 0.00758504 * t + 0.000192654 * t2 - 6.7043e-7 * t3 - 0.255774 / x;
  const W C = +0.00322308 + 0.000852637 * t - 0.0000154757 * t2 + 5.66661e-8 * t3 + 0.0338444 / x;
  const W D = +0.0474323 - 0.000625104 * t + 2.65066e-6 * t2 - 3.67471e-9 * t3 - 0.000267251 / x;
  const W E = -0.0125211 + 0.00580655 * t - 0.000101674 * t2 + 2.88195e-7 * t3 + 0.0942243 / x;
  const W F = -0.038546 - 0.000672316 * t + 2.60288e-6 * t2 + 1.19416e-8 * t3 - 0.00851515 / x;
  const W G = -0.0183749 + 0.000172072 * t - 3.71766e-7 * t2 - 5.14875e-10 * t3 + 0.00026866 / x;
  const W H = -0.0619974 + 0.000906958 * t - 9.11728e-7 * t2 - 5.36796e-9 * t3 - 0.00774234 / x;
  const W I = +0.0121827 - 0.00010665 * t + 2.5346e-7 * t2 - 3.63519e-10 * t3 + 0.000610065 / x;
  const W J = +0.000320184 - 0.0000174762 * t + 6.06504e-8 * t2 - 1.4177e-11 * t3 + 0.000135751 / x;
  cnum_tot = exp(A + B * lnrh + C * r2 + D * r3 + E * lnN + F * lnrh * lnN + G * r2 * lnN +
                 H * N2 + I * lnrh * N2 + J * N3);
  xcrit = x;
}

// mam_newnuc_1subarea (mam4_amicphys.hpp:916-1095) with
// mer07_veh02_nuc_mosaic_1box (nucleation.hpp:346-809) inlined for the path
// the chain takes: method 11, no NH3, one size bin [dplom, dphim].  The rate
// (ln J) is formed in the wide type.
template <class P>
__device__ __forceinline__ void newnuc(const ChainConsts &K, double deltat, typename P::R temp,
                                       typename P::R pmid, typename P::R zmid,
                                       typename P::R pblh, typename P::R relhum,
                                       typename P::R uptkrate_h2so4, typename P::W &h2so4,
                                       typename P::W h2so4_avg, typename P::R &num_ait,
                                       typename P::R &so4_ait) {
  using R = typename P::R;
  using W = typename P::W;
  if (h2so4_avg <= 4.0e-16) return;
  const R relhumnn = kmax(0.01, kmin(0.99, relhum));
  const double pi = PI;
  const double rgas = 8.31446759100000, avogad = 6.022140000000000E+023;
  const double mw_so4a = 96, mw_so4a_host = 115;
  const W qh2so4_cur = h2so4;
  // --- mer07_veh02_nuc_mosaic_1box ---
  const W cair = num::to<W>(pmid) / (num::to<W>(temp) * rgas);
  const W so4vol_in = h2so4_avg * cair * avogad * 1.0e-6;
  W rateloge = K.ln_1em38;
  W cnum_h2so4 = 0.0, cnum_tot = 0.0;
  R radius_cluster = 0.0;
  if (so4vol_in >= 1.0e4) {
    const W temp_bb = kmax(230.15, kmin(305.15, num::to<W>(temp)));
    const W rh_bb = kmax(1.0e-4, kmin(1.0, num::to<W>(relhumnn)));
    const W so4vol_bb = kmax(1.0e4, kmin(1.0e11, so4vol_in));
    W x;
    vehk2002(so4vol_bb, temp_bb, rh_bb, rateloge, cnum_tot, x);
    cnum_h2so4 = cnum_tot * x;
    radius_cluster = num::to<R>(exp(-1.6524245 + 0.42316402 * x + 0.3346648 * log(cnum_tot)));
  }
  rateloge = rateloge + ::log(kmax(1.0e-38, 1.0));
  if (zmid <= kmax(pblh, 100.0)) {  // Wang & Penner (2008) first-order PBL rate
    const W tmp_ratenucl = (1e-6 * so4vol_in) * 1.0;
    const W tmp_rateloge = log(kmax(1.0e-38, tmp_ratenucl));
    if (!(tmp_rateloge <= rateloge)) {
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
  const W ratenuclt_bb = exp(rateloge) * 1.0e6;
  const R tmpa0 = kmax(0.10, kmin(0.95, relhumnn));
  const R wetvol_dryvol = 1.0 - 0.56 / log(tmpa0);
  R voldry_clus = (num::to<R>(kmax(cnum_h2so4, 1.0)) * mw_so4a + 0.0 * 18) /
                  (1.0e3 * 1770.0 * avogad);
  voldry_clus = voldry_clus * (mw_so4a_host / mw_so4a);
  const R dpdry_clus = pow(voldry_clus * 6.0 / pi, R(1.0 / 3.0));
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
  const double voldry_part = (pi / 6.0) * ::pow(dpdry_part, 3);
  // composition: no NH3, so pure sulfuric acid when not growing and
  // ammonium bisulfate/sulfuric acid with zero NH3 (n2 = 0/h2so4) when growing
  W n1 = 0.0, n2 = 0.0, n3 = 1.0;
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
  const W m1 = n1 * 132.0, m2 = n2 * 114.0, m3 = n3 * 96.0;
  const W dens_part = (m1 + m2 + m3) / ((m1 / 1770.0) + (m2 / 1770.0) + (m3 / 1770.0));
  const W mass_part = voldry_part * dens_part;
  const W mole_nh4_so4 = 2.0 * n1 + n2;
  const W kgaero_per_moleso4a = 1.0e-3 * (m1 + m2 + m3) * (mw_so4a_host / mw_so4a);
  // Kerminen-Kulmala growth factor exp(nu/dfin - nu/dnuc): its argument can
  // reach tens, so it is formed in the wide type
  W factor_kk = 1.0;
  if (igrow) {
    const W tw = num::to<W>(temp);
    const W wvd = num::to<W>(wetvol_dryvol);
    const W wet_volfrac_so4a = 1.0 / (wvd * (1.0 + mole_nh4_so4 * 17.0 / 98.0));
    const W tmp_spd = 14.7 * sqrt(tw);
    const W gr_kk = 3.0e-9 * tmp_spd * 96.0 * so4vol_in / (dens_part * wet_volfrac_so4a);
    const W dfin_kk = 1.0e9 * dpdry_part * pow(wvd, W(1.0 / 3.0));
    const W dnuc_kk = kmax(2.0 * num::to<W>(radius_cluster), 1.0);
    const W gamma_kk = 0.23 * pow(dnuc_kk, W(0.2)) * pow(dfin_kk / 3.0, W(0.075)) *
                       pow(dens_part * 1.0e-3, W(-0.33)) * pow(tw / 293.0, W(-0.75));
    const W up = kmax(num::to<W>(uptkrate_h2so4) * 3600.0, 0.0);
    const W dif = (6.7037e-6 * pow(tw, W(0.75)) / cair) * 3600.0;
    const W cs_prime_kk = up / (4.0 * pi * dif * 0.65);
    const W nu_kk = gamma_kk * cs_prime_kk / gr_kk;
    factor_kk = exp((nu_kk / dfin_kk) - (nu_kk / dnuc_kk));
  }
  const W ratenuclt_kk = ratenuclt_bb * factor_kk;
  const W qmolso4a_del_max =
      kmax(0.0, (ratenuclt_kk * deltat * mass_part)) / (kgaero_per_moleso4a * cair);
  W freducea = 1.0;
  if (qmolso4a_del_max > qh2so4_cur) freducea = qh2so4_cur / qmolso4a_del_max;
  W freduceb = 1.0;
  if (mole_nh4_so4 >= 1.0e-10) {
    if (qmolso4a_del_max * mole_nh4_so4 > 0.0) freduceb = 0.0 / (qmolso4a_del_max * mole_nh4_so4);
  }
  const W freduce = kmin(freducea, freduceb);
  if ((freduce * ratenuclt_kk) <= 1.0e-12) return;
  const W qh2so4_del = kmin(0.9999 * qh2so4_cur, freduce * qmolso4a_del_max);
  const W qnh3_del = kmin(0.9999 * 0.0, qh2so4_del * mole_nh4_so4);
  const W qso4a_del = qh2so4_del;
  const W qnuma_del = 1.0e-3 * (qso4a_del * mw_so4a + qnh3_del * 18) / mass_part;
  // --- back in mam_newnuc_1subarea ---
  W dndt_ait = (qnuma_del * 1.0e3) / deltat;
  W dmdt_ait = kmax(0.0, ((qso4a_del * mw_so4a_host) / deltat));
  if (dndt_ait < 1.0e2) {
    dndt_ait = 0.0;
    dmdt_ait = 0.0;
  } else {
    const W mass1p = dmdt_ait / dndt_ait;
    if (mass1p < K.mass1p_aitlo)
      dndt_ait = dmdt_ait / K.mass1p_aitlo;
    else if (mass1p > K.mass1p_aithi)
      dmdt_ait = dndt_ait * K.mass1p_aithi;
  }
  num_ait = num::to<R>(num::to<W>(num_ait) + dndt_ait * deltat);
  const W dso4dt_ait = dmdt_ait * 1.0 / mw_so4a_host;
  if (dso4dt_ait > 0.0) {
    W dq = dso4dt_ait * deltat;
    so4_ait = num::to<R>(num::to<W>(so4_ait) + dq);
    dq = kmin(dq, h2so4);
    h2so4 = h2so4 - dq;
  }
}


// ------------------------------------------------------------------ coag
// getcoags_wrapper_f + getcoags (coagulation.hpp:650-817) for one pair, with
// the sigma-only factors taken from P.  Returns betaij0, betaij3 and, when
// asked for, the intramodal rates of the source (ii0) and destination (jj0).
template <class R, bool WANT_II, bool WANT_JJ>
__device__ __forceinline__ void coag_pair(const ChainConsts &K, const ChainConsts::CoagPair &P,
                                          R temp, R lamda, R knc, R dgatk, R dgacc, R rho_at,
                                          R rho_ac, R &ij0, R &ij3, R &ii0, R &jj0) {
  const double boltzmann = 1.3806500000000000e-023;
  const double a = 1.246;
  const R kfmatac = sqrt(6.0 * boltzmann * temp / (rho_at + rho_ac));
  const R dgat3 = dgatk * dgatk * dgatk;
  const R sqdgat = sqrt(dgatk), sqdgac = sqrt(dgacc);
  const R sqdgat7 = dgat3 * sqdgat;
  const R r1 = sqdgac / sqdgat;
  const R r2 = r1 * r1;
  const R rx4 = r2 * r2;
  const R ri1 = 1.0 / r1;
  const R ri2 = 1.0 / (r1 * r1);
  const R ri3 = 1.0 / (r1 * r1 * r1);
  const R kngat = 2.0 * lamda / dgatk;
  const R kngac = 2.0 * lamda / dgacc;
  const R rat = dgacc / dgatk;
  const int n1 = (int)num::val(kmax(1.0, kmin(10.0, 1 + round(K.dlgsqt2 * log(rat))))) - 1;
  const double bm0ij = P.bm0ij[n1], bm3i = P.bm3i[n1];
  // intermodal 0th moment (coagulation.hpp:297-328)
  {
    const R nc = knc * (2.0 + a * (kngat * (P.esat04 + r2 * P.esat16 * P.esac04) +
                                   kngac * (P.esac04 + ri2 * P.esac16 * P.esat04)) +
                        (r2 + ri2) * P.esat04 * P.esac04);
    const R fm = kfmatac * sqdgat * bm0ij *
                 (P.esat01 + r1 * P.esac01 + 2.0 * r2 * P.esat01 * P.esac04 +
                  rx4 * P.esat09 * P.esac16 + ri3 * P.esat16 * P.esac09 +
                  2.0 * ri1 * P.esat04 + P.esac01);
    ij0 = kmax(0.0, nc * fm / (nc + fm));
  }
  // intermodal 3rd moment (coagulation.hpp:548-579)
  {
    const R nc = knc * dgat3 *
                 (2.0 * P.esat36 + a * kngat * (P.esat16 + r2 * P.esat04 * P.esac04) +
                  a * kngac * (P.esat36 * P.esac04 + ri2 * P.esat64 * P.esac16) +
                  r2 * P.esat16 * P.esac04 + ri2 * P.esat64 * P.esac04);
    const R fm = kfmatac * sqdgat7 * bm3i *
                 (P.esat49 + r1 * P.esat36 * P.esac01 + 2.0 * r2 * P.esat25 * P.esac04 +
                  rx4 * P.esat09 * P.esac16 + ri3 * P.esat100 * P.esac09 +
                  2.0 * ri1 * P.esat64 * P.esac01);
    const R qv12 = nc * fm / (nc + fm);
    ij3 = kmax(0.0, qv12 / (cube(dgatk) * P.dumatk_fac));
  }
  // intramodal 0th moment (coagulation.hpp:581-609)
  if (WANT_II) {
    const R kfmat = sqrt(3.0 * boltzmann * temp / rho_at);
    const R nc = knc * (1.0 + P.esat08 + a * kngat * (P.esat20 + P.esat04));
    const R fm = kfmat * sqdgat * P.bm0_at * (P.esat01 + P.esat25 + 2.0 * P.esat05);
    ii0 = kmax(0.0, fm * nc / (fm + nc));
  }
  if (WANT_JJ) {
    const R kfmac = sqrt(3.0 * boltzmann * temp / rho_ac);
    const R nc = knc * (1.0 + P.esac08 + a * kngac * (P.esac20 + P.esac04));
    const R fm = kfmac * sqdgac * P.bm0_ac * (P.esac01 + P.esac25 + 2.0 * P.esac05);
    jj0 = kmax(0.0, fm * nc / (fm + nc));
  }
}

// coagulation.hpp:937-953
template <class R>
__device__ __forceinline__ R qnum_intra_inter(R bijdtqnumj, R biidt, R qbgn) {
  if (bijdtqnumj < 1.0e-5)
    return qbgn / (1.0 + (bijdtqnumj + biidt * qbgn) * (1.0 + 0.5 * bijdtqnumj));
  const R e = exp(-bijdtqnumj);
  return qbgn * e / (1.0 + (biidt * qbgn / bijdtqnumj) * (1.0 - e));
}

// mam_coag_1subarea (coagulation.hpp:1037-1119).  coag_in receives the SO4
// and SOA mass moved into the primary-carbon mode (pair 3) for aging.
template <class P>
__device__ __forceinline__ void coagulate(const ChainConsts &K, double deltat,
                                          typename P::R temp, typename P::R pmid,
                                          typename P::R aircon,
                                          const typename P::R dgn_awet[4],
                                          const typename P::R wetdens[4],
                                          Aer<typename P::R> &S, typename P::R &coag_in_so4,
                                          typename P::R &coag_in_soa) {
  using R = typename P::R;
  R bgn[7][4];
#pragma unroll
  for (int a = 0; a < 7; ++a)
#pragma unroll
    for (int m = 0; m < 4; ++m)
      if (live(a, m)) {
        S.q[a][m] = kmax(0.0, S.q[a][m]);
        bgn[a][m] = S.q[a][m];
      }
  R nb[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    S.n[m] = kmax(0.0, S.n[m]);
    nb[m] = S.n[m];
  }
  // getcoags_wrapper_f prologue (coagulation.hpp:756-779), pair independent.
  // The rate coefficients are formed in the wide type: in SI units the
  // third-moment products (~1e-66) are far below FP32's range.
  using W = typename P::W;
  const double t0 = 273.15 + 15.0;
  const W tw = num::to<W>(temp);
  const W sqrt_temp = sqrt(tw);
  const W lamda = 6.6328e-8 * 101325.0 * tw / (t0 * num::to<W>(pmid));
  const W amu = 1.458e-6 * tw * sqrt_temp / (tw + 110.4);
  const W knc = (2.0 / 3.0) * 1.3806500000000000e-023 * tw / amu;
  W dgw[4], rho[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    dgw[m] = num::to<W>(dgn_awet[m]);
    rho[m] = num::to<W>(wetdens[m]);
  }
  W wij0[3], wij3[3], wii0a = 0, wjj0a = 0, wii0p = 0, dummy = 0;
  coag_pair<W, true, true>(K, K.pair[0], tw, lamda, knc, dgw[NAIT], dgw[NACC], rho[NAIT],
                           rho[NACC], wij0[0], wij3[0], wii0a, wjj0a);
  coag_pair<W, true, false>(K, K.pair[1], tw, lamda, knc, dgw[NPCA], dgw[NACC], rho[NPCA],
                            rho[NACC], wij0[1], wij3[1], wii0p, dummy);
  coag_pair<W, false, false>(K, K.pair[2], tw, lamda, knc, dgw[NAIT], dgw[NPCA], rho[NAIT],
                             rho[NPCA], wij0[2], wij3[2], dummy, dummy);
  const W airw = num::to<W>(aircon);
  R ij0[3], ij3[3];
#pragma unroll
  for (int ip = 0; ip < 3; ++ip) {
    ij0[ip] = num::to<R>(wij0[ip] * airw);
    ij3[ip] = num::to<R>(wij3[ip] * airw);
  }
  const R ii0a = num::to<R>(wii0a * airw);
  const R jj0a = num::to<R>(wjj0a * airw);
  const R ii0p = num::to<R>(wii0p * airw);
  // mam_coag_num_update (coagulation.hpp:981-1027)
  R tavg[4];
  S.n[NACC] = nb[NACC] / (1.0 + jj0a * deltat * nb[NACC]);
  tavg[NACC] = (nb[NACC] + S.n[NACC]) * 0.5;
  {
    const R bijdt = kmax(0.0, deltat * ij0[1] * tavg[NACC]);
    const R biidt = kmax(0.0, deltat * ii0p);
    S.n[NPCA] = qnum_intra_inter(bijdt, biidt, nb[NPCA]);
    tavg[NPCA] = (nb[NPCA] + S.n[NPCA]) * 0.5;
  }
  {
    R bij = ij0[0] * tavg[NACC];
    bij = bij + ij0[2] * tavg[NPCA];
    const R bijdt = kmax(0.0, deltat * bij);
    const R biidt = kmax(0.0, deltat * ii0a);
    S.n[NAIT] = qnum_intra_inter(bijdt, biidt, nb[NAIT]);
  }
  // mam_coag_aer_update (coagulation.hpp:849-928)
  coag_in_so4 = 0.0;
  coag_in_soa = 0.0;
  const double epsx2 = 2.220446049250313e-16 * 2.0;
  {
    const R b1 = kmax(0.0, ij3[0] * tavg[NACC]);
    const R b2 = kmax(0.0, ij3[2] * tavg[NPCA]);
    const R dc = b1 + b2;
    const R df = deltat * dc;
    if (df > epsx2) {
      const R prtn2 = b2 / dc;
      const R prtn1 = 1.0 - prtn2;
      const R xf = 1.0 - exp(-df);
#pragma unroll
      for (int a = 0; a < 7; ++a) {
        if (!live(a, NAIT)) continue;
        const R dq = xf * bgn[a][NAIT];
        S.q[a][NAIT] -= dq;
        S.q[a][NACC] += dq * prtn1;
        S.q[a][NPCA] += dq * prtn2;
        if (a == A_SO4) coag_in_so4 = 0.0 + dq * prtn2;
        if (a == A_SOA) coag_in_soa = 0.0 + dq * prtn2;
      }
    }
  }
  {
    const R dc = kmax(0.0, ij3[1] * tavg[NACC]);
    const R df = deltat * dc;
    if (df > epsx2) {
      const R xf = 1.0 - exp(-df);
#pragma unroll
      for (int a = 0; a < 7; ++a) {
        if (!live(a, NPCA)) continue;
        const R dq = xf * bgn[a][NPCA];
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
// xf_out / frac_cond_out (optional): the transferred fraction and the share
// of the transfer the diagnostics attribute to condensation.
template <class R>
__device__ __forceinline__ void age_pcarbon(const ChainConsts &K, unsigned nmono, R dgn_a_pca,
                                            Aer<R> &S, R dcond_so4, R dcond_soa, R coag_so4,
                                            R coag_soa, R *xf_out = nullptr,
                                            R *frac_cond_out = nullptr) {
  const R vol_shell = S.q[A_SO4][NPCA] * K.so4_vol + S.q[A_SOA][NPCA] * K.fac_m2v_eqvhyg;
  R del_cond = dcond_so4 * K.so4_vol + dcond_soa * K.fac_m2v_eqvhyg;
  const R del_coag = coag_so4 * K.so4_vol + coag_soa * K.fac_m2v_eqvhyg;
  del_cond = kmax(del_cond, 1e-35);
  const R frac_cond = del_cond / (del_cond + kmax(del_coag, 0.0));
  if (frac_cond_out) *frac_cond_out = frac_cond;  // only splits the diagnostics
  R vol_core = 0.0;
#pragma unroll
  for (int a = 0; a < 7; ++a)
    if (valid(a, NPCA)) vol_core += S.q[a][NPCA] * mass_2_vol(a);
  const R t1 = vol_shell * dgn_a_pca * K.fac_volsfc;
  const R t2 = kmax(6.0 * (nmono * 4.76e-10) * vol_core, 0.0);
  R xf;
  if (t1 >= t2)
    xf = K.xferfrac_max;
  else
    xf = kmin(t1 / t2, K.xferfrac_max);
  if (xf_out) *xf_out = xf;
  // POM, BC, MOM: aged fraction (aging.hpp:166-194)
#pragma unroll
  for (int a = 0; a < 7; ++a) {
    if (!(a == A_POM || a == A_BC || a == A_MOM)) continue;
    const R qt = S.q[a][NPCA] * xf;
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
  const R nt = S.n[NPCA] * xf;
  S.n[NPCA] -= nt;
  S.n[NACC] += nt;
}

// ------------------------------------------------------------------ subarea
// mam_amicphys_1subarea (mam4_amicphys.hpp:1126-1694) for a clear subarea,
// split around the SOA loop.
//
// The SOA loop's inputs taken from the subarea's state: SOAG, SOA per mode
// and the oxygenated POM (the rest of SoaIn stays).
template <class P>
__device__ __forceinline__ void soa_in_state(const Aer<typename P::R> &S, typename P::W g_soag,
                                             SoaIn<typename P::W> &o) {
  using W = typename P::W;
  o.g = g_soag;
#pragma unroll
  for (int n = 0; n < 4; ++n) o.a[n] = num::to<W>(S.q[A_SOA][n]);
  // opoa_frac = 0.1, 0 for the pcarbon mode (gasaerexch_soaexch.hpp:123-138)
  o.aopoa[NACC] = num::to<W>(0.1 * kmax(S.q[A_POM][NACC], 0.0));
  o.aopoa[NAIT] = 0.1 * kmax(0.0, 0.0);
  o.aopoa[NCOA] = num::to<W>(0.1 * kmax(S.q[A_POM][NCOA], 0.0));
  o.aopoa[NPCA] = num::to<W>(0.0 * kmax(S.q[A_POM][NPCA], 0.0));
}

// clear_pre: the part of mam_gasaerexch_1subarea before the SOA loop
// (gasaerexch.hpp:387-477): H2SO4 uptake rates and the loop's inputs.  With
// more than one sub-step the rates are those of the first, held for the
// others (the Fortran original's scope; the C++ reference leaves them
// undefined after the first, mam4_amicphys.hpp:1377, gasaerexch.hpp:428),
// and soa_in_state refreshes the rest of the inputs before each later one.
template <class P>
__device__ __forceinline__ void clear_pre(const ChainConsts &K, typename P::R temp,
                                          typename P::R pmid, const typename P::R dgn_awet[4],
                                          const Aer<typename P::R> &S, typename P::W g_soag,
                                          SoaIn<typename P::W> &o) {
  using R = typename P::R;
  using W = typename P::W;
  const R aircon = pmid / (8314.46759100000 * temp);
  R uptk[4];
  h2so4_uptake<P>(K, temp, pmid, aircon, dgn_awet, S.n, uptk);
#pragma unroll
  for (int n = 0; n < 4; ++n) o.uptk[n] = num::to<W>(uptk[n]);
  soa_in_state<P>(S, g_soag, o);
  o.g0 = num::to<W>(soa_g0(K, temp, pmid));
}

// Before the sub-steps (mam4_amicphys.hpp:1279-1286): with simultaneous
// H2SO4 uptake, the net production by other processes over the step, and
// H2SO4 restarts from its value before gas chemistry.
template <class W>
__device__ __forceinline__ W clear_post_begin(const Cfg &cfg, double deltat, const W g1[2],
                                              W g[2]) {
  W netprod = 0.0;
  if (cfg.do_cond && cfg.gaexch_h2so4_uptake_optaa == 2) {
    netprod = (g[1] - g1[1]) / deltat;
    g[1] = (netprod >= 0) ? g1[1] : W(0);  // QUIRK (:1283)
  }
  return netprod;
}

// One sub-step of everything after the SOA loop -- the H2SO4 update,
// renaming, nucleation, coagulation and aging -- over dtsub.  uptk: H2SO4
// uptake rates from clear_pre; soa_g/soa_gavg/soa_a: the sub-step's SOA loop
// results (ignored without condensation).  g: gases {SOAG, H2SO4}, current.
// D: the diagnostics sink (P::diag only), with the clear area fraction set.
template <class P>
__device__ __forceinline__ void clear_post_sub(
    const ChainConsts &K, const Cfg &cfg, double dtsub, typename P::W netprod,
    typename P::R temp, typename P::R pmid, typename P::R zmid, typename P::R pblh,
    typename P::R relhum, const typename P::R dgn_a[4], const typename P::R dgn_awet[4],
    const typename P::R wetdens[4], typename P::W g[2], Aer<typename P::R> &S,
    const typename P::W uptk[4], typename P::W soa_g, typename P::W soa_gavg,
    const typename P::W soa_a[4], const ProcSink *D = nullptr) {
  using R = typename P::R;
  using W = typename P::W;
  const R aircon = pmid / (8314.46759100000 * temp);
  [[maybe_unused]] const W gsv[2] = {g[0], g[1]};  // the gases condensation starts from
  W uptkrate_h2so4 = 0;
  W gavg[2];
  R dc_so4[4] = {0, 0, 0, 0}, dc_soa[4] = {0, 0, 0, 0};
  if (cfg.do_cond) {
#pragma unroll
    for (int n = 0; n < 4; ++n) uptkrate_h2so4 += uptk[n];
    g[0] = soa_g;
    gavg[0] = soa_gavg;
    W so4[4], so4_0[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) so4[n] = so4_0[n] = num::to<W>(S.q[A_SO4][n]);
    h2so4_update(dtsub, netprod, uptk, g[1], gavg[1], so4);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      dc_soa[n] = num::to<R>(soa_a[n] - num::to<W>(S.q[A_SOA][n]));
      dc_so4[n] = num::to<R>(so4[n] - so4_0[n]);
      S.q[A_SOA][n] = num::to<R>(soa_a[n]);
      S.q[A_SO4][n] = num::to<R>(so4[n]);
    }
    if constexpr (P::diag) {  // gas_pcnst 5 SOAG, 2 H2SO4 (mam4_amicphys.hpp:66-69)
      D->add(0, 5, g[0] - (gsv[0] + 0.0 * dtsub), FCVT_SOAG);
      D->add(0, 2, g[1] - (gsv[1] + netprod * dtsub), 1.0);
    }
  } else {
    gavg[0] = g[0];
    gavg[1] = g[1];
  }
  if (cfg.do_rename) {
    R gi[7] = {0, 0, 0, 0, 0, 0, 0}, gc[7] = {0, 0, 0, 0, 0, 0, 0};
    gi[A_SO4] = dc_so4[NAIT];
    gi[A_SOA] = dc_soa[NAIT];
    Aer<R> none;
    if constexpr (P::diag) {
      const Aer<R> S0 = S;
      rename_ait_to_acc<P>(K, false, S, gi, none, gc);
      D->add_aer(1, S, S0);
    } else {
      rename_ait_to_acc<P>(K, false, S, gi, none, gc);
    }
  }
  if (cfg.do_newnuc) {
    [[maybe_unused]] const W h0 = g[1];
    [[maybe_unused]] const R n0 = S.n[NAIT], q0 = S.q[A_SO4][NAIT];
    newnuc<P>(K, dtsub, temp, pmid, zmid, pblh, relhum, num::to<R>(uptkrate_h2so4), g[1],
              gavg[1], S.n[NAIT], S.q[A_SO4][NAIT]);
    if constexpr (P::diag) {
      D->add(2, 2, g[1] - h0, 1.0);
      D->add(2, NUMPTR[NAIT], S.n[NAIT] - n0, 1.0);
      D->add(2, mass_slot(NAIT, spec_in_mode(A_SO4, NAIT)), S.q[A_SO4][NAIT] - q0,
             fcvt_aer(A_SO4));
    }
  }
  R cin_so4 = 0.0, cin_soa = 0.0;
  if (cfg.do_coag) {
    if constexpr (P::diag) {
      const Aer<R> S0 = S;
      coagulate<P>(K, dtsub, temp, pmid, aircon, dgn_awet, wetdens, S, cin_so4, cin_soa);
      D->add_aer(3, S, S0);
    } else {
      coagulate<P>(K, dtsub, temp, pmid, aircon, dgn_awet, wetdens, S, cin_so4, cin_soa);
    }
  }
  if constexpr (P::diag) {
    const Aer<R> S0 = S;
    R xf, fc;
    age_pcarbon(K, cfg.nmono, dgn_a[NPCA], S, dc_so4[NPCA], dc_soa[NPCA], cin_so4, cin_soa,
                &xf, &fc);
    if (cfg.do_cond) {
      // the condensation increments, with aging's transfer re-attributed:
      // the condensation share of the aged POM/BC/MOM and number, and all
      // of the condensed primary-carbon SO4/SOA (aging.hpp:166-231)
      Aer<R> d, z;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        d.n[m] = z.n[m] = 0.0;
#pragma unroll
        for (int a = 0; a < 7; ++a) d.q[a][m] = z.q[a][m] = 0.0;
        d.q[A_SO4][m] = dc_so4[m];
        d.q[A_SOA][m] = dc_soa[m];
      }
#pragma unroll
      for (int a = 0; a < 7; ++a) {
        if (a == A_POM || a == A_BC || a == A_MOM) {
          const R qt = S0.q[a][NPCA] * xf;
          d.q[a][NPCA] -= qt * fc;
          d.q[a][NACC] += qt * fc;
        } else {
          d.q[a][NACC] += d.q[a][NPCA];
          d.q[a][NPCA] = 0.0;
        }
      }
      const R nt = S0.n[NPCA] * xf;
      d.n[NPCA] -= nt * fc;
      d.n[NACC] += nt * fc;
      D->add_aer(0, d, z);
    }
  } else {
    age_pcarbon(K, cfg.nmono, dgn_a[NPCA], S, dc_so4[NPCA], dc_soa[NPCA], cin_so4, cin_soa);
  }
}

// clear_post: the whole of it for one sub-step over deltat (the reference's
// ntsubstep = 1).  g1/g: gases {SOAG, H2SO4} before gas chemistry / current.
template <class P>
__device__ __forceinline__ void clear_post(
    const ChainConsts &K, const Cfg &cfg, double deltat, typename P::R temp, typename P::R pmid,
    typename P::R zmid, typename P::R pblh, typename P::R relhum,
    const typename P::R dgn_a[4], const typename P::R dgn_awet[4],
    const typename P::R wetdens[4], const typename P::W g1[2], typename P::W g[2],
    Aer<typename P::R> &S, const typename P::W uptk[4], typename P::W soa_g,
    typename P::W soa_gavg, const typename P::W soa_a[4], const ProcSink *D = nullptr) {
  const typename P::W netprod = clear_post_begin(cfg, deltat, g1, g);
  clear_post_sub<P>(K, cfg, (double)(int)(deltat / 1), netprod, temp, pmid, zmid, pblh, relhum,
                    dgn_a, dgn_awet, wetdens, g, S, uptk, soa_g, soa_gavg, soa_a, D);
}

// ------------------------------------------------------------------ calcsize
// CalcSize::compute_tendencies for one level (calcsize.hpp:962-1218) on the
// post-chain prognostics (kg/kg, #/kg), in the wide type.  Tendencies are
// returned as separate contributions, zero where calcsize does not touch a
// field.
template <class W>
struct CalcOut {
  W dn_i[4], dn_c[4];
  W dq_i[4][7], dq_c[4][7];
  W dgn_i[4], dgn_c[4];
};

template <class W>
__device__ __forceinline__ W dgn_from_v2n(const ChainConsts &K, int m, W v2n) {
  return cbrt((1.0 / v2n) / (PI / 6)) * K.dgn_from_vol[m];
}

template <class W>
__device__ __forceinline__ void update_diameter(const ChainConsts &K, int m, W drv, W num,
                                                W &dgn, double v2nmin, double v2nmax) {
  if (!(drv > 0.0)) return;
  if (num <= drv * v2nmin)
    dgn = MODE_DMAX[m];
  else if (num >= drv * v2nmax)
    dgn = MODE_DMIN[m];
  else
    dgn = dgn_from_v2n(K, m, num / drv);
}

template <class T>
__device__ __forceinline__ T safe_den(T x) {
  const double tol = 2.220446049250313e-16;
  return x / (x * x + tol * tol);
}

// calcsize.hpp:94-311
template <class W>
__device__ __forceinline__ void adjust_num_sizes(W drv_i, W drv_c, W init_i, W init_c, double dt,
                                                 double v2nmin, double v2nmax, double tsinv,
                                                 W &ni, W &nc, W &dqdt, W &dqqcwdt) {
  const double frac = kmax(0.0, kmin(1.0, dt * tsinv));
  const double dtinv = safe_den(dt);
  const double minrl = v2nmin / 27.0, maxrl = v2nmax * 27.0;
  if (drv_i <= 0.0 && drv_c <= 0.0) {
    ni = 0.0;
    nc = 0.0;
  } else if (drv_c <= 0.0) {
    nc = 0.0;
    const W nb = mmb(drv_i * v2nmin, drv_i * v2nmax, ni);
    ni = ni + (nb - ni) * frac;
  } else if (drv_i <= 0.0) {
    const W nb = mmb(drv_c * v2nmin, drv_c * v2nmax, nc);
    nc = nc + (nb - nc) * frac;
    ni = 0.0;
  } else {
    const W ni1 = ni, nc1 = nc;
    W nb = mmb(drv_i * minrl, drv_i * maxrl, ni1);
    const W dni2 = (nb - ni1) * frac;
    W ni2 = ni1 + dni2;
    nb = mmb(drv_c * minrl, drv_c * maxrl, nc1);
    const W dnc2 = (nb - nc1) * frac;
    W nc2 = nc1 + dnc2;
    if (dni2 == 0.0 && dnc2 != 0.0)
      ni2 = mmb(drv_i * minrl, drv_i * maxrl, ni1 - dnc2);
    else if (dni2 != 0.0 && dnc2 == 0.0)
      nc2 = mmb(drv_c * minrl, drv_c * maxrl, nc1 - dni2);
    const W tdrv = drv_i + drv_c, tnum = ni2 + nc2;
    W dni3 = 0.0, dnc3 = 0.0;
    const W minb = tdrv * v2nmin, maxb = tdrv * v2nmax;
    if (tnum < minb) {
      const W d3 = (minb - tnum) * frac;
      if ((ni2 < drv_i * v2nmin) && (nc2 < drv_c * v2nmin)) {
        dni3 = d3 * (ni2 / tnum);
        dnc3 = d3 * (nc2 / tnum);
      } else if (nc2 < drv_c * v2nmin) {
        dnc3 = d3;
      } else if (ni2 < drv_i * v2nmin) {
        dni3 = d3;
      }
    } else if (tnum > maxb) {
      const W d3 = (maxb - tnum) * frac;
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
template <class W>
__device__ __forceinline__ void new_size_after_transfer(const ChainConsts &K, int m, W drv, W num,
                                                        W &dgn) {
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

// e3sm: modal_aero_calcsize_sub (modal_aero_calcsize.hpp:748-928), the
// E3SM-layout calcsize, instead of CalcSize::compute_tendencies: the same
// algorithm on the same quantities except that the diameter update divides
// the Aitken mode's num2vol_ratio_min and multiplies the accumulation mode's
// num2vol_ratio_max by szadj_block_fac = 1e6 (:373-382), and the
// accumulation -> Aitken transfer counts the non-transferable species'
// number with num2vol_ratio_max (:219-222).
template <class W>
__device__ __forceinline__ void calcsize(const ChainConsts &K, bool do_xfer, bool do_adjust,
                                         double dt, const W n_i[4], const W n_c[4],
                                         const W q_i[4][7], const W q_c[4][7], CalcOut<W> &o,
                                         bool e3sm = false) {
  const double tsinv = safe_den(kmax(86400.0, dt));
  W dv_i[4], dv_c[4], nk_i[4], nk_c[4];
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
    W vi = 0, vc = 0;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      if (s >= NSPEC_MODE[m]) continue;
      vi += kmax(0.0, q_i[m][s]) * K.inv_density[m][s];
      vc += kmax(0.0, q_c[m][s]) * K.inv_density[m][s];
    }
    dv_i[m] = vi;
    dv_c[m] = vc;
    W ni = n_i[m] < 0 ? W(0.0) : n_i[m];
    W nc = n_c[m] < 0 ? W(0.0) : n_c[m];
    const bool aa = (m == NACC || m == NAIT);
    if (do_adjust && !(aa && do_xfer))
      adjust_num_sizes(vi, vc, n_i[m], n_c[m], dt, K.v2n_min[m], K.v2n_max[m], tsinv, ni, nc,
                       o.dn_i[m], o.dn_c[m]);
    double v2nmin = K.v2n_min[m], v2nmax = K.v2n_max[m];
    if (e3sm) {
      if (m == NAIT) v2nmin /= 1.0e6;
      if (m == NACC) v2nmax *= 1.0e6;
    }
    update_diameter(K, m, vi, ni, o.dgn_i[m], v2nmin, v2nmax);
    update_diameter(K, m, vc, nc, o.dgn_c[m], v2nmin, v2nmax);
    nk_i[m] = ni;
    nk_c[m] = nc;
  }
  if (!do_xfer) return;
  // aitken_accum_exchange (calcsize.hpp:594-826)
  const double v2n_ait = K.v2n_nom[NAIT], v2n_acc = K.v2n_nom[NACC];
  const double gm = K.v2n_geomean;
  bool a2c = false, c2a = false;
  W xt00 = 0, xt01 = 0, xt10 = 0, xt11 = 0;
  W xfv_a2c = 0, xfv_c2a = 0;
  W drv_i_noxf = 0, drv_c_noxf = 0;
  {
    W fn = 0, fv = 0;
    const W drv_t = dv_i[NAIT] + dv_c[NAIT];
    const W num_t = nk_i[NAIT] + nk_c[NAIT];
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
      const W xfn = fn * tsinv;
      xfv_a2c = fv * tsinv;
      xt00 = nk_i[NAIT] * xfn;
      xt01 = nk_c[NAIT] * xfn;
    }
  }
  {
    W fn = 0, fv = 0, num_t0 = 0;
    W drv_t = dv_i[NACC] + dv_c[NACC];
    W num_t = nk_i[NACC] + nk_c[NACC];
    if (drv_t > 0.0 && num_t > drv_t * gm) {
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        // accumulation species absent from Aitken: pom, bc, dst
        if (!(s == 1 || s == 3 || s == 4)) continue;
        drv_i_noxf += kmax(0.0, q_i[NACC][s]) * K.inv_density[NACC][s];
        drv_c_noxf += kmax(0.0, q_c[NACC][s]) * K.inv_density[NACC][s];
      }
      const W drv_t_noxf = drv_i_noxf + drv_c_noxf;
      // QUIRK (calcsize.hpp:438): num2vol_ratio_min; modal_aero_calcsize_sub
      // uses num2vol_ratio_max
      const W num_t_noxf = drv_t_noxf * (e3sm ? K.v2n_max[NACC] : K.v2n_min[NACC]);
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
      const W xfn = fn * tsinv;
      xfv_c2a = fv * tsinv;
      xt10 = nk_i[NACC] * xfn;
      xt11 = nk_c[NACC] * xfn;
    }
  }
  if (!(a2c || c2a)) return;
  {
    const W ndi = (xt00 - xt10) * dt;
    const W vdi = (dv_i[NAIT] * xfv_a2c - (dv_i[NACC] - drv_i_noxf) * xfv_c2a) * dt;
    const W ndc = (xt01 - xt11) * dt;
    const W vdc = (dv_c[NAIT] * xfv_a2c - (dv_c[NACC] - drv_c_noxf) * xfv_c2a) * dt;
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
      const W ti = kmax(0.0, q_i[NAIT][sa]) * xfv_a2c;
      o.dq_i[NAIT][sa] -= ti;
      o.dq_i[NACC][sc] += ti;
      const W tc = kmax(0.0, q_c[NAIT][sa]) * xfv_a2c;
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
      const W ti = kmax(0.0, q_i[NACC][sc]) * xfv_c2a;
      o.dq_i[NACC][sc] -= ti;
      o.dq_i[NAIT][sa] += ti;
      const W tc = kmax(0.0, q_c[NACC][sc]) * xfv_c2a;
      o.dq_c[NACC][sc] -= tc;
      o.dq_c[NAIT][sa] += tc;
    }
  }
}

}  // namespace dev
}  // namespace mam4x
