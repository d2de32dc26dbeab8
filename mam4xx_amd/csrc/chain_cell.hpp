// One cell of the microphysics step: subarea partition, the chain on the
// clear and cloudy subareas, recombination, calcsize, tendencies.
// Follows modal_aero_amicphys_intr (mam4_amicphys.hpp:2119-2376) and
// CalcSize::compute_tendencies (calcsize.hpp:962-1218).
//
// The step is organised in three phases so the data-dependent SOA loop can
// run in a work-queue kernel of its own:
//   cell_setup + clear_build + clear_pre  -> SoaIn    (phase 1)
//   soa_init / soa_tick ...               -> results  (phase 2)
//   cell_setup + clear_build + clear_post + cell_finish (phase 3)
// chain_cell_fused runs the three phases back to back in one lane.
#pragma once

#include <type_traits>

#include "chain_device.hpp"

namespace mam4x {

// Kernel argument block (passed by value; every pointer is a device plane).
struct ChainArgs {
  const double *state[NSTATE];
  const double *T, *p, *z, *qv, *cld, *pblh;
  const double *dga[4], *dgw[4], *wd[4];
  const double *pre_h2so4, *pre_soag;
  const double *precld[NSTATE];
  double *tend[NSTATE];
  double *dgi[4], *dgc[4];
  double *sout[NSTATE];
  // tangent mode: directions of the inputs (NULL = 0) and the derivative of
  // the tendencies along them
  const double *dstate[NSTATE];
  const double *dT, *dp, *dqv;
  double *dtend[NSTATE];
  // vector-Jacobian product by a seeded tangent pass: when xbar is set, the
  // lane writes sum_f tbar[f] * d(tendency f) to xbar instead of dtend
  // (tbar[f] NULL = 0)
  const double *tbar[NSTATE];
  double *xbar;
  // per-process tendency diagnostics (dev::ProcSink; P::diag only): block k
  // of 31 gas_pcnst planes, tracer i at proc[k] + i * pstride + off
  double *proc[5];
  long pstride;
  // per cell: 1 when a SOA loop stopped at its sub-step cap (NULL = not formed)
  int *soacap;
  const ChainConsts *K;
  int ncol, nlev, ldc;
  int layout;  // 0: [nlev][ldc] column fastest; 1: [ncol][ldc] level fastest
  int lpw;     // layout 0: levels per run of 64 cells (1: 64 columns of one level)
  int do_calcsize, do_xfer, do_adjust, has_dgn, has_sout;
  double dt;
  dev::Cfg cfg;
};

namespace dev {

__device__ __forceinline__ ProcSink proc_sink(const ChainArgs &a, long off, double afrac) {
  ProcSink D;
#pragma unroll
  for (int k = 0; k < 5; ++k) D.blk[k] = a.proc[k];
  D.stride = a.pstride;
  D.off = off;
  D.deltat = a.dt;
  D.afrac = afrac;
  return D;
}

// an input value in the wide type: with its direction in tangent mode
template <class P>
__device__ __forceinline__ typename P::W ldw(const double *p, const double *dp, long off) {
  using W = typename P::W;
  if constexpr (std::is_same<W, num::Dual>::value)
    return num::Dual(p[off], dp ? dp[off] : 0.0);
  else
    return W(p[off]);
}

// get_partition_factors (mam4_amicphys.hpp:492-546)
template <class R>
__device__ __forceinline__ void partition_factors(R qi, R qc, double fcldy, double fclea, R &fa,
                                                  R &fc) {
  const R qc_cldy = qc / fcldy;
  const R qi_cldy = kmax(0.0, ((qi + qc) - qc_cldy));
  const R qi_clea = (qi - fcldy * qi_cldy) / fclea;
  R r = kmax(1.e-35, qi_clea * fclea) / kmax(1.e-35, qi);
  r = kmax(0.0, kmin(1.0, r));
  fa = r / fclea;
  fc = (1 - r) / fcldy;
}

// grid-cell means in chain units, clipped at 0 (mam4_amicphys.hpp:2258-2267);
// the product is formed in the wide type
template <class P>
__device__ __forceinline__ typename P::W ld3w(const ChainArgs &a, long off, int f, double c) {
  return kmax(0.0, ldw<P>(a.state[f], a.dstate[f], off) * c);
}
template <class P>
__device__ __forceinline__ typename P::R ld3(const ChainArgs &a, long off, int f, double c) {
  return num::to<typename P::R>(ld3w<P>(a, off, f, c));
}
template <bool PRE, class P>
__device__ __forceinline__ typename P::W ld2w(const ChainArgs &a, long off, int f, double c) {
  if (PRE && a.precld[f]) return kmax(0.0, ldw<P>(a.precld[f], nullptr, off) * c);
  return ld3w<P>(a, off, f, c);
}
template <bool PRE, class P>
__device__ __forceinline__ typename P::R ld2(const ChainArgs &a, long off, int f, double c) {
  return num::to<typename P::R>(ld2w<PRE, P>(a, off, f, c));
}

// Per-cell context shared by the phases.
template <class P>
struct CellCtx {
  using R = typename P::R;
  using W = typename P::W;
  R T, p, z, pblh;
  int cat;  // setup_subareas (mam4_amicphys.hpp:241-317): 0 clear, 1 overcast, 2 partly
  double fcldy, fclea;
  R rh_clear;
  R dga[4], dgw[4], wd[4];
  W h1, s1, h2, s2, h3, s3;  // gases before gas chem / before cloud chem / now
  R fa_n[4], fc_n[4], fa_q[4], fc_q[4];
  bool map_gas;
};

template <bool PRE, class P>
__device__ __forceinline__ void cell_setup(const ChainArgs &a, long off, int icol,
                                           CellCtx<P> &X) {
  using R = typename P::R;
  using W = typename P::W;
  const W Tw = ldw<P>(a.T, a.dT, off), pw = ldw<P>(a.p, a.dp, off);
  X.T = num::to<R>(Tw);
  X.p = num::to<R>(pw);
  X.z = a.z[off];
  X.pblh = a.pblh[icol];
  const R qv = num::to<R>(ldw<P>(a.qv, a.dqv, off));
  const double cld = a.cld[off];
  // relative humidity from the improved Magnus formula (conversions.hpp:148-180)
  const double epsqs = MW_H2O / MW_DRY_AIR;
  const R tc = X.T - 273.15;
  const R ew = 610.94 * exp(17.625 * tc / (234.04 + tc));
  const R ev = 1.00071 * exp(4.5e-8 * X.p) * ew;
  const R qvs = epsqs * ev / (X.p - (1 - epsqs) * ev);
  const R rhgcm = kmax(0.0, kmin(1.0, qv / qvs));
  if (cld < 1.0e-5) {
    X.cat = 0;
    X.fcldy = 0;
  } else if (cld > 0.999) {
    X.cat = 1;
    X.fcldy = 1;
  } else {
    X.cat = 2;
    X.fcldy = cld;
  }
  X.fclea = 1.0 - X.fcldy;
  X.rh_clear = (X.cat == 2) ? R(mmb(0.0, 1.0, (rhgcm - X.fcldy) / X.fclea)) : rhgcm;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    X.dga[m] = a.dga[m][off];
    X.dgw[m] = a.dgw[m][off];
    X.wd[m] = kmax(1000.0, R(a.wd[m][off]));
  }
  if (PRE) {
    X.h1 = a.pre_h2so4 ? kmax(0.0, W(a.pre_h2so4[off]) * TO_VMR_H2SO4)
                       : ld3w<P>(a, off, F_H2SO4, TO_VMR_H2SO4);
    X.s1 = a.pre_soag ? kmax(0.0, W(a.pre_soag[off]) * TO_VMR_SOAG)
                      : ld3w<P>(a, off, F_SOAG, TO_VMR_SOAG);
  } else {
    X.h1 = ld3w<P>(a, off, F_H2SO4, TO_VMR_H2SO4);
    X.s1 = ld3w<P>(a, off, F_SOAG, TO_VMR_SOAG);
  }
  X.h3 = ld3w<P>(a, off, F_H2SO4, TO_VMR_H2SO4);
  X.s3 = ld3w<P>(a, off, F_SOAG, TO_VMR_SOAG);
  X.h2 = ld2w<PRE, P>(a, off, F_H2SO4, TO_VMR_H2SO4);
  X.s2 = ld2w<PRE, P>(a, off, F_SOAG, TO_VMR_SOAG);
  // partition factors of the interstitial aerosol (partly cloudy cells)
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    X.fa_n[m] = X.fc_n[m] = X.fa_q[m] = X.fc_q[m] = 0.0;
    if (X.cat == 2) {
      partition_factors(ld2<PRE, P>(a, off, f_ni(m), TO_VMR_NUM),
                        ld2<PRE, P>(a, off, f_nc(m), TO_VMR_NUM), X.fcldy, X.fclea, X.fa_n[m],
                        X.fc_n[m]);
      R qi = 0.0, qc = 0.0;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        qi = qi + ld2<PRE, P>(a, off, f_qi(m, s), to_vmr_mass(m, s));
        qc = qc + ld2<PRE, P>(a, off, f_qc(m, s), to_vmr_mass(m, s));
      }
      partition_factors(qi, qc, X.fcldy, X.fclea, X.fa_q[m], X.fc_q[m]);
    }
  }
  X.map_gas = a.cfg.do_cond || a.cfg.do_newnuc;
}

// Clear-subarea state (set_subarea_gases_and_aerosols, mam4_amicphys.hpp:
// 678-911, and the mapping of mam_amicphys_1gridcell, :1784-1851).
template <class W>
struct ClearGas {
  W g1[2], g[2];  // {SOAG, H2SO4} before gas chem / current, chain units
  W hcl, scl;     // clear-subarea gases before mapping
  W hcd, scd;     // cloudy-subarea gases (partly cloudy cells)
};

template <class P>
__device__ __forceinline__ void clear_build(const ChainArgs &a, long off, const CellCtx<P> &X,
                                            Aer<typename P::R> &S, ClearGas<typename P::W> &G) {
  using R = typename P::R;
  using W = typename P::W;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const R nn = ld3<P>(a, off, f_ni(m), TO_VMR_NUM);
    S.n[m] = (X.cat == 2 ? nn * X.fa_n[m] : nn) * 1.0;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      if (s >= NSPEC_MODE[m]) continue;
      const int sp = mode_spec(m, s);
      const R qq = ld3<P>(a, off, f_qi(m, s), to_vmr_mass(m, s));
      S.q[sp][m] = (X.cat == 2 ? qq * X.fa_q[m] : qq) * fcvt_aer(sp);
    }
  }
  S.q[A_SO4][NPCA] = 0.0;
  S.q[A_SOA][NPCA] = 0.0;
  S.q[A_NCL][NPCA] = 0.0;
  G.hcl = (X.cat == 2) ? X.h2 : X.h3;
  G.scl = (X.cat == 2) ? X.s2 : X.s3;
  G.hcd = 0.0;
  G.scd = 0.0;
  if (X.cat == 2) {
    G.hcd = (X.h3 - X.fclea * G.hcl) / X.fcldy;
    if (G.hcd < 0) {
      G.hcd = 0;
      G.hcl = X.h3 / X.fclea;
    }
    G.scd = (X.s3 - X.fclea * G.scl) / X.fcldy;
    if (G.scd < 0) {
      G.scd = 0;
      G.scl = X.s3 / X.fclea;
    }
  }
  G.g1[0] = X.map_gas ? X.s1 * FCVT_SOAG : W(0.0);
  G.g1[1] = X.map_gas ? X.h1 * 1.0 : W(0.0);
  G.g[0] = X.map_gas ? G.scl * FCVT_SOAG : W(0.0);
  G.g[1] = X.map_gas ? G.hcl * 1.0 : W(0.0);
}

// Cloudy subarea (renaming only), recombination, calcsize and the writes.
// S/G: the clear subarea after clear_post (unused for overcast cells).  The
// grid-cell sums, the calcsize step and the tendencies are in the wide type.
template <bool PRE, class P, bool MULTI = false>
__device__ __forceinline__ void cell_finish(const ChainArgs &a, long off, const CellCtx<P> &X,
                                            const Aer<typename P::R> &S,
                                            const ClearGas<typename P::W> &G) {
  using R = typename P::R;
  using W = typename P::W;
  const ChainConsts &K = *a.K;
  const double dt = a.dt;
  W gi_out[NSTATE];  // grid-cell mean after the chain, chain units
  if (X.cat != 1) {
    // map back (mam4_amicphys.hpp:1940-1969) and weight by the clear fraction
    const W hout = X.map_gas ? G.g[1] / 1.0 : G.hcl;
    const W sout = X.map_gas ? G.g[0] / FCVT_SOAG : G.scl;
    gi_out[F_H2SO4] = hout * X.fclea;
    gi_out[F_SOAG] = sout * X.fclea;
    if (X.cat == 2) {
      gi_out[F_H2SO4] += G.hcd * X.fcldy;
      gi_out[F_SOAG] += G.scd * X.fcldy;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      gi_out[f_ni(m)] = (num::to<W>(S.n[m]) / 1.0) * X.fclea;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        gi_out[f_qi(m, s)] = (num::to<W>(S.q[sp][m]) / fcvt_aer(sp)) * X.fclea;
      }
    }
  } else {
    gi_out[F_H2SO4] = X.h3 * 1.0;
    gi_out[F_SOAG] = X.s3 * 1.0;
  }

  if (X.cat != 0) {
    // ---- cloudy subarea: renaming of interstitial + cloud-borne only ----
    Aer<R> C, Cc;
    R gi[7] = {0, 0, 0, 0, 0, 0, 0}, gc[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const R n3 = ld3<P>(a, off, f_ni(m), TO_VMR_NUM);
      const R nc3 = ld3<P>(a, off, f_nc(m), TO_VMR_NUM);
      C.n[m] = (X.cat == 2 ? n3 * X.fc_n[m] : n3) * 1.0;
      Cc.n[m] = (X.cat == 2 ? nc3 / X.fcldy : nc3) * 1.0;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        const R q3 = ld3<P>(a, off, f_qi(m, s), to_vmr_mass(m, s));
        const R qc3 = ld3<P>(a, off, f_qc(m, s), to_vmr_mass(m, s));
        C.q[sp][m] = (X.cat == 2 ? q3 * X.fc_q[m] : q3) * fcvt_aer(sp);
        Cc.q[sp][m] = (X.cat == 2 ? qc3 / X.fcldy : qc3) * fcvt_aer(sp);
        if (m == NAIT) {
          const R q2 = ld2<PRE, P>(a, off, f_qi(m, s), to_vmr_mass(m, s));
          const R qc2 = ld2<PRE, P>(a, off, f_qc(m, s), to_vmr_mass(m, s));
          const R a2 = (X.cat == 2 ? q2 * X.fc_q[m] : q2) * fcvt_aer(sp);
          const R c2 = (X.cat == 2 ? qc2 / X.fcldy : qc2) * fcvt_aer(sp);
          gi[sp] = (C.q[sp][m] - a2) / (MULTI ? nsub(a.cfg) : 1);
          gc[sp] = (Cc.q[sp][m] - c2) / (MULTI ? nsub(a.cfg) : 1);
        }
      }
    }
    // one renaming per sub-step, each with the sub-step's share of the
    // cloud-chemistry growth (mam4_amicphys.hpp:1446-1455)
    if (a.cfg.do_rename) {
      for (int j = 0; j < (MULTI ? nsub(a.cfg) : 1); ++j) {
        if constexpr (P::diag) {
          const Aer<R> C0 = C, Cc0 = Cc;
          rename_ait_to_acc<P>(K, true, C, gi, Cc, gc);
          const ProcSink D = proc_sink(a, off, X.fcldy);
          D.add_aer(1, C, C0);
          D.add_aer(4, Cc, Cc0);
        } else {
          rename_ait_to_acc<P>(K, true, C, gi, Cc, gc);
        }
      }
    }
    const double w = X.fcldy;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const W ni = (num::to<W>(C.n[m]) / 1.0) * w;
      gi_out[f_ni(m)] = (X.cat == 2) ? gi_out[f_ni(m)] + ni : ni;
      gi_out[f_nc(m)] = (num::to<W>(Cc.n[m]) / 1.0) * w;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        const W qi = (num::to<W>(C.q[sp][m]) / fcvt_aer(sp)) * w;
        gi_out[f_qi(m, s)] = (X.cat == 2) ? gi_out[f_qi(m, s)] + qi : qi;
        gi_out[f_qc(m, s)] = (num::to<W>(Cc.q[sp][m]) / fcvt_aer(sp)) * w;
      }
    }
  } else {
    // no cloudy subarea: cloud-borne aerosol keeps its (clipped) value
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      gi_out[f_nc(m)] = ld3w<P>(a, off, f_nc(m), TO_VMR_NUM);
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        gi_out[f_qc(m, s)] = ld3w<P>(a, off, f_qc(m, s), to_vmr_mass(m, s));
      }
    }
  }

  // back to prognostic units; interstitial and gases are clipped at 0
  // (form_gcm_of_gases_and_aerosols_from_subareas, :2045-2047)
  W post[NSTATE];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    post[f_ni(m)] = kmax(0.0, gi_out[f_ni(m)]) * TO_MMR_NUM;
    post[f_nc(m)] = gi_out[f_nc(m)] * TO_MMR_NUM;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      if (s >= NSPEC_MODE[m]) continue;
      post[f_qi(m, s)] = kmax(0.0, gi_out[f_qi(m, s)]) * to_mmr_mass(m, s);
      post[f_qc(m, s)] = gi_out[f_qc(m, s)] * to_mmr_mass(m, s);
    }
  }
  post[F_H2SO4] = kmax(0.0, gi_out[F_H2SO4]) * TO_MMR_H2SO4;
  post[F_SOAG] = kmax(0.0, gi_out[F_SOAG]) * TO_MMR_SOAG;

  // calcsize on the post-chain state, then the tendencies: (post - state)/dt
  // of the chain plus calcsize's own (zero where calcsize does not act)
  CalcOut<W> o;
  if (a.do_calcsize) {
    W n_i[4], n_c[4], q_i[4][7], q_c[4][7];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      n_i[m] = post[f_ni(m)];
      n_c[m] = post[f_nc(m)];
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        q_i[m][s] = s < NSPEC_MODE[m] ? post[f_qi(m, s)] : W(0.0);
        q_c[m][s] = s < NSPEC_MODE[m] ? post[f_qc(m, s)] : W(0.0);
      }
    }
    calcsize(K, a.do_xfer != 0, a.do_adjust != 0, dt, n_i, n_c, q_i, q_c, o);
  }
  double xbar = 0.0;  // tangent mode: the contraction with tbar
#pragma unroll
  for (int f = 0; f < NSTATE; ++f) {
    const W st = ldw<P>(a.state[f], a.dstate[f], off);
    W t = (post[f] - st) / dt;
    if (a.do_calcsize) {
      if (f < F_Q_I) {
        t += f < F_NUM_C ? o.dn_i[f] : o.dn_c[f - F_NUM_C];
      } else if (f < F_H2SO4) {
        const bool cw = f >= F_Q_C;
        const int r = f - (cw ? F_Q_C : F_Q_I);
        const int m = r < MODE_OFF[1] ? 0 : r < MODE_OFF[2] ? 1 : r < MODE_OFF[3] ? 2 : 3;
        const int sp = r - MODE_OFF[m];
        t += cw ? o.dq_c[m][sp] : o.dq_i[m][sp];
      }
    }
    if (a.tend[f]) a.tend[f][off] = num::val(t);
    if constexpr (std::is_same<W, num::Dual>::value) {
      if (a.xbar) {
        if (a.tbar[f]) xbar += a.tbar[f][off] * num::tan(t);
      } else {
        a.dtend[f][off] = num::tan(t);
      }
    }
    if (a.has_sout) a.sout[f][off] = num::val(st) + dt * num::val(t);
  }
  if constexpr (std::is_same<W, num::Dual>::value)
    if (a.xbar) a.xbar[off] = xbar;
  if (a.has_dgn) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double di = a.do_calcsize ? num::val(o.dgn_i[m]) : num::val(X.dga[m]);
      const double dc = a.do_calcsize ? num::val(o.dgn_c[m]) : num::val(X.dga[m]);
      if (a.dgi[m]) a.dgi[m][off] = di;
      if (a.dgc[m]) a.dgc[m][off] = dc;
    }
  }
}

// The whole step for one cell in one lane.  Returns the SOA sub-step count.
// MULTI: ntsubstep > 1 (a separate instantiation: the sub-step loop keeps
// the subarea's state live across sub-steps, which costs the one-sub-step
// kernel registers it does not need).
template <bool PRE, class P, bool MULTI = false>
__device__ __forceinline__ int chain_cell_fused(const ChainArgs &a, long off, int icol) {
  using W = typename P::W;
  const ChainConsts &K = *a.K;
  CellCtx<P> X;
  cell_setup<PRE, P>(a, off, icol, X);
  Aer<typename P::R> S;
  ClearGas<W> G;
  int niter = 0;
  bool capped = false;
  [[maybe_unused]] ProcSink D;
  if constexpr (P::diag) {
    D = proc_sink(a, off, X.fclea);
    D.zero();
  }
  if (X.cat != 1) {
    clear_build<P>(a, off, X, S, G);
    SoaIn<W> in = {};
    W sg = 0.0, sgavg = 0.0, sa[4] = {0, 0, 0, 0};
    if constexpr (!MULTI) {
      if (a.cfg.do_cond) {
        clear_pre<P>(K, X.T, X.p, X.dgw, S, G.g[0], in);
        SoaState<W> st;
        soa_run(st, in, (double)(int)(a.dt / 1));
        niter = st.niter;
        capped = st.niter > 1000;
        sg = st.g;
        sgavg = soa_gavg(st);
#pragma unroll
        for (int n = 0; n < 4; ++n) sa[n] = st.a[n];
      }
      clear_post<P>(K, a.cfg, a.dt, X.T, X.p, X.z, X.pblh, X.rh_clear, X.dga, X.dgw, X.wd, G.g1,
                    G.g, S, in.uptk, sg, sgavg, sa, &D);
    } else {
      if (a.cfg.do_cond) clear_pre<P>(K, X.T, X.p, X.dgw, S, G.g[0], in);
      const W netprod = clear_post_begin(a.cfg, a.dt, G.g1, G.g);
      const int ns = nsub(a.cfg);
      const double dtsub = dt_sub(a.cfg, a.dt);
      for (int j = 0; j < ns; ++j) {
        if (a.cfg.do_cond) {
          if (j > 0) soa_in_state<P>(S, G.g[0], in);
          SoaState<W> st;
          soa_run(st, in, dtsub);
          niter += st.niter;
          capped = capped || st.niter > 1000;
          sg = st.g;
          sgavg = soa_gavg(st);
#pragma unroll
          for (int n = 0; n < 4; ++n) sa[n] = st.a[n];
        }
        clear_post_sub<P>(K, a.cfg, dtsub, netprod, X.T, X.p, X.z, X.pblh, X.rh_clear, X.dga,
                          X.dgw, X.wd, G.g, S, in.uptk, sg, sgavg, sa, &D);
      }
    }
  }
  cell_finish<PRE, P, MULTI>(a, off, X, S, G);
  if (a.soacap) a.soacap[off] = capped ? 1 : 0;
  return niter;
}

// chain_cell_fused with the instantiation the configuration needs (host
// builds; the kernels choose it at launch).
template <bool PRE, class P>
__device__ __forceinline__ int chain_cell_step(const ChainArgs &a, long off, int icol) {
  return nsub(a.cfg) > 1 ? chain_cell_fused<PRE, P, true>(a, off, icol)
                         : chain_cell_fused<PRE, P, false>(a, off, icol);
}

}  // namespace dev
}  // namespace mam4x
