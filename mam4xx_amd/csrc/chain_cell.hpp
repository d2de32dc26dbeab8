// One cell of the microphysics step: subarea partition, the chain on the
// clear and cloudy subareas, recombination, calcsize, tendencies.
// Follows modal_aero_amicphys_intr (mam4_amicphys.hpp:2119-2376) and
// CalcSize::compute_tendencies (calcsize.hpp:962-1218).
#pragma once

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
  const ChainConsts *K;
  unsigned long long *stats;  // optional: SOA sub-step count accumulator
  int ncol, nlev, ldc;
  int do_calcsize, do_xfer, do_adjust, has_dgn, has_sout;
  double dt;
  dev::Cfg cfg;
};

namespace dev {

// get_partition_factors (mam4_amicphys.hpp:492-546)
__device__ __forceinline__ void partition_factors(double qi, double qc, double fcldy,
                                                  double fclea, double &fa, double &fc) {
  const double qc_cldy = qc / fcldy;
  const double qi_cldy = kmax(0.0, ((qi + qc) - qc_cldy));
  const double qi_clea = (qi - fcldy * qi_cldy) / fclea;
  double r = kmax(1.e-35, qi_clea * fclea) / kmax(1.e-35, qi);
  r = kmax(0.0, kmin(1.0, r));
  fa = r / fclea;
  fc = (1 - r) / fcldy;
}

template <bool PRE>
__device__ __forceinline__ int chain_cell(const ChainArgs &a, long off, int icol) {
  const ChainConsts &K = *a.K;
  const Cfg cfg = a.cfg;
  const double dt = a.dt;
  const double T = a.T[off], p = a.p[off], z = a.z[off], qv = a.qv[off];
  const double cld = a.cld[off], pblh = a.pblh[icol];

  // relative humidity from the improved Magnus formula (conversions.hpp:148-180)
  const double epsqs = MW_H2O / MW_DRY_AIR;
  const double tc = T - 273.15;
  const double ew = 610.94 * exp(17.625 * tc / (234.04 + tc));
  const double ev = 1.00071 * exp(4.5e-8 * p) * ew;
  const double qvs = epsqs * ev / (p - (1 - epsqs) * ev);
  const double rhgcm = kmax(0.0, kmin(1.0, qv / qvs));

  // setup_subareas (mam4_amicphys.hpp:241-317): 0 clear, 1 overcast, 2 partly
  int cat;
  double fcldy;
  if (cld < 1.0e-5) {
    cat = 0;
    fcldy = 0;
  } else if (cld > 0.999) {
    cat = 1;
    fcldy = 1;
  } else {
    cat = 2;
    fcldy = cld;
  }
  const double fclea = 1.0 - fcldy;
  const double rh_clear = (cat == 2) ? mmb(0.0, 1.0, (rhgcm - fcldy) / fclea) : rhgcm;

  double dga[4], dgw[4], wd[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    dga[m] = a.dga[m][off];
    dgw[m] = a.dgw[m][off];
    wd[m] = kmax(1000.0, a.wd[m][off]);
  }

  // grid-cell means in chain units, clipped at 0 (mam4_amicphys.hpp:2258-2267)
  auto ld3 = [&](int f, double c) { return kmax(0.0, a.state[f][off] * c); };
  auto ld2 = [&](int f, double c) {
    return PRE ? kmax(0.0, (a.precld[f] ? a.precld[f][off] : a.state[f][off]) * c)
               : ld3(f, c);
  };
  double h1, s1;
  if (PRE) {
    h1 = kmax(0.0, (a.pre_h2so4 ? a.pre_h2so4[off] : a.state[F_H2SO4][off]) * TO_VMR_H2SO4);
    s1 = kmax(0.0, (a.pre_soag ? a.pre_soag[off] : a.state[F_SOAG][off]) * TO_VMR_SOAG);
  } else {
    h1 = ld3(F_H2SO4, TO_VMR_H2SO4);
    s1 = ld3(F_SOAG, TO_VMR_SOAG);
  }
  const double h3 = ld3(F_H2SO4, TO_VMR_H2SO4), s3 = ld3(F_SOAG, TO_VMR_SOAG);
  const double h2 = ld2(F_H2SO4, TO_VMR_H2SO4), s2 = ld2(F_SOAG, TO_VMR_SOAG);

  // partition factors of the interstitial aerosol (partly cloudy cells)
  double fa_n[4], fc_n[4], fa_q[4], fc_q[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    fa_n[m] = fc_n[m] = fa_q[m] = fc_q[m] = 0.0;
    if (cat == 2) {
      partition_factors(ld2(f_ni(m), TO_VMR_NUM), ld2(f_nc(m), TO_VMR_NUM), fcldy, fclea,
                        fa_n[m], fc_n[m]);
      double qi = 0.0, qc = 0.0;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        qi = qi + ld2(f_qi(m, s), to_vmr_mass(m, s));
        qc = qc + ld2(f_qc(m, s), to_vmr_mass(m, s));
      }
      partition_factors(qi, qc, fcldy, fclea, fa_q[m], fc_q[m]);
    }
  }

  // subarea results in chain units before the area weighting:
  // interstitial + gases of the clear (oc_*) and cloudy (od_*) subareas
  int niter = 0;
  double gi_out[NSTATE];  // grid-cell mean after the chain, chain units
  const bool map_gas = cfg.do_cond || cfg.do_newnuc;

  if (cat != 1) {
    // ---- clear subarea ----
    Aer S;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double nn = ld3(f_ni(m), TO_VMR_NUM);
      S.n[m] = (cat == 2 ? nn * fa_n[m] : nn) * 1.0;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        const double qq = ld3(f_qi(m, s), to_vmr_mass(m, s));
        S.q[sp][m] = (cat == 2 ? qq * fa_q[m] : qq) * fcvt_aer(sp);
      }
    }
    S.q[A_SO4][NPCA] = 0.0;
    S.q[A_SOA][NPCA] = 0.0;
    S.q[A_NCL][NPCA] = 0.0;
    // gases: before gas chem (1) and current (3) in the clear subarea
    double hcl = (cat == 2) ? h2 : h3, scl = (cat == 2) ? s2 : s3;
    double hcd = 0.0, scd = 0.0;  // cloudy-subarea gases (cat 2)
    if (cat == 2) {
      hcd = (h3 - fclea * hcl) / fcldy;
      if (hcd < 0) {
        hcd = 0;
        hcl = h3 / fclea;
      }
      scd = (s3 - fclea * scl) / fcldy;
      if (scd < 0) {
        scd = 0;
        scl = s3 / fclea;
      }
    }
    double g[2] = {0.0, 0.0};
    const double g1[2] = {map_gas ? s1 * FCVT_SOAG : 0.0, map_gas ? h1 * 1.0 : 0.0};
    if (map_gas) {
      g[0] = scl * FCVT_SOAG;
      g[1] = hcl * 1.0;
    }
    niter = clear_subarea(K, cfg, dt, T, p, z, pblh, rh_clear, dga, dgw, wd, g1, g, S);
    // map back (mam4_amicphys.hpp:1940-1969) and weight by the clear fraction
    const double hout = map_gas ? g[1] / 1.0 : hcl;
    const double sout = map_gas ? g[0] / FCVT_SOAG : scl;
    gi_out[F_H2SO4] = hout * fclea;
    gi_out[F_SOAG] = sout * fclea;
    if (cat == 2) {
      gi_out[F_H2SO4] += hcd * fcldy;
      gi_out[F_SOAG] += scd * fcldy;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      gi_out[f_ni(m)] = (S.n[m] / 1.0) * fclea;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        gi_out[f_qi(m, s)] = (S.q[sp][m] / fcvt_aer(sp)) * fclea;
      }
    }
  } else {
    gi_out[F_H2SO4] = h3 * 1.0;
    gi_out[F_SOAG] = s3 * 1.0;
  }

  if (cat != 0) {
    // ---- cloudy subarea: renaming of interstitial + cloud-borne only ----
    Aer C, Cc;
    double gi[7] = {0, 0, 0, 0, 0, 0, 0}, gc[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double n3 = ld3(f_ni(m), TO_VMR_NUM);
      const double nc3 = ld3(f_nc(m), TO_VMR_NUM);
      C.n[m] = (cat == 2 ? n3 * fc_n[m] : n3) * 1.0;
      Cc.n[m] = (cat == 2 ? nc3 / fcldy : nc3) * 1.0;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        const double q3 = ld3(f_qi(m, s), to_vmr_mass(m, s));
        const double qc3 = ld3(f_qc(m, s), to_vmr_mass(m, s));
        C.q[sp][m] = (cat == 2 ? q3 * fc_q[m] : q3) * fcvt_aer(sp);
        Cc.q[sp][m] = (cat == 2 ? qc3 / fcldy : qc3) * fcvt_aer(sp);
        if (m == NAIT) {
          const double q2 = ld2(f_qi(m, s), to_vmr_mass(m, s));
          const double qc2 = ld2(f_qc(m, s), to_vmr_mass(m, s));
          const double a2 = (cat == 2 ? q2 * fc_q[m] : q2) * fcvt_aer(sp);
          const double c2 = (cat == 2 ? qc2 / fcldy : qc2) * fcvt_aer(sp);
          gi[sp] = (C.q[sp][m] - a2) / 1;
          gc[sp] = (Cc.q[sp][m] - c2) / 1;
        }
      }
    }
    if (cfg.do_rename) rename_ait_to_acc(K, true, C, gi, Cc, gc);
    const double w = fcldy;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double ni = (C.n[m] / 1.0) * w;
      gi_out[f_ni(m)] = (cat == 2) ? gi_out[f_ni(m)] + ni : ni;
      gi_out[f_nc(m)] = (Cc.n[m] / 1.0) * w;
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        const int sp = mode_spec(m, s);
        const double qi = (C.q[sp][m] / fcvt_aer(sp)) * w;
        gi_out[f_qi(m, s)] = (cat == 2) ? gi_out[f_qi(m, s)] + qi : qi;
        gi_out[f_qc(m, s)] = (Cc.q[sp][m] / fcvt_aer(sp)) * w;
      }
    }
  } else {
    // no cloudy subarea: cloud-borne aerosol keeps its (clipped) value
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      gi_out[f_nc(m)] = ld3(f_nc(m), TO_VMR_NUM);
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        gi_out[f_qc(m, s)] = ld3(f_qc(m, s), to_vmr_mass(m, s));
      }
    }
  }

  // back to prognostic units; interstitial and gases are clipped at 0
  // (form_gcm_of_gases_and_aerosols_from_subareas, :2045-2047)
  double post[NSTATE];
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

  // tendencies of the chain, plus calcsize on the post-chain state
  double tend[NSTATE];
#pragma unroll
  for (int f = 0; f < NSTATE; ++f) tend[f] = (post[f] - a.state[f][off]) / dt;
  double dgi[4], dgc[4];
  if (a.do_calcsize) {
    double n_i[4], n_c[4], q_i[4][7], q_c[4][7];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      n_i[m] = post[f_ni(m)];
      n_c[m] = post[f_nc(m)];
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        q_i[m][s] = s < NSPEC_MODE[m] ? post[f_qi(m, s)] : 0.0;
        q_c[m][s] = s < NSPEC_MODE[m] ? post[f_qc(m, s)] : 0.0;
      }
    }
    CalcOut o;
    calcsize(K, a.do_xfer != 0, a.do_adjust != 0, dt, n_i, n_c, q_i, q_c, o);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      tend[f_ni(m)] += o.dn_i[m];
      tend[f_nc(m)] += o.dn_c[m];
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        if (s >= NSPEC_MODE[m]) continue;
        tend[f_qi(m, s)] += o.dq_i[m][s];
        tend[f_qc(m, s)] += o.dq_c[m][s];
      }
      dgi[m] = o.dgn_i[m];
      dgc[m] = o.dgn_c[m];
    }
  } else {
#pragma unroll
    for (int m = 0; m < 4; ++m) dgi[m] = dgc[m] = dga[m];
  }
#pragma unroll
  for (int f = 0; f < NSTATE; ++f) a.tend[f][off] = tend[f];
  if (a.has_sout) {
#pragma unroll
    for (int f = 0; f < NSTATE; ++f) a.sout[f][off] = a.state[f][off] + dt * tend[f];
  }
  if (a.has_dgn) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      if (a.dgi[m]) a.dgi[m][off] = dgi[m];
      if (a.dgc[m]) a.dgc[m][off] = dgc[m];
    }
  }
  return niter;
}

}  // namespace dev
}  // namespace mam4x
