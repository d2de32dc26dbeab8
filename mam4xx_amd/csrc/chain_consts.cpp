// Host evaluation of the chain's derived constants (see chain_consts.hpp).
// Each value is computed with the same expression the reference evaluates
// per call, so the kernels see the value the reference would have.
#include <cmath>

#include "chain_consts.hpp"
#include "whitby_coag_tables.inc"

namespace mam4x {

static double sq(double x) { return x * x; }

void chain_consts_init(ChainConsts *c) {
  for (int m = 0; m < 4; ++m) {
    c->alnsg[m] = std::log(MODE_SIGMA[m]);
    c->dgn_from_vol[m] = std::exp(-1.5 * sq(std::log(MODE_SIGMA[m])));
    c->vol_from_dgn[m] = std::exp(4.5 * sq(std::log(MODE_SIGMA[m])));
  }
  // gasaerexch.hpp:222-236 with mw = 98.0784, vd = 42.88 (gasaerexch.hpp:404-407)
  const double onethird = 1.0 / 3.0;
  c->h2so4_sqrt_mw = std::sqrt(1. / 9.807840000000000E+01 + 0.035);
  c->h2so4_vd_term = std::pow((std::pow(4.288000000000000E+01, onethird) + 2.7189), 2.0);
  c->tworootpi = 2 * std::sqrt(PI);
  c->root2 = std::sqrt(2.0);
  // mam4_amicphys.hpp:991-1003
  c->dplom = std::exp(0.67 * std::log(MODE_DMIN[NAIT]) + 0.33 * std::log(MODE_DNOM[NAIT]));
  c->dphim = MODE_DMAX[NAIT];
  const double tmpa = 1770 * PI / 6.0;
  c->mass1p_aitlo = tmpa * (std::pow(c->dplom, 3.0));
  c->mass1p_aithi = tmpa * (std::pow(c->dphim, 3.0));
  c->ln_1e38 = std::log(1e38);
  c->ln_1em38 = std::log(1.0e-38);
  // coagulation.hpp:650-717 (sigma-only part of getcoags)
  c->dlgsqt2 = 1.0 / std::log(std::sqrt(2.0));
  static const double bm0[10] = {0.707106785165097, 0.726148960080488, 0.766430744110958,
                                 0.814106389441342, 0.861679526483207, 0.903600509090092,
                                 0.936578814219156, 0.960098926735545, 0.975646823342881,
                                 0.985397173215326};  // coagulation.hpp:588-591
  const int src[3] = {NAIT, NPCA, NAIT};
  const int dst[3] = {NACC, NACC, NPCA};
  for (int ip = 0; ip < 3; ++ip) {
    ChainConsts::CoagPair &p = c->pair[ip];
    const double sgat = MODE_SIGMA[src[ip]], sgac = MODE_SIGMA[dst[ip]];
    const double xxlsgat = std::log(sgat), xxlsgac = std::log(sgac);
    p.esat01 = std::exp(0.125 * xxlsgat * xxlsgat);
    p.esac01 = std::exp(0.125 * xxlsgac * xxlsgac);
    p.esat04 = std::pow(p.esat01, 4.0);
    p.esac04 = std::pow(p.esac01, 4.0);
    p.esat05 = p.esat04 * p.esat01;
    p.esac05 = p.esac04 * p.esac01;
    p.esat08 = p.esat04 * p.esat04;
    p.esac08 = p.esac04 * p.esac04;
    p.esat09 = p.esat08 * p.esat01;
    p.esac09 = p.esac08 * p.esac01;
    p.esat16 = p.esat08 * p.esat08;
    p.esac16 = p.esac08 * p.esac08;
    p.esat20 = p.esat16 * p.esat04;
    p.esac20 = p.esac16 * p.esac04;
    const double esat24 = p.esat20 * p.esat04;
    p.esat25 = p.esat20 * p.esat05;
    p.esac25 = p.esac20 * p.esac05;
    p.esat36 = p.esat20 * p.esat16;
    p.esat49 = esat24 * p.esat25;
    p.esat64 = p.esat20 * p.esat20 * esat24;
    p.esat100 = p.esat64 * p.esat36;
    const int n2n = (int)(std::fmax(1.0, std::fmin(10.0, std::round(4.0 * (sgat - 0.75))))) - 1;
    const int n2a = (int)(std::fmax(1.0, std::fmin(10.0, std::round(4.0 * (sgac - 0.75))))) - 1;
    for (int n1 = 0; n1 < 10; ++n1) {
      // the accessors are called as bm0ij_data(n1, n2n, n2a) on a table
      // declared [n1][n2a][n2n] (coagulation.hpp:306, 559)
      p.bm0ij[n1] = WHITBY_BM0IJ[n1][n2n][n2a];
      p.bm3i[n1] = WHITBY_BM3I[n1][n2n][n2a];
    }
    p.bm0_at = bm0[n2n];
    p.bm0_ac = bm0[n2a];
    p.dumatk_fac = std::exp(4.5 * xxlsgat * xxlsgat);
  }
  // aging.hpp:99-151
  {
    const double mw_soa = 150 / 1000.0, mw_so4 = 115 / 1000.0;
    c->so4_vol = mw_so4 * 1000.0 / AER_DENS[A_SO4];
    const double soa_vol = mw_soa * 1000.0 / AER_DENS[A_SOA];
    const double hygro_soa = 0.14000000000000001, hygro_so4 = 0.50700000000000001;
    c->fac_m2v_eqvhyg = soa_vol * hygro_soa / hygro_so4;
    c->fac_volsfc = std::exp(2.5 * sq(std::log(MODE_SIGMA[NPCA])));
    c->xferfrac_max = 1.0 - 10.0 * 2.220446049250313e-16;
  }
  // rename.hpp:392-474 for the Aitken -> accumulation pair
  {
    const int m = NAIT, d = NACC;
    const double alnsg = std::log(MODE_SIGMA[m]);
    c->rn_tail_fac = std::sqrt(0.5) / alnsg;
    const double pio6 = PI / 6;
    const double vmin = (MODE_DMIN[m] * MODE_DMIN[m] * MODE_DMIN[m]) *
                        std::exp(4.5 * sq(std::log(MODE_SIGMA[m]))) * pio6;
    const double vmax = (MODE_DMAX[m] * MODE_DMAX[m] * MODE_DMAX[m]) *
                        std::exp(4.5 * sq(std::log(MODE_SIGMA[m]))) * pio6;
    c->rn_v2n_lo_rlx = 1.0 / vmin * 27.0;
    c->rn_v2n_hi_rlx = 1.0 / vmax / 27.0;
    c->rn_ln_tail_fac = 3.0 * sq(alnsg);
    const double alnsg_d = std::log(MODE_SIGMA[d]);
    const double cutoff = std::sqrt(MODE_DNOM[m] * std::exp(1.5 * sq(alnsg)) * MODE_DNOM[d] *
                                    std::exp(1.5 * sq(alnsg_d)));
    c->rn_ln_cutoff = std::log(cutoff);
    c->rn_threshold = 0.99 * cutoff;
  }
  // calcsize.hpp:893-960
  for (int m = 0; m < 4; ++m) {
    const double pio6 = PI / 6;
    const double f = std::exp(4.5 * sq(std::log(MODE_SIGMA[m])));
    c->v2n_nom[m] = 1.0 / ((MODE_DNOM[m] * MODE_DNOM[m] * MODE_DNOM[m]) * f * pio6);
    c->v2n_min[m] = 1.0 / ((MODE_DMAX[m] * MODE_DMAX[m] * MODE_DMAX[m]) * f * pio6);
    c->v2n_max[m] = 1.0 / ((MODE_DMIN[m] * MODE_DMIN[m] * MODE_DMIN[m]) * f * pio6);
    for (int s = 0; s < 7; ++s)
      c->inv_density[m][s] = s < NSPEC_MODE[m] ? 1.0 / AER_DENS[mode_spec(m, s)] : 0.0;
  }
  c->v2n_geomean = std::sqrt(c->v2n_nom[NAIT]) * std::sqrt(c->v2n_nom[NACC]);
  c->soa_c1 = 156.0e3 / 8.3144;
  c->soa_inv298 = 1.0 / 298.0;
}

}  // namespace mam4x
