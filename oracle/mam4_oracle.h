/*
 * CPU FP64 oracle for the MAM4 aerosol-microphysics chain.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (eagles-project/mam4xx, src/mam4xx/mam4_amicphys.hpp and the
 * process headers it calls, plus calcsize.hpp), written to be read side by
 * side with it.  It is the checker for the HIP library and the CPU baseline
 * that bench.py times; nothing in the shipped library links or calls it.
 *
 * Units inside the chain are the reference's: gases and aerosol masses in
 * mol/mol-air, numbers in #/kmol-air.  The composed step (oracle_chain_step)
 * takes the AeroProcess prognostic units (kg/kg, #/kg) and converts.
 */
#ifndef MAM4X_ORACLE_H
#define MAM4X_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NMODE 4
#define ORC_NAER 7
#define ORC_NGAS 2
#define ORC_GAS_PCNST 31
#define ORC_NQTENDAA 4
#define ORC_NSTATE 52

typedef struct {
  int do_cond;
  int do_rename;
  int do_newnuc;
  int do_coag;
  int gaexch_h2so4_uptake_optaa;
  int newnuc_h2so4_conc_optaa;
  unsigned n_so4_monolayers_pcage;
  int do_calcsize;
  int calcsize_do_aitacc_transfer;
  int calcsize_do_adjust;
  /* condensation sub-steps per step (mam4_amicphys.hpp:1288 fixes 1; 0 is
   * taken as 1): the subarea chain runs ntsubstep times with an int sub-step
   * of deltat / ntsubstep, the H2SO4 uptake rates of the first held */
  int ntsubstep;
} orc_config;

/* Per-cell chain driver: modal_aero_amicphys_intr (mam4_amicphys.hpp:2119).
 * qq/qqcw are gas_pcnst(31)-vectors in chain units, updated in place.
 * qgcm_tendaa / qqcwgcm_tendaa (may be NULL) receive the per-process
 * tendency diagnostics [31][4] (cond, rename, newnuc, coag) and [31]
 * (cloud-borne rename), as written to the optional View2Ds at :2365-2374.
 * stats (may be NULL): [0] SOA sub-step iterations, [1] nsubarea. */
void orc_amicphys_intr(const orc_config *cfg, double deltat, double temp,
                       double pmid, double pdel, double zm, double pblh,
                       double qv, double cld, double qq[31], double qqcw[31],
                       const double q_pregaschem[31],
                       const double q_precldchem[31],
                       const double qqcw_precldchem[31],
                       const double dgncur_a[4], const double dgncur_awet[4],
                       const double wetdens_host[4], double *qgcm_tendaa,
                       double *qqcwgcm_tendaa, int stats[2]);

/* One call of mam_amicphys_1subarea (mam4_amicphys.hpp:1126-1694): gases
 * {SOAG, H2SO4}, masses [7][4] in the amicphys species order x mode,
 * increments over (cond, rename, newnuc, coag).  Returns SOA sub-steps. */
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
                          double *qnum_coag_aged, double *qaer_coag_aged);

/* Test switch for the air terms of the gas diffusivity: 0 = the reference's
 * constants (default), 1 = the Fortran MAM4 values (1/28.966, 20.1^(1/3),
 * 1.01325e5 Pa/atm). */
void orc_set_fortran_air_constants(int on);

/* Improved Magnus saturation vapour pressure over water with the moist-air
 * enhancement factor [Pa] (conversions.hpp:148-180). */
double orc_magnus_es(double T, double P);

/* Renaming's erfc tail fraction of the Aitken -> accumulation pair at
 * geometric-mean diameter `diameter` (number, or volume with volume != 0);
 * *cutoff (may be NULL) receives the pair's cutoff diameter. */
double orc_rename_tail_fraction(double diameter, int volume, double *cutoff);

/* Per-level calcsize: the body of CalcSize::compute_tendencies
 * (calcsize.hpp:962-1218) for one level k.  Inputs are prognostics in the
 * AeroProcess layout (n[mode], q[mode][species-in-mode]); tendencies must be
 * zeroed by the caller (the reference accumulates into them). */
void orc_calcsize_level(const orc_config *cfg, double dt, const double n_i[4],
                        const double n_c[4], const double q_i[4][7],
                        const double q_c[4][7], double dn_i[4], double dn_c[4],
                        double dq_i[4][7], double dq_c[4][7], double dgn_i[4],
                        double dgn_c[4]);

/* Composed microphysics step on [field][level][column] buffers (column
 * fastest, leading dimension ncol).  state: ORC_NSTATE fields in kg/kg and
 * #/kg, ordered as documented in include/mam4x.h.  Outputs: tend (ORC_NSTATE
 * fields, per second), dgn_i/dgn_c (4 fields each).  Optional extended
 * pre-chemistry inputs (NULL = use the current state).  nthreads <= 0: use
 * the OpenMP default.  Returns total SOA sub-step iterations. */
long orc_chain_step(const orc_config *cfg, int ncol, int nlev, double dt,
                    const double *state, const double *temperature,
                    const double *pressure, const double *height,
                    const double *qv, const double *cloud_fraction,
                    const double *pblh, const double *dgncur_a,
                    const double *dgncur_awet, const double *wetdens,
                    const double *pregaschem_gas, const double *precldchem,
                    double *tend, double *dgn_i, double *dgn_c, int nthreads);

/* orc_chain_step plus the per-process tendency diagnostics of
 * modal_aero_amicphys_intr (mam4_amicphys.hpp:2343-2374): proc holds 5
 * blocks (condensation with aging, renaming, nucleation, coagulation,
 * cloud-borne renaming) of 31 gas_pcnst planes, [5][31][nlev][ncol], in
 * mol/mol/s (#/kmol/s for numbers); NULL skips them. */
long orc_chain_step_proc(const orc_config *cfg, int ncol, int nlev, double dt,
                         const double *state, const double *temperature,
                         const double *pressure, const double *height,
                         const double *qv, const double *cloud_fraction,
                         const double *pblh, const double *dgncur_a,
                         const double *dgncur_awet, const double *wetdens,
                         const double *pregaschem_gas, const double *precldchem,
                         double *tend, double *dgn_i, double *dgn_c,
                         double *proc, int nthreads);

/* Conditioning probe: seed != 0 perturbs every transcendental result by a
 * random -1/0/+1 ulp (per-cell reproducible); 0 switches it off. */
void orc_set_libm_noise(unsigned long long seed);

/* SOA exchange of one subarea (gasaerexch_soaexch.hpp:79-341) in isolation;
 * species index 0 = SOA, 2 = POM in the reference's qaer.  Returns the
 * sub-step count. */
int orc_soaexch(double dt, double temp, double pmid, double *q_soag, double *q_soag_avg,
                double soa[4], const double pom[4], const double uptk_soag[4]);

/* ---- process-level entry points used by the known-answer tests ---- */

/* H2SO4 non-volatile analytic update (gasaerexch.hpp:479-537; identical math
 * to mam_gasaerexch_1subarea_1gas_nonvolatile, gasaerexch.hpp:139-218). */
void orc_nonvolatile_update(double dt, double netprod, const double uptkaer[4],
                            double *qgas_cur, double *qgas_avg,
                            double qaer_cur[4]);

/* Coagulation coefficient wrapper getcoags_wrapper_f (coagulation.hpp:751). */
void orc_getcoags_wrapper(double airtemp, double airprs, double dgatk,
                          double dgacc, double sgatk, double sgacc,
                          double xxlsgat, double xxlsgac, double pdensat,
                          double pdensac, double out[4]);
double orc_bm0ij(int n1, int n2a, int n2n);
double orc_bm3i(int n1, int n2a, int n2n);

/* Vehkamaki (2002) eqs 11-15 (vehkamaki2002.hpp:53-188). */
double orc_vehk_xcrit(double c_h2so4, double temp, double rh);
double orc_vehk_logj(double c_h2so4, double temp, double rh, double xcrit);
double orc_vehk_ntot(double c_h2so4, double temp, double rh, double xcrit);
double orc_vehk_radius(double xcrit, double ntot);
double orc_vehk_threshold(double temp, double rh);

/* Nucleation box model mer07_veh02_nuc_mosaic_1box (nucleation.hpp:346). */
void orc_nuc_mosaic_1box(int flag, double dtnuc, double temp, double rh,
                         double press, double zm, double pblh,
                         double qh2so4_cur, double qh2so4_avg, double qnh3_cur,
                         double h2so4_uptkrate, double mw_so4a_host, int nsize,
                         double dp_lo, double dp_hi, double rgas, double avogad,
                         double mw_nh4a, double mw_so4a, double pi,
                         double out[8]);

/* Aging transfer helpers (aging.hpp:166-231), arrays of length 4. */
void orc_transfer_aged_pcarbon_to_accum(int nsrc, int ndest, double xferfrac,
                                        double frac_cond, double frac_coag,
                                        double q[4], double dcond[4],
                                        double dcoag[4]);
void orc_transfer_cond_coag_mass_to_accum(int nsrc, int ndest, double q[4],
                                          double dcond[4], double dcoag[4]);

/* Full pcarbon aging of one subarea (aging.hpp:238-358). */
void orc_pcarbon_aging_1subarea(unsigned nmono, const double dgn_a[4],
                                double qnum[4], double qnum_del_cond[4],
                                double qnum_del_coag[4], double qaer[7][4],
                                double qaer_del_cond[7][4],
                                double qaer_del_coag[7][4],
                                double qaer_del_coag_in[7][1]);

/* Lognormal conversions (conversions.hpp:299-317). */
double orc_mean_particle_diameter_from_volume(double vol, double sigma);
double orc_mean_particle_volume_from_diameter(double dgn, double sigma);

/* ---- water uptake (water_uptake.hpp) ---- */

/* makoh_quartic (water_uptake.hpp:124-169): the four complex roots of
 * x^4 + p3 x^3 + p2 x^2 + p1 x + p0. */
void orc_makoh_quartic(double p3, double p2, double p1, double p0, double re[4],
                       double im[4]);

/* The Kohler solve of modal_aero_kohler (water_uptake.hpp:178-229) for a
 * given Kelvin coefficient aa [micron]: dry radius [micron] -> wet radius
 * [micron] (small-particle approximation or quartic + find_real_solution),
 * before the 30-micron bound. */
double orc_kohler_root(double aa, double rdry, double hygro, double rh);

/* wv_sat_qsat_water (wv_sat_methods.hpp:114-133): Goff & Gratch saturation
 * vapour pressure over water es [Pa] and saturation specific humidity qs. */
void orc_qsat_water(double t, double p, double *es, double *qs);

/* One cell of modal_aero_water_uptake_dr with wet density
 * (water_uptake.hpp:362-505): interstitial mass mixing ratios [kg/kg] per
 * mode in mode_aero_species order, T [K], p [Pa], qv [kg/kg], cloud
 * fraction, dry diameters [m] -> wet diameters [m], aerosol water [kg/kg],
 * wet density [kg/m3]. */
void orc_water_uptake_cell(const double qaer[4][7], double temperature, double pmid,
                           double qv, double cldn, const double dgncur_a[4],
                           double dgncur_awet[4], double qaerwat[4], double wetdens[4]);

/* orc_water_uptake_cell over [field][level][column] buffers (state as for
 * orc_chain_step; dgncur_a and the outputs [4][nlev][ncol]; wetdens may be
 * NULL). */
void orc_water_uptake(int ncol, int nlev, const double *state, const double *temperature,
                      const double *pressure, const double *qv, const double *cloud_fraction,
                      const double *dgncur_a, double *dgncur_awet, double *qaerwat,
                      double *wetdens);

/* ---- the wet-size producer ---- */

/* modal_aero_calcsize_sub (modal_aero_calcsize.hpp:748-928) for one level:
 * the E3SM-layout calcsize (arguments as orc_calcsize_level; do_adjust and
 * do_aitacc_transfer on).  Differs from CalcSize in the size adjustment's
 * szadj_block_fac and the non-transferable number's num2vol ratio. */
void orc_calcsize_sub_level(double dt, const double n_i[4], const double n_c[4],
                            const double q_i[4][7], const double q_c[4][7], double dn_i[4],
                            double dn_c[4], double dq_i[4][7], double dq_c[4][7],
                            double dgn_i[4], double dgn_c[4]);

/* compute_calcsize_and_water_uptake_dr (modal_aero_opt.hpp:510-533) over
 * [field][level][column] buffers; dgn_c, wetdens, tend may be NULL. */
void orc_wet_size(int ncol, int nlev, double dt, const double *state,
                  const double *temperature, const double *pressure, const double *qv,
                  const double *cloud_fraction, double *dgn_i, double *dgn_c,
                  double *dgncur_awet, double *qaerwat, double *wetdens, double *tend);

/* ---- stratospheric ozone (linoz) and invariants ---- */

/* compute_o3_column_density (compute_o3_column_density.hpp:59-83) */
void orc_o3_column_density(int nlev, const double *pdel, const double *vmr_o3, double col0,
                           double *o3_col_dens);
/* lin_strat_chem_solve_kk with psc_activation (lin_strat_chem.hpp:31-222);
 * clim[8] in LinozData order; diag = do3_linoz, do3_linoz_psc, ss_o3 */
void orc_linoz_solve_kk(double o3col, double temperature, double sza, double pmid,
                        double delta_t, double rlats, const double clim[8],
                        double chlorine_loading, double psc_T, double *o3_vmr, double diag[3]);
/* lin_strat_sfcsink (lin_strat_chem.hpp:263-316); returns Tg/yr/m2 */
double orc_linoz_sfcsink(int nlev, double delta_t, const double *pdel, double *o3l_vmr,
                         double o3_sfc, int o3_lbl, double o3_tau);
/* one column: column density, solve above the tropopause, surface sink */
void orc_linoz_column(int nlev, double dt, double sza, double rlats, int ltrop, double col0,
                      double chlorine_loading, double psc_T, double o3_sfc, double o3_tau,
                      int o3_lbl, const double *temp, const double *pmid, const double *pdel,
                      const double *clim, double *o3, double *o3col, double *do3,
                      double *do3_psc, double *ss, double *sfcsink);
/* setinv_single_level (mo_setinv.hpp:45-86) and its loop over cells */
void orc_setinv_single_level(double tfld, double h2ovmr, double pmid,
                             const double cnst_offline[4], double inv[8]);
void orc_setinv(long ncell, const double *tfld, const double *pmid, const double *qv,
                const double *cnst_offline, double *inv);
void orc_linoz_domain(int ncol, int nlev, double dt, double chlorine_loading, double psc_T,
                      double o3_sfc, double o3_tau, int o3_lbl, const double *temp,
                      const double *pmid, const double *pdel, const double *clim,
                      const double *sza, const double *rlats, const double *col0,
                      const int *ltrop, double *o3, double *o3col, double *do3,
                      double *do3_psc, double *ss, double *sfcsink);

/* ---- washout (mo_sethet.hpp) and external forcing (mo_setext.hpp) ---- */
void orc_precip_rescale(int nlev, const double *cmfdqr, const double *nrain,
                        const double *nevapr, double *precip);
void orc_gas_washout(int nlev, int ktop, double xkgm, const double *xliq, const double *rain,
                     const double *xhen_i, const double *tfld_i, const double *delz_i,
                     double *xgas);
int orc_find_ktop(int nlev, double rlat, const double *press);
double orc_sethet_xkgm(void);
void orc_sethet_column(int nlev, double rlat, double phis, double delt, const double *press,
                       const double *zmid, const double *tfld, const double *cmfdqr,
                       const double *nrain, const double *nevapr, const double *inv_m,
                       const double *qin_h2o2, const double *qin_so2, double *het);
void orc_sethet_domain(int ncol, int nlev, double delt, const double *rlat, const double *phis,
                       const double *press, const double *zmid, const double *tfld,
                       const double *cmfdqr, const double *nrain, const double *nevapr,
                       const double *inv_m, const double *qin_h2o2, const double *qin_so2,
                       double *het);
void orc_setext_column(int nlev, const int frc_ndx[9], const int file_alt_data[9],
                       const int nsectors[9], const double *const sectors[9][4],
                       double *frcing);

/* ---- photolysis (mo_photo.hpp) ---- */
/* the photolysis tables (PhotoTableData, mo_photo.hpp:28-56) for the one
 * reaction (phtcnt = 1): row-major rsf_tab [nw][nump][numsza][numcolo3]
 * [numalb] and xsqy [numj][nw][nt][np_xs]; pht_alias_mult and lng_index are
 * the reaction's pht_alias_mult_1 and lng_indexer entries */
typedef struct {
  int nw, nt, np_xs, numj, nump, numsza, numcolo3, numalb;
  const double *xsqy;
  const double *sza, *del_sza, *alb, *del_alb, *press, *del_p, *colo3, *o3rat, *del_o3rat;
  const double *etfphot, *prs, *dprs;
  const double *rsf_tab;
  double pht_alias_mult;
  int lng_index;
} orc_photo_table;
int orc_photo_find_index(const double *var, int len, double vmin, int idx0);
void orc_photo_cloud_mod(int nlev, double zen_angle, const double *clouds, const double *lwc,
                         const double *delp, double srf_alb, double *eff_alb, double *cld_mult);
void orc_photo_interpolate_rsf(const orc_photo_table *t, int nlev, const double *alb_in,
                               double sza_in, const double *p_in, const double *colo3_in,
                               double *rsf);
void orc_photo_jlong(const orc_photo_table *t, int nlev, double sza_in, const double *alb_in,
                     const double *p_in, const double *t_in, const double *colo3_in,
                     double *j_long);
void orc_table_photo(const orc_photo_table *t, int nlev, double *photo, const double *pmid,
                     const double *pdel, const double *temper, const double *colo3_in,
                     double zen_angle, double srf_alb, const double *lwc, const double *clouds,
                     double esfact);
void orc_table_photo_domain(const orc_photo_table *t, int ncol, int nlev, double *photo,
                            const double *pmid, const double *pdel, const double *temper,
                            const double *colo3_in, const double *zen_angle,
                            const double *srf_alb, const double *lwc, const double *clouds,
                            double esfact);

#ifdef __cplusplus
}
#endif
#endif
