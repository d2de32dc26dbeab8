/*
 * mam4x.h -- C ABI of the MI355X MAM4 aerosol-microphysics chain.
 *
 * The chain is the per-cell sequence of mam4::microphysics::
 * modal_aero_amicphys_intr (eagles-project/mam4xx, src/mam4xx/
 * mam4_amicphys.hpp:2119-2376): gas-aerosol exchange (H2SO4 + SOAG), Aitken->
 * accumulation renaming, new-particle nucleation, coagulation and
 * primary-carbon aging, followed by the modal size recalculation of
 * CalcSize::compute_tendencies (calcsize.hpp:962-1218).
 *
 * In the reference the chain runs inside one Kokkos team per column, behind
 * AeroProcess<Impl>::compute_tendencies(team, t, dt, atm, sfc, progs, diags,
 * tends) (src/validation/aero_process.hpp:113-122).  A GPU kernel cannot be
 * launched from inside a device team, so this boundary sits one level up: one
 * host call covers every column of a domain, with the same argument set.  The
 * team argument becomes a HIP stream.
 *
 * Conventions
 *  - Every function returns an int status (MAM4X_OK = 0); no C++ exception
 *    crosses this boundary.  mam4x_last_error() describes the last failure.
 *  - All array arguments are caller-owned DEVICE pointers to double.  A
 *    "plane" is one field over all cells, in one of two layouts:
 *      MAM4X_LAYOUT_LEV_COL  [nlev][ld], column index fastest (ld >= ncol);
 *      MAM4X_LAYOUT_COL_LEV  [ncol][ld], level index fastest (ld >= nlev),
 *                            the layout of EAMxx fields (ncol x nlev).
 *    ld = 0 means the fastest extent.  Field tables list one pointer per
 *    field, so the prognostic fields may live in separate allocations (as
 *    EAMxx fields do) or in one [field][...][ld] block.  Cells are processed
 *    in memory order of the chosen layout, so both are read coalesced.
 *  - Work is enqueued on the given stream (a hipStream_t; NULL = the null
 *    stream).  Calls are stream-ordered and never synchronize the host,
 *    except mam4x_validate, which returns a host-side count.
 *  - One host thread per handle at a time; distinct handles are independent.
 *    The handle owns its device workspace.
 */
#ifndef MAM4X_H
#define MAM4X_H

#ifdef __cplusplus
extern "C" {
#endif

#define MAM4X_ABI_VERSION 11
#define MAM4X_MAX_NTSUBSTEP 64

/* per-process tendency diagnostics (mam4x_outputs.process_tendencies) */
#define MAM4X_GAS_PCNST 31
#define MAM4X_NPROCESS 5
#define MAM4X_PROC_CONDENSATION 0        /* gas_aero_exchange_condensation   */
#define MAM4X_PROC_RENAMING 1            /* gas_aero_exchange_renaming       */
#define MAM4X_PROC_NUCLEATION 2          /* gas_aero_exchange_nucleation     */
#define MAM4X_PROC_COAGULATION 3         /* gas_aero_exchange_coagulation    */
#define MAM4X_PROC_RENAMING_CLOUD_BORNE 4 /* ..._renaming_cloud_borne        */

/* column budgets (mam4x_outputs.column_budgets): entries 0..51 follow the
 * state fields; then the two closures the microphysics conserves */
#define MAM4X_NBUDGET 54
#define MAM4X_BUDGET_SULFUR 52  /* H2SO4 gas + every SO4 species  [mol/m2/s] */
#define MAM4X_BUDGET_ORGANIC 53 /* SOAG + every SOA species       [mol/m2/s] */

/* status codes */
#define MAM4X_OK 0
#define MAM4X_ERR_ARG 1         /* invalid argument (NULL pointer, size)   */
#define MAM4X_ERR_HIP 2         /* a HIP runtime call failed               */
#define MAM4X_ERR_UNSUPPORTED 3 /* option not implemented by this build    */
#define MAM4X_ERR_NOMEM 4       /* device workspace allocation failed      */

/* ---------------------------------------------------------------------- */
/* State layout: 52 prognostic fields per cell, AeroProcess units.         */
/*   [0,4)   interstitial number, modes acc, ait, coarse, pca    [#/kg]    */
/*   [4,8)   cloud-borne number                                  [#/kg]    */
/*   [8,29)  interstitial species mass, mode-major; species in   [kg/kg]   */
/*           mode_aero_species order (aero_modes.hpp:214-243):             */
/*           acc: so4 pom soa bc dst ncl mom; ait: so4 soa ncl mom;        */
/*           coarse: dst ncl so4 bc pom soa mom; pca: pom bc mom           */
/*   [29,50) cloud-borne species mass, same order                [kg/kg]   */
/*   50      H2SO4 gas, 51 SOAG gas                              [kg/kg]   */
/* Mass <-> molar conversion uses the E3SM molecular weights               */
/* (gas_chem_mechanism.hpp:31-37) and M_air = 28.966 g/mol; numbers        */
/* convert as #/kmol = #/kg * 28.966.                                      */
/* ---------------------------------------------------------------------- */
#define MAM4X_NSTATE 52
#define MAM4X_NMODE 4
#define MAM4X_F_NUM_I 0
#define MAM4X_F_NUM_C 4
#define MAM4X_F_Q_I 8
#define MAM4X_F_Q_C 29
#define MAM4X_F_H2SO4 50
#define MAM4X_F_SOAG 51

/* precision of a handle */
#define MAM4X_PREC_FP64 0  /* everything in FP64 (the drop-in)             */
#define MAM4X_PREC_MIXED 1 /* FP32 bulk; FP64 for the ill-conditioned parts */

/* plane layouts (see Conventions) */
#define MAM4X_LAYOUT_LEV_COL 0
#define MAM4X_LAYOUT_COL_LEV 1

/* schedule of a step (performance only; results are identical) */
#define MAM4X_SCHED_AUTO 0  /* fused below 16,384 cells, else work queue   */
#define MAM4X_SCHED_FUSED 1 /* one lane runs one cell's whole step         */
#define MAM4X_SCHED_QUEUE 2 /* per-cell phases around a work-queue kernel  */
                            /* that runs the SOA sub-stepping              */

/* Process switches.  Mirrors microphysics::AmicPhysConfig
 * (mam4_amicphys.hpp:223-237), Aging's n_so4_monolayers_pcage
 * (aging.hpp:20-32) and CalcSize::Config (calcsize.hpp:835-846). */
typedef struct mam4x_config {
  int do_cond;
  int do_rename;
  int do_newnuc;
  int do_coag;
  int gaexch_h2so4_uptake_optaa; /* 1 sequential, 2 simultaneous           */
  int newnuc_h2so4_conc_optaa;
  unsigned n_so4_monolayers_pcage; /* default 8                            */
  int do_calcsize;                 /* run calcsize after the chain         */
  int calcsize_do_aitacc_transfer;
  int calcsize_do_adjust;
  int precision; /* MAM4X_PREC_*                                          */
  int schedule;  /* MAM4X_SCHED_*                                         */
  int soa_blocks_per_cu; /* persistent SOA kernel: resident blocks per CU,
                            at most as many as fit; 0 = 3 (performance
                            only)                                        */
  int ntsubstep; /* condensation sub-steps per step, 1..MAM4X_MAX_NTSUBSTEP
                    (0 is taken as 1).  The reference fixes 1
                    (mam4_amicphys.hpp:1288).  With n > 1 each subarea's
                    chain (condensation, renaming, nucleation, coagulation,
                    aging) runs n times over the int sub-step dt / n; the
                    H2SO4 uptake rates and their sum are those of the first
                    sub-step, held for the others (the Fortran original's
                    scope of uptkaer / uptkrate_h2so4, which the C++ leaves
                    undefined after the first: :1344, :1377,
                    gasaerexch.hpp:428); the cloudy subarea's renaming takes
                    1/n of the cloud-chemistry growth per sub-step
                    (:1446-1455).  n > 1 runs the one-lane-per-cell schedule. */
  /* Performance only (results are identical; 0 = default):
   *  soa_chunk        queue entries a wave takes per refill from the long
   *                   part of the SOA work queue (default 64);
   *  soa_short_chunk  the same for the short part (default 256);
   *  soa_short_bucket first cost bucket of the long part, 1..24 (default 10:
   *                   predicted <= ~30 sub-steps go to the short part);
   *  levels_per_wave  1, 2, 4 or 8: the per-cell kernels give each wave a
   *                   tile of that many levels x 64 / n columns (layout
   *                   MAM4X_LAYOUT_LEV_COL, where the tiles cover the domain
   *                   exactly; otherwise 1 = 64 columns of one level). */
  int soa_chunk;
  int soa_short_chunk;
  int soa_short_bucket;
  int levels_per_wave;
} mam4x_config;

/* Atmosphere (atmosphere.hpp:16-137) and modal diagnostics the chain reads. */
typedef struct mam4x_inputs {
  int ncol;
  int nlev;
  int ldc;                            /* leading dimension of every plane   */
  int layout;                         /* MAM4X_LAYOUT_*                     */
  const double *state[MAM4X_NSTATE];  /* prognostics, one plane each        */
  const double *temperature;          /* [K]                                */
  const double *pressure;             /* mid-level pressure [Pa]            */
  const double *height;               /* mid-level height above surface [m] */
  const double *vapor_mixing_ratio;   /* qv [kg/kg]                         */
  const double *cloud_fraction;       /* [-]                                */
  const double *pblh;                 /* per column [ncol], PBL height [m]  */
  const double *dgncur_a[MAM4X_NMODE];    /* dry geo-mean diameter [m]      */
  const double *dgncur_awet[MAM4X_NMODE]; /* wet geo-mean diameter [m]      */
  const double *wetdens[MAM4X_NMODE];     /* wet density [kg/m3]            */
  /* Optional pre-chemistry states (mam4_amicphys.hpp:2136-2137).  NULL
   * entries mean "equal to the current state", the pure-microphysics step.
   * pregaschem_h2so4/soag: gases before gas chemistry; precldchem: all 52
   * fields before cloud chemistry (interstitial and cloud-borne). */
  const double *pregaschem_h2so4;
  const double *pregaschem_soag;
  const double *precldchem[MAM4X_NSTATE];
  /* Layer pressure thickness [Pa] (Atmosphere::hydrostatic_dp,
   * atmosphere.hpp:98).  Read only to form mam4x_outputs.column_budgets;
   * may be NULL otherwise. */
  const double *hydrostatic_dp;
} mam4x_inputs;

typedef struct mam4x_outputs {
  /* d(state)/dt of the whole step [units/s]: (chain result - state)/dt plus
   * the calcsize tendencies evaluated on the chain result.  Required. */
  double *tendencies[MAM4X_NSTATE];
  /* CalcSize diagnostics dry_geometric_mean_diameter_{i,c} (aero_config.hpp:
   * 209-213).  Optional (NULL to skip). */
  double *dgncur_i[MAM4X_NMODE];
  double *dgncur_c[MAM4X_NMODE];
  /* Optional: state after the step, state + dt * tendencies.  May alias the
   * input state planes (in-place update). */
  double *state_out[MAM4X_NSTATE];
  /* Optional per-process tendency diagnostics: the gas_aero_exchange_*
   * arrays of MicrophysDiagnosticArrays (diagnostic_arrays.hpp:13-18) that
   * modal_aero_amicphys_intr fills (mam4_amicphys.hpp:2343-2374).  Block
   * process_tendencies[k] (MAM4X_PROC_*; NULL = not formed) holds
   * MAM4X_GAS_PCNST planes, tracer i at process_tendencies[k] + i *
   * process_stride (0 = the plane extent: slow extent x ld), indexed by the
   * reference's gas_pcnst tracer order (mam4_amicphys.hpp:35-69; slots 0, 1,
   * 3, 4 are gases outside MAM4 and stay 0).  Units: the reference's tracer
   * mixing-ratio tendencies, mol/mol/s (#/kmol/s for numbers), area-weighted
   * over the clear and cloudy subareas.  The condensation block includes its
   * share of primary-carbon aging; the coagulation block is taken before
   * aging.  FP64 precision, single-domain calls only. */
  double *process_tendencies[MAM4X_NPROCESS];
  long long process_stride;
  /* Optional: per cell, 1 when a SOA sub-stepping loop of the step stopped
   * at its 1000-sub-step cap before the end of the (sub-)step, else 0 (the
   * reference warns and goes on: gasaerexch_soaexch.hpp:93, 170-173).  One
   * int plane in the tendencies' layout and leading dimension; NULL = not
   * formed. */
  int *soa_capped;
  /* Optional: per column, FP64 column integrals of the step's tendencies
   * over the levels, sum_k x(k) * dp(k) / g (dp = inputs.hydrostatic_dp,
   * required when this is set; g = 9.80616 m/s2, mam4_constants.hpp:48).
   * Entry b of column c at column_budgets[b * ncol + c], MAM4X_NBUDGET
   * entries:
   *   b < 52  x = tendency of state field b: kg/m2/s for masses, #/m2/s
   *           for numbers;
   *   MAM4X_BUDGET_SULFUR   x = sum over H2SO4 gas and the six SO4 fields
   *           (interstitial and cloud-borne) of tendency * 1000 / molecular
   *           weight [g/mol]: mol of sulfur /m2/s;
   *   MAM4X_BUDGET_ORGANIC  x = the same over SOAG and the SOA fields, in
   *           the gas-aerosol exchange's SOA basis (mam4_amicphys.hpp:
   *           131-154: SOAG * 12.011/150, SOA * 0.08, per 12.011 g/mol).
   * The chain moves sulfur and organics between the gas and the modes
   * without sources or sinks, so the two closures vanish to rounding, also
   * with the pre-chemistry inputs set (the chemistry's production is already
   * in the state the tendencies are taken against).  The level sum has a
   * fixed order in both
   * layouts, so results agree bit for bit across layouts and launches: the
   * levels k = l, l + 8, l + 16, ... form partial sum l (l = 0..7, fused
   * multiply-adds in ascending k, weight dp(k) * (1/g)), and the eight
   * partials are added as ((p0+p4)+(p2+p6))+((p1+p5)+(p3+p7)).  The closure
   * terms are summed over their fields in state order before the weight.
   * Formed by mam4x_compute_tendencies, the allocation-free variant and
   * batches (per scene; launches only, so a captured run forms them too), in
   * every precision; the tangent rejects it. */
  double *column_budgets;
} mam4x_outputs;

/* Forward-mode tangent: the direction of the inputs and the derivative of
 * the tendencies along it (mam4x_compute_tangent).  NULL input planes are
 * zero directions. */
typedef struct mam4x_tangent {
  const double *state[MAM4X_NSTATE];  /* d(prognostic), units of the state   */
  const double *temperature;          /* dT [K]                              */
  const double *pressure;             /* dp [Pa]                             */
  const double *vapor_mixing_ratio;   /* dqv [kg/kg]                         */
  double *tendencies[MAM4X_NSTATE];   /* out: d(tendency) along the direction */
} mam4x_tangent;

/* Vector-Jacobian product (mam4x_compute_adjoint): a cotangent of the
 * tendencies in, the gradients of <cotangent, tendencies> with respect to
 * the inputs out, cell by cell (cells are independent, so the Jacobian of a
 * step is block diagonal: one 52 x 55 block per cell).  NULL cotangent
 * planes are zero; NULL gradient planes are not formed. */
typedef struct mam4x_adjoint {
  const double *tendencies[MAM4X_NSTATE];  /* cotangent of each tendency     */
  double *state[MAM4X_NSTATE];             /* out: gradient w.r.t. the state */
  double *temperature;                     /* out: gradient w.r.t. T         */
  double *pressure;                        /* out: gradient w.r.t. p         */
  double *vapor_mixing_ratio;              /* out: gradient w.r.t. qv        */
} mam4x_adjoint;

/* Outputs of the aerosol water uptake (mam4x_compute_water_uptake), one
 * plane per mode (mode order of the state). */
typedef struct mam4x_water_outputs {
  double *dgncur_awet[4]; /* wet geometric mean diameter [m]           */
  double *qaerwat[4];     /* aerosol water mass mixing ratio [kg/kg]   */
  double *wetdens[4];     /* wet density [kg/m3] (NULL: not formed)    */
} mam4x_water_outputs;

/* Outputs of the wet-size producer (mam4x_compute_wet_size). */
typedef struct mam4x_wetsize_outputs {
  double *dgncur_i[4];    /* dry diameter, interstitial [m] (calcsize)         */
  double *dgncur_c[4];    /* dry diameter, cloud-borne [m] (NULL: not written) */
  double *dgncur_awet[4]; /* wet diameter [m] (water uptake)                   */
  double *qaerwat[4];     /* aerosol water [kg/kg]                             */
  double *wetdens[4];     /* wet density [kg/m3] (NULL: not formed)            */
  /* calcsize tendencies [1/s] (ptend / dqqcwdt) in state order, gases 0;
   * all NULL: not written */
  double *tendencies[MAM4X_NSTATE];
} mam4x_wetsize_outputs;

typedef struct mam4x_handle_s *mam4x_handle;

/* Library / handle lifetime.  mam4x_create replaces AeroProcess<Impl>'s
 * constructor + Impl::init (aero_process.hpp:34-43). */
int mam4x_abi_version(void);
int mam4x_create(const mam4x_config *config, mam4x_handle *out);
int mam4x_destroy(mam4x_handle h);
int mam4x_get_config(mam4x_handle h, mam4x_config *out);
int mam4x_set_precision(mam4x_handle h, int precision);
const char *mam4x_last_error(mam4x_handle h);
int mam4x_default_config(mam4x_config *out);

/* AeroProcess::init(const ProcessConfig &) (aero_process.hpp:76-80):
 * re-initializes the handle with a new configuration (NULL = defaults),
 * validated as by mam4x_create.  Later calls use it; the handle keeps its
 * workspace.  A handle with live batches (mam4x_batch_create) is refused
 * with MAM4X_ERR_ARG, since a batch fixes the configuration it was created
 * with.  On any error the handle keeps its previous configuration.
 * Host-side: call it between steps, not while one is in flight on another
 * thread. */
int mam4x_init(mam4x_handle h, const mam4x_config *config);

/* AeroProcess<Impl>::validate (aero_process.hpp:90-95; checks of
 * GasAerExch::validate, gasaerexch.hpp:96-120): counts cells with T < 0,
 * p < 0, qv < 0 or any negative prognostic.  Reads only the state planes,
 * temperature, pressure and vapor_mixing_ratio of *in (the reference's
 * validate takes no diagnostics): the other pointers may be NULL.
 * Synchronizes the stream and writes the count to *violations. */
int mam4x_validate(mam4x_handle h, const mam4x_inputs *in, long long *violations,
                   void *stream);

/* AeroProcess<Impl>::compute_tendencies (aero_process.hpp:113-122) for all
 * columns: one chain step of length dt [s] followed by calcsize.  t is the
 * model time, unused (as in the reference). */
int mam4x_compute_tendencies(mam4x_handle h, double t, double dt,
                             const mam4x_inputs *in, const mam4x_outputs *out,
                             void *stream);

/* mam4x_compute_tendencies without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same step and results as mam4x_compute_tendencies on the same handle; the
 * work-queue workspace lives in a fixed device array of the library
 * (MAM4X_NOALLOC_MAX_CELLS cells, 61.3 MB per GPU, held until the process
 * ends) instead of the handle's growing one, and no phase-timing events are
 * recorded.  Domains above MAM4X_NOALLOC_MAX_CELLS cells (ncol x nlev) are
 * rejected with MAM4X_ERR_ARG.  Every call on one GPU shares that array, whose
 * contents persist between calls: calls that overlap on one GPU (other
 * streams, other handles, other threads) race and must be ordered. */
#define MAM4X_NOALLOC_MAX_CELLS 524288
int mam4x_compute_tendencies_noalloc(mam4x_handle h, double t, double dt,
                                     const mam4x_inputs *in, const mam4x_outputs *out,
                                     void *stream);

/* Directional derivative of mam4x_compute_tendencies (no counterpart in the
 * reference, which has no derivative code): the same step, evaluated in
 * forward-mode dual numbers, writes the tendencies and diameters to *out
 * and their derivative along *dir to dir->tendencies.  The derivative is the
 * one of the branch each cell takes (the chain is piecewise smooth: clips,
 * regime switches and the SOA sub-step count are held at the primal's
 * choice).  FP64 only; one lane runs one cell's whole step. */
int mam4x_compute_tangent(mam4x_handle h, double t, double dt, const mam4x_inputs *in,
                          const mam4x_tangent *dir, const mam4x_outputs *out, void *stream);

/* Vector-Jacobian product of mam4x_compute_tendencies (no counterpart in the
 * reference): a reverse sweep of the FP64 step, one lane per cell.  The lane
 * recomputes its cell's step, keeping the state each SOA sub-step starts
 * from (6 doubles per sub-step), then back-propagates adj->tendencies
 * through every process in reverse order; all requested gradient planes
 * come out of that one sweep.  Derivatives are those of the branch each cell
 * takes, as for the tangent.  A first forward work-queue pass counts every
 * cell's SOA sub-steps; the stream is synchronised once to size the
 * sub-step records, so the call cannot be captured in a graph.  The records
 * and counters are held by the handle (grown on demand); calls on one handle
 * must be stream-ordered.  A cell whose recomputed loop does not retrace the
 * counted sub-steps gets NaN gradients rather than wrong ones.  The SOA
 * sub-steps are recomputed with the fast FP64 quotient throughout; where the
 * step itself falls back to IEEE quotients (SOA masses below 2^-400, inputs
 * out of range) the linearisation point can differ from the step's in the
 * last place of those masses. */
int mam4x_compute_adjoint(mam4x_handle h, double t, double dt, const mam4x_inputs *in,
                          const mam4x_adjoint *adj, void *stream);

/* Aerosol water uptake, the producer of the step's dgncur_awet / wetdens
 * inputs: modal_aero_water_uptake_dr with the wet density
 * (water_uptake.hpp:481-505; its callers are compute_calcsize_and_water_
 * uptake_dr, modal_aero_opt.hpp:510-533, and wet deposition) for every cell.
 * Reads the interstitial mass planes of in->state (kg/kg), temperature,
 * pressure, vapor_mixing_ratio (kg/kg), cloud_fraction and the dry
 * diameters in->dgncur_a; the other input planes may be NULL.  Hygroscopic
 * growth follows Kohler theory (closed-form quartic, as the reference), with
 * the clear-air RH from the Goff & Gratch saturation humidity and the
 * crystallisation / deliquescence hysteresis.  FP64; stream-ordered; no
 * workspace. */
int mam4x_compute_water_uptake(mam4x_handle h, const mam4x_inputs *in,
                               const mam4x_water_outputs *out, void *stream);

/* The wet-size producer, compute_calcsize_and_water_uptake_dr
 * (modal_aero_opt.hpp:510-533; also wet_dep.hpp:2064): the E3SM-layout
 * calcsize modal_aero_calcsize_sub (modal_aero_calcsize.hpp:748-928) on the
 * state -- dry diameters and its number / mass tendencies, with
 * do_adjust and do_aitacc_transfer on, over the adjustment time dt -- then
 * water uptake (as mam4x_compute_water_uptake) on the dry diameters it
 * returns.  Reads all number and aerosol-mass planes of in->state,
 * temperature, pressure, vapor_mixing_ratio and cloud_fraction.  FP64;
 * stream-ordered; no workspace. */
int mam4x_compute_wet_size(mam4x_handle h, double dt, const mam4x_inputs *in,
                           const mam4x_wetsize_outputs *out, void *stream);

/* ---------------------------------------------------------------------- */
/* E3SM tracer arrays (the host model's wire format, utils.hpp:183-426):   */
/* state_q and qqcw hold MAM4X_PCNST tracers per cell:                      */
/*   0-4 qv qc qi nc ni; 5-8 precipitation; 9-14 gases O3 H2O2 H2SO4 SO2    */
/*   DMS SOAG; 15-39 aerosol: per mode (acc, ait, coarse, pca) its species  */
/*   in mode_aero_species order, then its number.                           */
/* qqcw is indexed like state_q and holds cloud-borne aerosol at 15-39.     */
/* Tracer 9 + i is gas_pcnst tracer i (mam4_amicphys.hpp:35-63).            */
#define MAM4X_PCNST 40
#define MAM4X_PACK_STATE 0    /* extract_stateq_from_prognostics            */
#define MAM4X_PACK_TENDENCY 1 /* extract_ptend_from_tendencies              */

/* One domain's tracer arrays.  The tracers of cell (icol, k) start at
 * state_q + icol * col_stride + k * lev_stride (the same offset in qqcw),
 * tracer index fastest: the reference's per-column (nlev, pcnst) views,
 * stacked.  Strides in elements; 0 = lev_stride MAM4X_PCNST and col_stride
 * nlev * lev_stride.  Rows must not overlap. */
typedef struct mam4x_tracers {
  double *state_q;
  double *qqcw;  /* NULL: cloud-borne fields not transferred */
  long long col_stride, lev_stride;
} mam4x_tracers;

/* The 52 state fields as planes, extents and layout as in mam4x_inputs. */
typedef struct mam4x_fields {
  int ncol, nlev;
  int ldc;
  int layout;
  double *field[MAM4X_NSTATE]; /* NULL: field not transferred */
} mam4x_fields;

/* inject_stateq_to_prognostics and inject_qqcw_to_prognostics
 * (utils.hpp:297-334, 388-420) over a domain: state_q's H2SO4 (11), SOAG
 * (14) and interstitial aerosol (15-39), and qqcw's cloud-borne aerosol
 * (15-39), into the state planes.  Stream-ordered; no workspace. */
int mam4x_tracers_to_fields(mam4x_handle h, const mam4x_tracers *src, const mam4x_fields *dst,
                            void *stream);

/* extract_stateq_from_prognostics and extract_qqcw_from_prognostics
 * (utils.hpp:217-253, 364-386), mode MAM4X_PACK_STATE, or
 * extract_ptend_from_tendencies (utils.hpp:255-295), mode
 * MAM4X_PACK_TENDENCY: the planes into state_q 11, 14 and 15-39 and qqcw
 * 15-39.  The host's other tracers are left as they are; with
 * MAM4X_PACK_TENDENCY the gases the chain does not carry (O3, H2O2, SO2,
 * DMS: 9, 10, 12, 13) receive its zero tendency.  Stream-ordered. */
int mam4x_fields_to_tracers(mam4x_handle h, const mam4x_fields *src, const mam4x_tracers *dst,
                            int mode, void *stream);

/* ---------------------------------------------------------------------- */
/* Aqueous sulfur chemistry: mo_setsox::setsox_single_level               */
/* (mo_setsox.hpp:943-1259) for every cell of a domain -- in-cloud pH,     */
/* S(IV) oxidation by H2O2 and O3, H2SO4 uptake by cloud drops, cloud-     */
/* borne sulfate gain split over the modes by activated number.  It is    */
/* the cloud chemistry whose change of qqcw the microphysics reads as its */
/* pre-cloud-chemistry input.  Tracer tables are indexed by gas_pcnst, in */
/* the units setsox_single_level takes: mol/mol, cloud-borne number       */
/* #/kmol.  Planes: extents and layout as in mam4x_inputs.                 */
typedef struct mam4x_setsox_args {
  int ncol, nlev;
  int ldc;
  int layout;
  const double *pressure;           /* [Pa]                                */
  const double *temperature;        /* [K]                                 */
  const double *liquid_water;       /* cloud liquid water [kg/kg]          */
  const double *cloud_fraction;     /* [-]                                 */
  const double *droplet_number;     /* cloud droplet number [#/kg]         */
  const double *air_number_density; /* xhnm [#/cm3]                        */
  /* cloud-borne tracers: reads 6, 14, 21 (so4_c1-3) and 13, 18, 26
   * (num_c1-3); updates 6, 14, 21.  Other entries are not used (the
   * reference looks up num_c4, 30, but the primary-carbon mode has no
   * sulfate slot, so it never enters the result). */
  double *qcw[MAM4X_GAS_PCNST];
  /* gas-phase tracers: reads 0 (O3); updates 1 (H2O2), 2 (H2SO4), 3 (SO2).
   * Other entries are not used. */
  double *qin[MAM4X_GAS_PCNST];
  /* optional outputs [mol/mol/s] (NULL planes skipped): so4_c tendency from
   * aqueous oxidation and from H2SO4 uptake, non-zero at 6, 14, 21; the
   * reference zeroes the other entries, and so does this call for every
   * plane given. */
  double *dqdt_aqso4[MAM4X_GAS_PCNST];
  double *dqdt_aqh2so4[MAM4X_GAS_PCNST];
} mam4x_setsox_args;

/* setsox_single_level over a domain with time step dt [s]: qcw / qin updated
 * in place.  The reference's column driver mo_setsox::setsox
 * (mo_setsox.hpp:1262-1300) passes per-level copies of qcw / qin and drops
 * them; the single-level routine is what a host calls.  FP64;
 * stream-ordered; no workspace. */
int mam4x_compute_setsox(mam4x_handle h, double dt, const mam4x_setsox_args *args, void *stream);

/* ---------------------------------------------------------------------- */
/* Gas-phase chemistry: mam4::microphysics::gas_phase_chemistry            */
/* (gas_phase_chemistry.hpp:37-116) with the pp_linoz_mam4_resus_mom_soag  */
/* mechanism (gas_chem_mechanism.hpp) and its implicit Euler solver        */
/* (gas_chem.hpp:73-425), for every cell of a domain: rate constants, the  */
/* Newton-Raphson solve of the 30 implicit species with the reference's   */
/* step-cutting.  Its H2SO4 change is the gas production the microphysics */
/* reads as pregaschem.  Planes: extents and layout as in mam4x_inputs.   */
#define MAM4X_EXTCNT 9 /* species with external forcing                 */
#define MAM4X_NFS 8    /* fixed species (invariants)                     */
typedef struct mam4x_gas_chem_args {
  int ncol, nlev;
  int ldc;
  int layout;
  const double *temperature;     /* [K]                                    */
  const double *photolysis_rate; /* H2O2 photolysis, photo_rates[0] [1/s] */
  /* external in-situ forcing [#/cm3/s], the reference's extfrc order
   * (NULL = 0); divided by the air density as the reference does */
  const double *extfrc[MAM4X_EXTCNT];
  /* fixed species [#/cm3]: 0 total density M, 3 H2O, 4 OH, 5 NO3, 6 HO2
   * are read (others unused) */
  const double *invariants[MAM4X_NFS];
  /* washout rates [1/s] by gas_pcnst; entries 1-30 read, NULL = 0 */
  const double *het_rates[MAM4X_GAS_PCNST];
  /* mixing ratios [mol/mol] by gas_pcnst: 1-30 advanced in place (O3, 0,
   * is not in the implicit class and is not used) */
  double *qq[MAM4X_GAS_PCNST];
} mam4x_gas_chem_args;

/* gas_phase_chemistry over a domain with time step dt [s].  FP64;
 * stream-ordered; no workspace. */
int mam4x_compute_gas_chemistry(mam4x_handle h, double dt, const mam4x_gas_chem_args *args,
                                void *stream);

/* ---------------------------------------------------------------------- */
/* Invariant densities: mam4::mo_setinv::setinv (mo_setinv.hpp:45-122) for */
/* every cell -- the fixed species the gas-phase chemistry reads, from T,  */
/* p, water vapour and four offline species.  Planes: extents and layout   */
/* as in mam4x_inputs.                                                     */
typedef struct mam4x_setinv_args {
  int ncol, nlev;
  int ldc;
  int layout;
  const double *temperature;        /* [K]                                */
  const double *pressure;           /* mid-level [Pa]                     */
  const double *vapor_mixing_ratio; /* qv [kg/kg]                         */
  /* offline species [mol/mol], the reference's cnst_offline order: they
   * fill invariants 7, 4 (OH), 5 (NO3), 6 (HO2).  NULL = 0. */
  const double *cnst_offline[4];
  /* [#/cm3]: 0 M, 1 N2, 2 O2, 3 H2O, 4-7 the offline species.  NULL planes
   * are not written. */
  double *invariants[MAM4X_NFS];
} mam4x_setinv_args;

/* setinv over a domain.  FP64; stream-ordered; no workspace. */
int mam4x_compute_invariants(mam4x_handle h, const mam4x_setinv_args *args, void *stream);

/* ---------------------------------------------------------------------- */
/* Linearized stratospheric ozone (LINOZ): per column, the overhead ozone  */
/* column (compute_o3_column_density, compute_o3_column_density.hpp:59-83) */
/* from the ozone before the step, lin_strat_chem_solve at the levels above */
/* the tropopause (lin_strat_chem.hpp:225-261: relaxation to the           */
/* climatological steady state, polar-stratospheric-cloud loss), then      */
/* lin_strat_sfcsink in the lowest o3_lbl levels (:289-316).  Level 0 is   */
/* the model top.  Planes: extents and layout as in mam4x_inputs.          */
#define MAM4X_NLINOZ 8 /* climatology fields, LinozData order              */
typedef struct mam4x_linoz_args {
  int ncol, nlev;
  int ldc;
  int layout;
  const double *temperature; /* [K]                                       */
  const double *pressure;    /* mid-level [Pa]                            */
  const double *pdel;        /* layer thickness [Pa]                      */
  /* o3_clim [mol/mol], t_clim [K], o3col_clim [DU], PmL_clim [mol/mol/s],
   * dPmL_dO3 [1/s], dPmL_dT [mol/mol/s/K], dPmL_dO3col [mol/mol/s/DU],
   * cariolle_pscs [1/s] (mo_gas_phase_chemdr.hpp:35-52) */
  const double *climatology[MAM4X_NLINOZ];
  /* per column [ncol]: solar zenith angle [rad], latitude [rad], number of
   * levels above the tropopause (levels 0 .. ltrop-1 are solved), ozone
   * column above the model top [molecules/cm2] (NULL = 0) */
  const double *zenith_angle;
  const double *latitude;
  const int *ltrop;
  const double *column_above_top;
  /* LinozConf (mo_gas_phase_chemdr.hpp:55-70) */
  double chlorine_loading; /* EESC [ppbv]; PSC loss only when > 0         */
  double psc_T;            /* PSC temperature threshold [K]              */
  double o3_sfc;           /* surface-sink target [mol/mol]              */
  double o3_tau;           /* surface-sink time scale [s]                */
  int o3_lbl;              /* surface-sink levels (0: no sink)           */
  double *o3;              /* ozone [mol/mol], advanced in place          */
  /* optional outputs (NULL: not written): planes of the overhead column
   * [molecules/cm2] (every level), the linoz and PSC tendencies
   * [mol/mol/s] and the steady-state ozone (0 below the tropopause); per
   * column the surface sink's ozone loss [Tg/yr/m2] */
  double *o3_column;
  double *do3_linoz;
  double *do3_linoz_psc;
  double *ss_o3;
  double *o3_sfcsink;
} mam4x_linoz_args;

/* linoz over a domain with time step dt [s].  One lane per column walks
 * its levels.  FP64; stream-ordered; no workspace. */
int mam4x_compute_linoz(mam4x_handle h, double dt, const mam4x_linoz_args *args, void *stream);

/* ---------------------------------------------------------------------- */
/* Washout rates: mam4::mo_sethet::sethet (mo_sethet.hpp:217-585) for every */
/* column -- the het_rates the gas-phase chemistry reads.  H2O2 (1), H2SO4  */
/* (2) and SO2 (3) get rates; every other species 0.  Level 0 is the model  */
/* top.  Planes: extents and layout as in mam4x_inputs.                    */
typedef struct mam4x_sethet_args {
  int ncol, nlev;
  int ldc;
  int layout;
  const double *pressure;    /* mid-level [Pa]                              */
  /* the reference's zmid: sethet passes Atmosphere::height, and
   * sethet_detail's layer depths and fall distances treat it as km */
  const double *zmid;
  const double *temperature; /* [K]                                         */
  const double *cmfdqr;      /* convective precipitation production [kg/kg/s] */
  const double *nrain;       /* stratiform precipitation production [kg/kg/s] */
  const double *nevapr;      /* precipitation evaporation [kg/kg/s]          */
  const double *air_density; /* invariant M [#/cm3]                          */
  const double *h2o2;        /* H2O2 [mol/mol] (the reference also reads SO2
                                only for a washout whose result it drops)  */
  const double *latitude;    /* per column [ncol] [rad]                      */
  const double *phis;        /* per column [ncol] surface geopotential [m2/s2] */
  /* washout rates [1/s] by gas_pcnst: planes 1, 2, 3 required (they hold
   * the column's scratch during the call); other non-NULL planes are
   * zeroed */
  double *het_rates[MAM4X_GAS_PCNST];
} mam4x_sethet_args;

/* sethet over a domain with time step dt [s].  One lane per column.  FP64;
 * stream-ordered; no workspace. */
int mam4x_compute_washout(mam4x_handle h, double dt, const mam4x_sethet_args *args, void *stream);

/* ---------------------------------------------------------------------- */
/* External forcing: mam4::mo_setext::setext (mo_setext.hpp:22-117) for    */
/* every cell -- the extfrc the gas-phase chemistry reads.  Each forced     */
/* species' rate is the sum of its sectors' data, in sector order; data     */
/* stored bottom-up (the reference's file_alt_data) are read flipped.       */
/* Planes: extents and layout as in mam4x_inputs.                          */
#define MAM4X_MAX_SECTORS 4
typedef struct mam4x_setext_args {
  int ncol, nlev;
  int ldc;
  int layout;
  /* per forced species mm (the reference's Forcing): destination entry
   * frc_ndx[mm] (1-based, 1..MAM4X_EXTCNT), sector count (0..4), and
   * whether its data run bottom-up */
  int frc_ndx[MAM4X_EXTCNT];
  int nsectors[MAM4X_EXTCNT];
  int file_alt_data[MAM4X_EXTCNT];
  const double *sectors[MAM4X_EXTCNT][MAM4X_MAX_SECTORS]; /* [#/cm3/s] planes */
  double *extfrc[MAM4X_EXTCNT];                            /* out, [#/cm3/s] */
} mam4x_setext_args;

/* setext over a domain.  FP64; stream-ordered; no workspace. */
int mam4x_compute_forcing(mam4x_handle h, const mam4x_setext_args *args, void *stream);

/* ---------------------------------------------------------------------- */
/* Photolysis rates: mam4::mo_photo::table_photo (mo_photo.hpp:798-893)    */
/* for every column -- the photolysis_rate the gas-phase chemistry reads.  */
/* cloud_mod (:169-427) forms each level's effective albedo and cloud      */
/* multiplier from the cloud optical depths above and below it; jlong     */
/* (:647-796) interpolates the radiative source function (interpolate_rsf, */
/* :498-645) in zenith angle, ozone column, albedo and pressure, and the   */
/* cross sections in temperature and pressure, and sums their product over */
/* wavelength for the mechanism's one photolysis reaction (phtcnt = 1);    */
/* the rate becomes cld_mult esfact (photo + pht_alias_mult j).  Columns   */
/* in darkness (zenith angle >= 88.85 degrees) are left as they are.       */
/* Level 0 is the model top.  Planes: extents and layout as in mam4x_inputs. */
typedef struct mam4x_photo_table {
  /* PhotoTableData (mo_photo.hpp:28-56) dimensions: wavelengths, the cross
   * sections' temperatures (>= 201: the index runs to 200) and pressures,
   * photorates in xsqy, and the source function's pressures, zenith
   * angles, ozone-column ratios and albedos (each >= 2) */
  int nw, nt, np_xs, numj, nump, numsza, numcolo3, numalb;
  /* device arrays, row-major: xsqy [numj][nw][nt][np_xs], rsf_tab [nw][nump]
   * [numsza][numcolo3][numalb]; axes sza [numsza] (deg, increasing), alb
   * [numalb] (increasing), press [nump] (hPa, decreasing), colo3 [nump]
   * (the ozone column at each press), o3rat [numcolo3] (increasing), prs
   * [np_xs] (hPa, decreasing); del_* and dprs: the reciprocals of each
   * axis's spacings (n - 1 each, positive); etfphot [nw] */
  const double *xsqy;
  const double *sza, *del_sza, *alb, *del_alb, *press, *del_p, *colo3, *o3rat, *del_o3rat;
  const double *etfphot, *prs, *dprs;
  const double *rsf_tab;
  double pht_alias_mult; /* pht_alias_mult_1 of the reaction                  */
  int lng_index;         /* its lng_indexer entry: the xsqy photorate
                            (0 .. numj - 1); -1 leaves every rate as it is   */
} mam4x_photo_table;

typedef struct mam4x_photo_args {
  int ncol, nlev;
  int ldc;
  int layout;
  const double *pressure;       /* mid-level [Pa]                           */
  const double *pdel;           /* layer thickness [Pa]                     */
  const double *temperature;    /* [K]                                      */
  const double *o3_column;      /* colo3 [molecules/cm2], the overhead ozone
                                   column (mam4x_compute_linoz's o3_column)  */
  const double *liquid_water;   /* lwc [kg/kg]                              */
  const double *cloud_fraction; /* [fraction]                               */
  const double *zenith_angle;   /* per column [ncol] [rad]                  */
  const double *surface_albedo; /* per column [ncol]                        */
  double esfact;                /* earth-sun distance factor               */
  double *photo;                /* rates [1/s], in and out                  */
  /* device scratch of mam4x_photo_work_len doubles (the reference's
   * PhotoTableWorkArrays): two planes of the domain's layout and ldc, one
   * after the other, for the effective albedo and cloud multiplier */
  double *work;
} mam4x_photo_args;

/* doubles of work a call needs (get_photo_table_work_len,
 * mo_photo.hpp:116-122): 2 x slow x ld, with slow the slower extent of the
 * layout and ld = ldc (0: the fastest extent); -1 for invalid arguments.
 * No device work. */
long long mam4x_photo_work_len(const mam4x_photo_table *table, int ncol, int nlev, int ldc,
                               int layout);

/* table_photo over a domain: a column kernel (one lane per column) forms
 * the albedos and multipliers into work, then a level kernel (one lane per
 * cell) the rates.  FP64; stream-ordered; no allocation. */
int mam4x_compute_photolysis(mam4x_handle h, const mam4x_photo_table *table,
                             const mam4x_photo_args *args, void *stream);

/* Batches: many scenes (domains of any column count, each with its own
 * inputs, outputs and layout) stepped by one set of launches.  The scenes'
 * cells form one work queue, so uneven scenes and uneven per-cell cost share
 * the GPU.  A batch is planned once for fixed pointers and dt; runs only
 * launch kernels -- no allocation, copy, memset or synchronisation -- so a
 * run can be captured in a hipGraph and replayed (with stats disabled).
 *   mam4x_batch_create: validates and uploads the argument blocks and
 *     allocates the batch's own work-queue workspace (allocates and
 *     synchronizes).  A batch never uses the handle's workspace, so a graph
 *     that captured its run stays valid whatever the handle does later
 *     (larger steps, other batches); it is valid until mam4x_batch_destroy.
 *   mam4x_batch_run: one step of every scene; dt must equal the planned dt.
 * A run reads the handle's launch settings, so a handle must outlive its
 * batches.  No counterpart in the reference, which steps one column per Kokkos team. */
typedef struct mam4x_batch_s *mam4x_batch;
int mam4x_batch_create(mam4x_handle h, int nscene, const mam4x_inputs *in,
                       const mam4x_outputs *out, double dt, mam4x_batch *batch);
int mam4x_batch_run(mam4x_batch batch, double t, double dt, void *stream);
int mam4x_batch_destroy(mam4x_batch batch);

/* Cost diagnostic: with stats enabled, every compute_tendencies call adds
 * the SOA sub-steps (gasaerexch_soaexch.hpp:170-335) its cells took to a
 * device counter.  mam4x_last_stats returns the sum since stats were enabled
 * or last read, resets it, and synchronizes the stream; it returns
 * MAM4X_ERR_UNSUPPORTED while stats are disabled. */
int mam4x_enable_stats(mam4x_handle h, int enable);
int mam4x_last_stats(mam4x_handle h, long long *soa_iterations, void *stream);

/* Phase timing, also while stats are enabled: HIP events recorded on the
 * launch stream between the kernels of every compute_tendencies call.
 * mam4x_last_phase_times synchronizes the stream and returns, per phase, the
 * milliseconds summed over the calls since stats were enabled or the times
 * last read (which resets them), and the number of calls in *calls.
 * The fused schedule counts its single kernel as MAM4X_PHASE_FINISH. */
#define MAM4X_NPHASE 3
#define MAM4X_PHASE_SETUP 0  /* per-cell set-up and the work-queue build     */
#define MAM4X_PHASE_SOA 1    /* work-queue kernel of the SOA sub-stepping    */
#define MAM4X_PHASE_FINISH 2 /* per-cell finish: rest of the chain, calcsize */
int mam4x_last_phase_times(mam4x_handle h, double ms[MAM4X_NPHASE], long long *calls,
                           void *stream);

#ifdef __cplusplus
}
#endif
#endif /* MAM4X_H */
