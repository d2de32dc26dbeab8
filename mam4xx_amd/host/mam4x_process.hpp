// C++ host interface of the MI355X microphysics chain, shaped like the
// reference's AeroProcess<Impl> (src/validation/aero_process.hpp:19-129):
// a process object built from a configuration, validate(...) and
// compute_tendencies(t, dt, atm, sfc, progs, diags, tends).  It runs over a
// whole domain per call, on a HIP stream, where the reference runs one
// Kokkos team per column.
//
// The containers mirror the reference's (atmosphere.hpp:16-137,
// aero_config.hpp:74-260) with each ColumnView replaced by a device pointer
// to the field over all columns.  Species are indexed [mode][AeroId] and
// gases [GasId] as in the reference; a (mode, species) pair the mode does not
// carry is ignored.  Everything goes through the C ABI of include/mam4x.h.
//
// Errors: the C ABI returns status codes; this layer turns a failed call into
// a std::runtime_error carrying mam4x_last_error().
#pragma once

#include <stdexcept>
#include <string>

#include "../../include/mam4x.h"

namespace mam4x {
namespace host {

// index enums of aero_modes.hpp:62-68, 126-135, 265-273
enum class ModeIndex { Accumulation = 0, Aitken = 1, Coarse = 2, PrimaryCarbon = 3 };
enum class AeroId { SOA = 0, SO4 = 1, POM = 2, BC = 3, NaCl = 4, DST = 5, MOM = 6 };
enum class GasId { O3 = 0, H2O2 = 1, H2SO4 = 2, SO2 = 3, DMS = 4, SOAG = 5 };
constexpr int num_modes = 4, num_aerosol_ids = 7, num_gas_ids = 6;

// Extent and layout shared by every field of a call (include/mam4x.h).
struct Domain {
  int ncol = 0, nlev = 0;
  int ld = 0;                           // leading dimension, 0 = fastest extent
  int layout = MAM4X_LAYOUT_COL_LEV;    // EAMxx fields: level index fastest
};

struct Atmosphere {  // atmosphere.hpp:16-137 (fields the chain reads)
  const double *temperature = nullptr;         // [K]
  const double *pressure = nullptr;            // [Pa]
  const double *vapor_mixing_ratio = nullptr;  // [kg/kg]
  const double *height = nullptr;              // [m]
  const double *cloud_fraction = nullptr;      // [-]
  const double *planetary_boundary_layer_height = nullptr;  // [m], one per column
  const double *hydrostatic_dp = nullptr;  // [Pa], read only for column budgets
};

struct Surface {};  // surface.hpp: the chain reads nothing from it

struct Prognostics {  // aero_config.hpp:74-168
  double *n_mode_i[num_modes] = {};
  double *n_mode_c[num_modes] = {};
  double *q_aero_i[num_modes][num_aerosol_ids] = {};
  double *q_aero_c[num_modes][num_aerosol_ids] = {};
  double *q_gas[num_gas_ids] = {};
};
using Tendencies = Prognostics;  // aero_config.hpp:18

struct Diagnostics {  // aero_config.hpp:170-260 (fields the chain reads / writes)
  double *dry_geometric_mean_diameter_i[num_modes] = {};  // in: dgncur_a; out: calcsize
  double *dry_geometric_mean_diameter_c[num_modes] = {};  // out: calcsize
  double *wet_geometric_mean_diameter_i[num_modes] = {};  // in: dgncur_awet
  double *wet_density[num_modes] = {};                    // in
};

// The fields of MicrophysDiagnosticArrays (diagnostic_arrays.hpp:5-48) the
// microphysics fills (mam4_amicphys.hpp:2343-2374): per process, a device
// block of MAM4X_GAS_PCNST tracer planes in the reference's gas_pcnst order,
// tracer i at pointer + i * tracer_stride (0 = one plane of the domain).
// Units mol/mol/s (#/kmol/s for numbers).  NULL members are not formed.
struct MicrophysDiagnosticArrays {
  double *gas_aero_exchange_condensation = nullptr;
  double *gas_aero_exchange_renaming = nullptr;
  double *gas_aero_exchange_nucleation = nullptr;
  double *gas_aero_exchange_coagulation = nullptr;
  double *gas_aero_exchange_renaming_cloud_borne = nullptr;
  long long tracer_stride = 0;
};

// species of each mode, in state order (aero_modes.hpp:214-243)
constexpr int kModeSpecies[num_modes][num_aerosol_ids] = {
    {1, 2, 0, 3, 5, 4, 6}, {1, 0, 4, 6, -1, -1, -1}, {5, 4, 1, 3, 2, 0, 6}, {2, 3, 6, -1, -1, -1, -1}};
constexpr int kModeNspec[num_modes] = {7, 4, 7, 3};

class MicrophysicsProcess {
 public:
  using Config = mam4x_config;

  static Config default_config() {
    Config c;
    mam4x_default_config(&c);
    return c;
  }

  explicit MicrophysicsProcess(const Config &config = default_config()) {
    check(mam4x_create(&config, &h_), "mam4x_create");
  }
  ~MicrophysicsProcess() {
    if (h_) mam4x_destroy(h_);
  }
  MicrophysicsProcess(const MicrophysicsProcess &) = delete;
  MicrophysicsProcess &operator=(const MicrophysicsProcess &) = delete;

  std::string name() const { return "MAM4 microphysics chain (MI355X)"; }

  Config process_config() const {
    Config c;
    check(mam4x_get_config(h_, &c), "mam4x_get_config");
    return c;
  }

  // (re-)initializes the process with the given configuration
  // (aero_process.hpp:76-80); the handle keeps its device workspace
  void init(const Config &config) { check(mam4x_init(h_, &config), "mam4x_init"); }

  // true when the inputs are physically consistent (aero_process.hpp:90-95):
  // no negative T, p, qv or prognostic.  Synchronizes the stream.
  bool validate(const Domain &d, const Atmosphere &atm, const Surface &,
                const Prognostics &progs, const Diagnostics &diags, void *stream = nullptr) const {
    mam4x_inputs in = inputs(d, atm, progs, diags);
    long long bad = 0;
    check(mam4x_validate(h_, &in, &bad, stream), "mam4x_validate");
    return bad == 0;
  }

  // aero_process.hpp:113-122 for every column of the domain.  Writes the
  // tendencies and diags.dry_geometric_mean_diameter_{i,c}; enqueued on
  // `stream`, no host synchronisation.
  void compute_tendencies(const Domain &d, double t, double dt, const Atmosphere &atm,
                          const Surface &sfc, const Prognostics &progs, const Diagnostics &diags,
                          const Tendencies &tends, void *stream = nullptr) const {
    compute_tendencies(d, t, dt, atm, sfc, progs, diags, tends, MicrophysDiagnosticArrays{},
                       stream);
  }
  // the same step, also filling the per-process diagnostic arrays (FP64
  // precision only) and, when column_budgets is set, the MAM4X_NBUDGET x
  // ncol column budgets of the tendencies (needs atm.hydrostatic_dp)
  void compute_tendencies(const Domain &d, double t, double dt, const Atmosphere &atm,
                          const Surface &, const Prognostics &progs, const Diagnostics &diags,
                          const Tendencies &tends, const MicrophysDiagnosticArrays &diag_arrays,
                          void *stream = nullptr, double *column_budgets = nullptr) const {
    mam4x_inputs in = inputs(d, atm, progs, diags);
    mam4x_outputs out = {};
    out.column_budgets = column_budgets;
    state_table(tends, out.tendencies);
    for (int m = 0; m < num_modes; ++m) {
      out.dgncur_i[m] = diags.dry_geometric_mean_diameter_i[m];
      out.dgncur_c[m] = diags.dry_geometric_mean_diameter_c[m];
    }
    out.process_tendencies[MAM4X_PROC_CONDENSATION] = diag_arrays.gas_aero_exchange_condensation;
    out.process_tendencies[MAM4X_PROC_RENAMING] = diag_arrays.gas_aero_exchange_renaming;
    out.process_tendencies[MAM4X_PROC_NUCLEATION] = diag_arrays.gas_aero_exchange_nucleation;
    out.process_tendencies[MAM4X_PROC_COAGULATION] = diag_arrays.gas_aero_exchange_coagulation;
    out.process_tendencies[MAM4X_PROC_RENAMING_CLOUD_BORNE] =
        diag_arrays.gas_aero_exchange_renaming_cloud_borne;
    out.process_stride = diag_arrays.tracer_stride;
    check(mam4x_compute_tendencies(h_, t, dt, &in, &out, stream), "mam4x_compute_tendencies");
  }

  mam4x_handle handle() const { return h_; }

 private:
  template <class P>
  static void state_table(const Prognostics &p, P *f) {
    for (int m = 0; m < num_modes; ++m) {
      f[MAM4X_F_NUM_I + m] = p.n_mode_i[m];
      f[MAM4X_F_NUM_C + m] = p.n_mode_c[m];
    }
    int k = 0;
    for (int m = 0; m < num_modes; ++m)
      for (int s = 0; s < kModeNspec[m]; ++s, ++k) {
        f[MAM4X_F_Q_I + k] = p.q_aero_i[m][kModeSpecies[m][s]];
        f[MAM4X_F_Q_C + k] = p.q_aero_c[m][kModeSpecies[m][s]];
      }
    f[MAM4X_F_H2SO4] = p.q_gas[static_cast<int>(GasId::H2SO4)];
    f[MAM4X_F_SOAG] = p.q_gas[static_cast<int>(GasId::SOAG)];
  }

  static mam4x_inputs inputs(const Domain &d, const Atmosphere &atm, const Prognostics &progs,
                             const Diagnostics &diags) {
    mam4x_inputs in = {};
    in.ncol = d.ncol;
    in.nlev = d.nlev;
    in.ldc = d.ld;
    in.layout = d.layout;
    state_table(progs, in.state);
    in.temperature = atm.temperature;
    in.pressure = atm.pressure;
    in.height = atm.height;
    in.vapor_mixing_ratio = atm.vapor_mixing_ratio;
    in.cloud_fraction = atm.cloud_fraction;
    in.pblh = atm.planetary_boundary_layer_height;
    in.hydrostatic_dp = atm.hydrostatic_dp;
    for (int m = 0; m < num_modes; ++m) {
      in.dgncur_a[m] = diags.dry_geometric_mean_diameter_i[m];
      in.dgncur_awet[m] = diags.wet_geometric_mean_diameter_i[m];
      in.wetdens[m] = diags.wet_density[m];
    }
    return in;
  }

  void check(int rc, const char *what) const {
    if (rc != MAM4X_OK)
      throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                               "): " + (h_ ? mam4x_last_error(h_) : "no handle"));
  }

  mam4x_handle h_ = nullptr;
};

}  // namespace host
}  // namespace mam4x
