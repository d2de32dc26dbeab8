// Test driver of the C++ host interface (mam4xx_amd/host/mam4x_process.hpp),
// loaded by tests/test_gpu_adapter.py.  It takes a scene in host memory,
// [field][nlev][ncol], stores every field in its own device allocation in the
// EAMxx layout (ncol x nlev, level fastest), fills the AeroProcess-shaped
// containers species by species, calls validate and compute_tendencies
// through MicrophysicsProcess, and returns tendencies and diameters in the
// scene's layout; with proc set, also the per-process diagnostic arrays
// ([5][31][nlev][ncol]), each block one allocation of 31 EAMxx-layout planes.
// Test infrastructure only.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../mam4xx_amd/host/mam4x_process.hpp"

using namespace mam4x::host;

namespace {

struct DevField {
  double *p = nullptr;
  DevField() = default;
  DevField(const DevField &) = delete;
  ~DevField() {
    if (p) (void)hipFree(p);
  }
};

// [nlev][ncol] host plane -> [ncol][nlev] device field
bool upload(DevField &f, const double *plane, int ncol, int nlev) {
  std::vector<double> t((size_t)ncol * nlev);
  for (int k = 0; k < nlev; ++k)
    for (int i = 0; i < ncol; ++i) t[(size_t)i * nlev + k] = plane[(size_t)k * ncol + i];
  if (hipMalloc(&f.p, t.size() * sizeof(double)) != hipSuccess) return false;
  return hipMemcpy(f.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice) ==
         hipSuccess;
}

// [ncol][nlev] device field -> [nlev][ncol] host plane
bool download(double *plane, const double *dev, int ncol, int nlev) {
  std::vector<double> t((size_t)ncol * nlev);
  if (hipMemcpy(t.data(), dev, t.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return false;
  for (int k = 0; k < nlev; ++k)
    for (int i = 0; i < ncol; ++i) plane[(size_t)k * ncol + i] = t[(size_t)i * nlev + k];
  return true;
}

}  // namespace

// Returns 0 on success, 1 on a HIP failure, 2 when validate reports invalid
// inputs, 3 on an exception (message in err).  Optional: hdp (layer
// thicknesses, [nlev][ncol]) with budgets (MAM4X_NBUDGET x ncol out), and
// do_coag >= 0 to re-initialize the process with that switch
// (MicrophysicsProcess::init) before the step.
extern "C" int adapter_run(int ncol, int nlev, double dt, const double *state, const double *T,
                           const double *p, const double *z, const double *qv,
                           const double *cld, const double *pblh, const double *dga,
                           const double *dgw, const double *wd, double *tend, double *dgi,
                           double *dgc, double *proc, char *err, int errlen, const double *hdp,
                           double *budgets, int do_coag) {
  const size_t n = (size_t)ncol * nlev;
  try {
    DevField st[MAM4X_NSTATE], td[MAM4X_NSTATE], met[5], dga_d[4], dgw_d[4], wd_d[4], dgc_d[4];
    DevField pbl;
    for (int f = 0; f < MAM4X_NSTATE; ++f) {
      if (!upload(st[f], state + f * n, ncol, nlev)) return 1;
      if (hipMalloc(&td[f].p, n * sizeof(double)) != hipSuccess) return 1;
    }
    const double *mp[5] = {T, p, z, qv, cld};
    for (int i = 0; i < 5; ++i)
      if (!upload(met[i], mp[i], ncol, nlev)) return 1;
    if (hipMalloc(&pbl.p, ncol * sizeof(double)) != hipSuccess ||
        hipMemcpy(pbl.p, pblh, ncol * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return 1;
    for (int m = 0; m < 4; ++m) {
      if (!upload(dga_d[m], dga + m * n, ncol, nlev) || !upload(dgw_d[m], dgw + m * n, ncol, nlev) ||
          !upload(wd_d[m], wd + m * n, ncol, nlev) ||
          hipMalloc(&dgc_d[m].p, n * sizeof(double)) != hipSuccess)
        return 1;
    }
    Domain d;
    d.ncol = ncol;
    d.nlev = nlev;
    Atmosphere atm;
    atm.temperature = met[0].p;
    atm.pressure = met[1].p;
    atm.height = met[2].p;
    atm.vapor_mixing_ratio = met[3].p;
    atm.cloud_fraction = met[4].p;
    atm.planetary_boundary_layer_height = pbl.p;
    Prognostics progs;
    Tendencies tends;
    for (int m = 0; m < 4; ++m) {
      progs.n_mode_i[m] = st[MAM4X_F_NUM_I + m].p;
      progs.n_mode_c[m] = st[MAM4X_F_NUM_C + m].p;
      tends.n_mode_i[m] = td[MAM4X_F_NUM_I + m].p;
      tends.n_mode_c[m] = td[MAM4X_F_NUM_C + m].p;
    }
    for (int m = 0, k = 0; m < 4; ++m)
      for (int s = 0; s < kModeNspec[m]; ++s, ++k) {
        const int a = kModeSpecies[m][s];
        progs.q_aero_i[m][a] = st[MAM4X_F_Q_I + k].p;
        progs.q_aero_c[m][a] = st[MAM4X_F_Q_C + k].p;
        tends.q_aero_i[m][a] = td[MAM4X_F_Q_I + k].p;
        tends.q_aero_c[m][a] = td[MAM4X_F_Q_C + k].p;
      }
    progs.q_gas[(int)GasId::H2SO4] = st[MAM4X_F_H2SO4].p;
    progs.q_gas[(int)GasId::SOAG] = st[MAM4X_F_SOAG].p;
    tends.q_gas[(int)GasId::H2SO4] = td[MAM4X_F_H2SO4].p;
    tends.q_gas[(int)GasId::SOAG] = td[MAM4X_F_SOAG].p;
    Diagnostics diags;
    for (int m = 0; m < 4; ++m) {
      // calcsize overwrites the dry diameter it reads, as in the reference
      diags.dry_geometric_mean_diameter_i[m] = dga_d[m].p;
      diags.dry_geometric_mean_diameter_c[m] = dgc_d[m].p;
      diags.wet_geometric_mean_diameter_i[m] = dgw_d[m].p;
      diags.wet_density[m] = wd_d[m].p;
    }
    DevField blk[5];
    MicrophysDiagnosticArrays da;
    if (proc) {
      for (int k = 0; k < 5; ++k)
        if (hipMalloc(&blk[k].p, 31 * n * sizeof(double)) != hipSuccess) return 1;
      da.gas_aero_exchange_condensation = blk[0].p;
      da.gas_aero_exchange_renaming = blk[1].p;
      da.gas_aero_exchange_nucleation = blk[2].p;
      da.gas_aero_exchange_coagulation = blk[3].p;
      da.gas_aero_exchange_renaming_cloud_borne = blk[4].p;
    }
    DevField dp_d, bud_d;
    if (hdp) {
      if (!upload(dp_d, hdp, ncol, nlev)) return 1;
      atm.hydrostatic_dp = dp_d.p;
    }
    if (budgets && hipMalloc(&bud_d.p, (size_t)MAM4X_NBUDGET * ncol * sizeof(double)) != hipSuccess)
      return 1;
    MicrophysicsProcess process;
    if (do_coag >= 0) {
      auto cfg = process.process_config();
      cfg.do_coag = do_coag;
      process.init(cfg);
    }
    if (!process.validate(d, atm, Surface{}, progs, diags)) return 2;
    if (proc || budgets)
      process.compute_tendencies(d, 0.0, dt, atm, Surface{}, progs, diags, tends, da, nullptr,
                                 bud_d.p);
    else
      process.compute_tendencies(d, 0.0, dt, atm, Surface{}, progs, diags, tends);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    if (budgets && hipMemcpy(budgets, bud_d.p, (size_t)MAM4X_NBUDGET * ncol * sizeof(double),
                             hipMemcpyDeviceToHost) != hipSuccess)
      return 1;
    if (proc)
      for (int k = 0; k < 5; ++k)
        for (int i = 0; i < 31; ++i)
          if (!download(proc + (k * 31 + i) * n, blk[k].p + i * n, ncol, nlev)) return 1;
    for (int f = 0; f < MAM4X_NSTATE; ++f)
      if (!download(tend + f * n, td[f].p, ncol, nlev)) return 1;
    for (int m = 0; m < 4; ++m)
      if (!download(dgi + m * n, dga_d[m].p, ncol, nlev) || !download(dgc + m * n, dgc_d[m].p, ncol, nlev))
        return 1;
    return 0;
  } catch (const std::exception &e) {
    std::strncpy(err, e.what(), errlen - 1);
    err[errlen - 1] = 0;
    return 3;
  }
}
