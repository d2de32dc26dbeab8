// MI355X (gfx950) implementation of the MAM4 microphysics chain and its C ABI
// (include/mam4x.h).
//
// Launch shape: one lane per (column, level) cell, cells enumerated level-
// major / column-minor so consecutive lanes read consecutive addresses of each
// [level][column] plane (fully coalesced 64-lane loads).  The chain is VALU
// bound (no dense contraction anywhere), so there is no MFMA or LDS tiling;
// the sigma-only constants come from a small device buffer through the
// scalar cache.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "../../include/mam4x.h"
#include "chain_cell.hpp"

using namespace mam4x;

namespace {

constexpr int kBlock = 256;

template <bool PRE>
__global__ void __launch_bounds__(kBlock) chain_fp64_kernel(ChainArgs a) {
  const long ncell = (long)a.ncol * a.nlev;
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  unsigned long long niter = 0;
  if (c < ncell) {
    const int k = (int)(c / a.ncol);
    const int icol = (int)(c - (long)k * a.ncol);
    const long off = (long)k * a.ldc + icol;
    niter = (unsigned long long)dev::chain_cell<PRE>(a, off, icol);
  }
  if (a.stats) {  // uniform branch: wave-reduce, one atomic per wave
    for (int d = 32; d > 0; d >>= 1) niter += __shfl_xor(niter, d, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(a.stats, niter);
  }
}

// AeroProcess::validate checks (gasaerexch.hpp:96-120): negative T, p, qv or
// any negative prognostic.  One count per offending cell.
__global__ void __launch_bounds__(kBlock) validate_kernel(ChainArgs a,
                                                          unsigned long long *count) {
  const long ncell = (long)a.ncol * a.nlev;
  const long c = (long)blockIdx.x * kBlock + threadIdx.x;
  unsigned long long bad = 0;
  if (c < ncell) {
    const int k = (int)(c / a.ncol);
    const int icol = (int)(c - (long)k * a.ncol);
    const long off = (long)k * a.ldc + icol;
    bool v = (a.T[off] < 0) || (a.p[off] < 0) || (a.qv[off] < 0);
    for (int f = 0; f < NSTATE; ++f) v = v || (a.state[f][off] < 0);
    bad = v ? 1ull : 0ull;
  }
  for (int d = 32; d > 0; d >>= 1) bad += __shfl_xor(bad, d, 64);
  if ((threadIdx.x & 63) == 0 && bad) atomicAdd(count, bad);
}

}  // namespace

struct mam4x_handle_s {
  mam4x_config cfg;
  int device = 0;
  ChainConsts *d_consts = nullptr;
  unsigned long long *d_count = nullptr;  // validate counter
  unsigned long long *d_stats = nullptr;  // SOA sub-steps since the last read
  bool stats = false;
  std::string err;
};

static int fail(mam4x_handle h, int code, const std::string &msg) {
  if (h) h->err = msg;
  return code;
}

#define HIP_TRY(h, expr)                                                     \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail((h), MAM4X_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

static int check_inputs(mam4x_handle h, const mam4x_inputs *in) {
  if (!in) return fail(h, MAM4X_ERR_ARG, "inputs is NULL");
  if (in->ncol < 0 || in->nlev < 0) return fail(h, MAM4X_ERR_ARG, "negative ncol or nlev");
  const int ldc = in->ldc ? in->ldc : in->ncol;
  if (ldc < in->ncol) return fail(h, MAM4X_ERR_ARG, "ldc < ncol");
  if ((long long)ldc * in->nlev > 0x7fffffffLL * 8)
    return fail(h, MAM4X_ERR_ARG, "domain too large");
  for (int f = 0; f < MAM4X_NSTATE; ++f)
    if (!in->state[f]) return fail(h, MAM4X_ERR_ARG, "state plane " + std::to_string(f) + " is NULL");
  if (!in->temperature || !in->pressure || !in->height || !in->vapor_mixing_ratio ||
      !in->cloud_fraction || !in->pblh)
    return fail(h, MAM4X_ERR_ARG, "an atmosphere plane is NULL");
  for (int m = 0; m < 4; ++m)
    if (!in->dgncur_a[m] || !in->dgncur_awet[m] || !in->wetdens[m])
      return fail(h, MAM4X_ERR_ARG, "a modal diameter/density plane is NULL");
  return MAM4X_OK;
}

static void fill_args(mam4x_handle h, const mam4x_inputs *in, ChainArgs &a) {
  std::memset(&a, 0, sizeof(a));
  for (int f = 0; f < NSTATE; ++f) {
    a.state[f] = in->state[f];
    a.precld[f] = in->precldchem[f];
  }
  a.T = in->temperature;
  a.p = in->pressure;
  a.z = in->height;
  a.qv = in->vapor_mixing_ratio;
  a.cld = in->cloud_fraction;
  a.pblh = in->pblh;
  for (int m = 0; m < 4; ++m) {
    a.dga[m] = in->dgncur_a[m];
    a.dgw[m] = in->dgncur_awet[m];
    a.wd[m] = in->wetdens[m];
  }
  a.pre_h2so4 = in->pregaschem_h2so4;
  a.pre_soag = in->pregaschem_soag;
  a.K = h->d_consts;
  a.ncol = in->ncol;
  a.nlev = in->nlev;
  a.ldc = in->ldc ? in->ldc : in->ncol;
  const mam4x_config &c = h->cfg;
  a.cfg.do_cond = c.do_cond;
  a.cfg.do_rename = c.do_rename;
  a.cfg.do_newnuc = c.do_newnuc;
  a.cfg.do_coag = c.do_coag;
  // QUIRK (mam4_amicphys.hpp:1129 vs :1914): the two optaa switches arrive
  // swapped in mam_amicphys_1subarea, so the H2SO4 production treatment
  // follows newnuc_h2so4_conc_optaa.
  a.cfg.gaexch_h2so4_uptake_optaa = c.newnuc_h2so4_conc_optaa;
  a.cfg.nmono = c.n_so4_monolayers_pcage;
  a.do_calcsize = c.do_calcsize;
  a.do_xfer = c.calcsize_do_aitacc_transfer;
  a.do_adjust = c.calcsize_do_adjust;
}

extern "C" {

int mam4x_abi_version(void) { return MAM4X_ABI_VERSION; }

int mam4x_default_config(mam4x_config *out) {
  if (!out) return MAM4X_ERR_ARG;
  out->do_cond = 1;
  out->do_rename = 1;
  out->do_newnuc = 1;
  out->do_coag = 1;
  out->gaexch_h2so4_uptake_optaa = 2;
  out->newnuc_h2so4_conc_optaa = 2;
  out->n_so4_monolayers_pcage = 8;
  out->do_calcsize = 1;
  out->calcsize_do_aitacc_transfer = 1;
  out->calcsize_do_adjust = 1;
  out->precision = MAM4X_PREC_FP64;
  return MAM4X_OK;
}

int mam4x_create(const mam4x_config *config, mam4x_handle *out) {
  if (!out) return MAM4X_ERR_ARG;
  *out = nullptr;
  mam4x_handle h = new (std::nothrow) mam4x_handle_s();
  if (!h) return MAM4X_ERR_NOMEM;
  if (config)
    h->cfg = *config;
  else
    mam4x_default_config(&h->cfg);
  if (h->cfg.precision != MAM4X_PREC_FP64 && h->cfg.precision != MAM4X_PREC_MIXED) {
    delete h;
    return MAM4X_ERR_ARG;
  }
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return MAM4X_ERR_HIP;
  }
  ChainConsts k;
  chain_consts_init(&k);
  if (hipMalloc(&h->d_consts, sizeof(ChainConsts)) != hipSuccess ||
      hipMalloc(&h->d_count, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&h->d_stats, sizeof(unsigned long long)) != hipSuccess) {
    if (h->d_consts) (void)hipFree(h->d_consts);
    if (h->d_count) (void)hipFree(h->d_count);
    delete h;
    return MAM4X_ERR_NOMEM;
  }
  if (hipMemcpy(h->d_consts, &k, sizeof(k), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(h->d_stats, 0, sizeof(unsigned long long)) != hipSuccess) {
    (void)hipFree(h->d_consts);
    (void)hipFree(h->d_count);
    (void)hipFree(h->d_stats);
    delete h;
    return MAM4X_ERR_HIP;
  }
  *out = h;
  return MAM4X_OK;
}

int mam4x_destroy(mam4x_handle h) {
  if (!h) return MAM4X_ERR_ARG;
  if (h->d_consts) (void)hipFree(h->d_consts);
  if (h->d_count) (void)hipFree(h->d_count);
  if (h->d_stats) (void)hipFree(h->d_stats);
  delete h;
  return MAM4X_OK;
}

int mam4x_get_config(mam4x_handle h, mam4x_config *out) {
  if (!h || !out) return MAM4X_ERR_ARG;
  *out = h->cfg;
  return MAM4X_OK;
}

int mam4x_set_precision(mam4x_handle h, int precision) {
  if (!h) return MAM4X_ERR_ARG;
  if (precision != MAM4X_PREC_FP64 && precision != MAM4X_PREC_MIXED)
    return fail(h, MAM4X_ERR_ARG, "unknown precision");
  h->cfg.precision = precision;
  return MAM4X_OK;
}

const char *mam4x_last_error(mam4x_handle h) { return h ? h->err.c_str() : "NULL handle"; }

int mam4x_enable_stats(mam4x_handle h, int enable) {
  if (!h) return MAM4X_ERR_ARG;
  if (enable && !h->stats)
    HIP_TRY(h, hipMemset(h->d_stats, 0, sizeof(unsigned long long)));
  h->stats = enable != 0;
  return MAM4X_OK;
}

int mam4x_last_stats(mam4x_handle h, long long *soa_iterations, void *stream) {
  if (!h || !soa_iterations) return MAM4X_ERR_ARG;
  if (!h->stats) return fail(h, MAM4X_ERR_UNSUPPORTED, "stats not enabled");
  unsigned long long v = 0;
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(h, hipMemcpyAsync(&v, h->d_stats, sizeof(v), hipMemcpyDeviceToHost, s));
  HIP_TRY(h, hipMemsetAsync(h->d_stats, 0, sizeof(v), s));
  HIP_TRY(h, hipStreamSynchronize(s));
  *soa_iterations = (long long)v;
  return MAM4X_OK;
}

int mam4x_validate(mam4x_handle h, const mam4x_inputs *in, long long *violations,
                   void *stream) {
  if (!h || !violations) return MAM4X_ERR_ARG;
  int rc = check_inputs(h, in);
  if (rc) return rc;
  ChainArgs a;
  fill_args(h, in, a);
  const long ncell = (long)in->ncol * in->nlev;
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(h, hipMemsetAsync(h->d_count, 0, sizeof(unsigned long long), s));
  if (ncell > 0) {
    const unsigned grid = (unsigned)((ncell + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(validate_kernel, dim3(grid), dim3(kBlock), 0, s, a, h->d_count);
    HIP_TRY(h, hipGetLastError());
  }
  unsigned long long v = 0;
  HIP_TRY(h, hipMemcpyAsync(&v, h->d_count, sizeof(v), hipMemcpyDeviceToHost, s));
  HIP_TRY(h, hipStreamSynchronize(s));
  *violations = (long long)v;
  return MAM4X_OK;
}

int mam4x_compute_tendencies(mam4x_handle h, double t, double dt, const mam4x_inputs *in,
                             const mam4x_outputs *out, void *stream) {
  (void)t;
  if (!h) return MAM4X_ERR_ARG;
  int rc = check_inputs(h, in);
  if (rc) return rc;
  if (!out) return fail(h, MAM4X_ERR_ARG, "outputs is NULL");
  if (!(dt > 0.0)) return fail(h, MAM4X_ERR_ARG, "dt must be > 0");
  if (h->cfg.precision != MAM4X_PREC_FP64)
    return fail(h, MAM4X_ERR_UNSUPPORTED, "mixed precision not built yet");
  ChainArgs a;
  fill_args(h, in, a);
  for (int f = 0; f < NSTATE; ++f) {
    if (!out->tendencies[f])
      return fail(h, MAM4X_ERR_ARG, "tendency plane " + std::to_string(f) + " is NULL");
    a.tend[f] = out->tendencies[f];
    a.sout[f] = out->state_out[f];
  }
  int nsout = 0;
  for (int f = 0; f < NSTATE; ++f) nsout += out->state_out[f] != nullptr;
  if (nsout != 0 && nsout != NSTATE)
    return fail(h, MAM4X_ERR_ARG, "state_out must be all NULL or all set");
  a.has_sout = nsout == NSTATE;
  a.has_dgn = 0;
  for (int m = 0; m < 4; ++m) {
    a.dgi[m] = out->dgncur_i[m];
    a.dgc[m] = out->dgncur_c[m];
    a.has_dgn |= (a.dgi[m] != nullptr) || (a.dgc[m] != nullptr);
  }
  a.dt = dt;
  hipStream_t s = (hipStream_t)stream;
  if (h->stats) a.stats = h->d_stats;
  bool pre = in->pregaschem_h2so4 || in->pregaschem_soag;
  for (int f = 0; f < NSTATE; ++f) pre = pre || in->precldchem[f];
  const long ncell = (long)in->ncol * in->nlev;
  if (ncell == 0) return MAM4X_OK;
  const unsigned grid = (unsigned)((ncell + kBlock - 1) / kBlock);
  if (pre)
    hipLaunchKernelGGL(chain_fp64_kernel<true>, dim3(grid), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(chain_fp64_kernel<false>, dim3(grid), dim3(kBlock), 0, s, a);
  HIP_TRY(h, hipGetLastError());
  return MAM4X_OK;
}

}  // extern "C"
