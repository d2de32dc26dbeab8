// CPU build of the HIP kernel's per-cell code (mam4xx_amd/csrc/chain_cell.hpp)
// for the test suite.  It separates logic errors from libm differences: run
// on the CPU with glibc, the kernel's restructured chain must agree with the
// oracle to round-off.  Test infrastructure only; the library never uses it.
#include <cmath>
#include <cstring>

#include "../../mam4xx_amd/csrc/chain_cell.hpp"

using namespace mam4x;

extern "C" long cell_host_run(const int *cfgv, int ncol, int nlev, double dt,
                              const double *state, const double *T, const double *p,
                              const double *z, const double *qv, const double *cld,
                              const double *pblh, const double *dga, const double *dgw,
                              const double *wd, double *tend, double *dgi, double *dgc) {
  static ChainConsts K;
  chain_consts_init(&K);
  const long n = (long)ncol * nlev;
  ChainArgs a;
  std::memset(&a, 0, sizeof(a));
  for (int f = 0; f < NSTATE; ++f) {
    a.state[f] = state + f * n;
    a.tend[f] = tend + f * n;
  }
  a.T = T; a.p = p; a.z = z; a.qv = qv; a.cld = cld; a.pblh = pblh;
  for (int m = 0; m < 4; ++m) {
    a.dga[m] = dga + m * n;
    a.dgw[m] = dgw + m * n;
    a.wd[m] = wd + m * n;
    a.dgi[m] = dgi + m * n;
    a.dgc[m] = dgc + m * n;
  }
  a.K = &K;
  a.ncol = ncol; a.nlev = nlev; a.ldc = ncol; a.layout = 0;
  a.cfg.do_cond = cfgv[0]; a.cfg.do_rename = cfgv[1]; a.cfg.do_newnuc = cfgv[2];
  a.cfg.do_coag = cfgv[3]; a.cfg.gaexch_h2so4_uptake_optaa = cfgv[4];
  a.cfg.nmono = (unsigned)cfgv[5];
  a.do_calcsize = cfgv[6]; a.do_xfer = cfgv[7]; a.do_adjust = cfgv[8];
  a.has_dgn = 1; a.has_sout = 0; a.dt = dt;
  long iters = 0;
  for (int k = 0; k < nlev; ++k)
    for (int i = 0; i < ncol; ++i) iters += dev::chain_cell_fused<false>(a, (long)k * ncol + i, i);
  return iters;
}
