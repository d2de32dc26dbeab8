// CPU build of the HIP kernel's per-cell code (mam4xx_amd/csrc/chain_cell.hpp)
// for the test suite, in the FP64 (prec 0), mixed (prec 1) or tangent
// (prec 2) policy, and of its reverse sweep (chain_adjoint.hpp).  It
// separates logic errors from libm differences: run on the CPU with glibc,
// the kernel's restructured FP64 chain must agree with the oracle to
// round-off.  Test infrastructure only; the library never uses
// it.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../mam4xx_amd/csrc/chain_adjoint.hpp"

using namespace mam4x;

// condensation sub-steps of the next calls (the config vectors predate it)
static int g_ntsubstep = 1;
extern "C" void cell_host_set_ntsubstep(int n) { g_ntsubstep = n; }

// The next cell_host_run* calls' SOA-cap flag plane ([nlev][ncol] ints; NULL
// = not formed).
static int *g_soacap = nullptr;
extern "C" void cell_host_set_soacap(int *p) { g_soacap = p; }

// The reverse sweep's SOA states: 0 every state in a vector, 1 the GPU's
// fixed two-level storage (dev::rev::SoaTape2<32>).
static int g_tape = 0;
extern "C" void cell_host_set_tape(int k) { g_tape = k; }

struct VecSoaTape {
  std::vector<dev::rev::SoaCkpt> v;
  int record(dev::SoaState<double> &st, double dtfull) {
    v.clear();
    while (true) {
      const dev::rev::SoaCkpt c = dev::rev::soa_ckpt(st);
      if (!dev::soa_tick<true>(st, dtfull)) break;
      v.push_back(c);
    }
    return (int)v.size();
  }
  dev::rev::SoaCkpt get(int i) const { return v[i]; }
};

struct VecSubTape {
  std::vector<dev::Aer<double>> S;
  std::vector<double> g;
  void save(int j, const dev::Aer<double> &S_, const double g_[2]) {
    if ((int)S.size() <= j) {
      S.resize(j + 1);
      g.resize(2 * (j + 1));
    }
    S[j] = S_;
    g[2 * j] = g_[0];
    g[2 * j + 1] = g_[1];
  }
  void load(int j, dev::Aer<double> &S_, double g_[2]) const {
    S_ = S[j];
    g_[0] = g[2 * j];
    g_[1] = g[2 * j + 1];
  }
};

// Per-process tendency diagnostics (the Fp64Diag policy): proc receives
// [5][31][nlev][ncol]; precld (NULL = none) is the optional pre-cloud-
// chemistry state [52][nlev][ncol].
extern "C" long cell_host_proc(const int *cfgv, int ncol, int nlev, double dt,
                               const double *state, const double *T, const double *p,
                               const double *z, const double *qv, const double *cld,
                               const double *pblh, const double *dga, const double *dgw,
                               const double *wd, const double *precld, double *tend,
                               double *proc) {
  static ChainConsts K;
  chain_consts_init(&K);
  const long n = (long)ncol * nlev;
  ChainArgs a;
  std::memset(&a, 0, sizeof(a));
  for (int f = 0; f < NSTATE; ++f) {
    a.state[f] = state + f * n;
    a.tend[f] = tend + f * n;
    if (precld) a.precld[f] = precld + f * n;
  }
  for (int k = 0; k < 5; ++k) a.proc[k] = proc + (long)k * 31 * n;
  a.pstride = n;
  a.T = T; a.p = p; a.z = z; a.qv = qv; a.cld = cld; a.pblh = pblh;
  for (int m = 0; m < 4; ++m) {
    a.dga[m] = dga + m * n;
    a.dgw[m] = dgw + m * n;
    a.wd[m] = wd + m * n;
  }
  a.K = &K;
  a.ncol = ncol; a.nlev = nlev; a.ldc = ncol; a.layout = 0;
  a.cfg.do_cond = cfgv[0]; a.cfg.do_rename = cfgv[1]; a.cfg.do_newnuc = cfgv[2];
  a.cfg.do_coag = cfgv[3]; a.cfg.gaexch_h2so4_uptake_optaa = cfgv[4];
  a.cfg.nmono = (unsigned)cfgv[5];
  a.cfg.ntsubstep = g_ntsubstep;
  a.do_calcsize = cfgv[6]; a.do_xfer = cfgv[7]; a.do_adjust = cfgv[8];
  a.dt = dt;
  long iters = 0;
  for (int k = 0; k < nlev; ++k)
    for (int i = 0; i < ncol; ++i) {
      const long off = (long)k * ncol + i;
      iters += precld ? dev::chain_cell_step<true, Fp64Diag>(a, off, i)
                      : dev::chain_cell_step<false, Fp64Diag>(a, off, i);
    }
  return iters;
}

extern "C" long cell_host_run3(const int *cfgv, int prec, int ncol, int nlev, double dt,
                               const double *state, const double *T, const double *p,
                               const double *z, const double *qv, const double *cld,
                               const double *pblh, const double *dga, const double *dgw,
                               const double *wd, double *tend, double *dgi, double *dgc,
                               const double *dstate, const double *dT, const double *dp,
                               const double *dqv, double *dtend, const double *tbar,
                               double *xbar) {
  static ChainConsts K;
  chain_consts_init(&K);
  const long n = (long)ncol * nlev;
  ChainArgs a;
  std::memset(&a, 0, sizeof(a));
  for (int f = 0; f < NSTATE; ++f) {
    a.state[f] = state + f * n;
    if (tend) a.tend[f] = tend + f * n;
    if (dstate) a.dstate[f] = dstate + f * n;
    if (dtend) a.dtend[f] = dtend + f * n;
    if (tbar) a.tbar[f] = tbar + f * n;
  }
  a.xbar = xbar;
  a.T = T; a.p = p; a.z = z; a.qv = qv; a.cld = cld; a.pblh = pblh;
  a.dT = dT; a.dp = dp; a.dqv = dqv;
  for (int m = 0; m < 4; ++m) {
    a.dga[m] = dga + m * n;
    a.dgw[m] = dgw + m * n;
    a.wd[m] = wd + m * n;
    if (dgi) a.dgi[m] = dgi + m * n;
    if (dgc) a.dgc[m] = dgc + m * n;
  }
  a.K = &K;
  a.ncol = ncol; a.nlev = nlev; a.ldc = ncol; a.layout = 0;
  a.cfg.do_cond = cfgv[0]; a.cfg.do_rename = cfgv[1]; a.cfg.do_newnuc = cfgv[2];
  a.cfg.do_coag = cfgv[3]; a.cfg.gaexch_h2so4_uptake_optaa = cfgv[4];
  a.cfg.nmono = (unsigned)cfgv[5];
  a.cfg.ntsubstep = g_ntsubstep;
  a.do_calcsize = cfgv[6]; a.do_xfer = cfgv[7]; a.do_adjust = cfgv[8];
  a.has_dgn = dgi || dgc; a.has_sout = 0; a.dt = dt;
  a.soacap = g_soacap;
  long iters = 0;
  for (int k = 0; k < nlev; ++k)
    for (int i = 0; i < ncol; ++i) {
      const long off = (long)k * ncol + i;
      iters += prec == 2   ? dev::chain_cell_step<false, Tangent>(a, off, i)
               : prec == 1 ? dev::chain_cell_step<false, Mixed>(a, off, i)
                           : dev::chain_cell_step<false, Fp64>(a, off, i);
    }
  return iters;
}

// Tangent mode: dstate ([52][nlev][ncol]), dT, dp, dqv are the input
// direction (NULL = 0); dtend receives the directional derivative of the
// tendencies.  With xbar set, the lane writes sum_f tbar[f] * d(tendency f)
// to xbar instead (tbar: [52][nlev][ncol], NULL = 0).
extern "C" long cell_host_run2(const int *cfgv, int prec, int ncol, int nlev, double dt,
                               const double *state, const double *T, const double *p,
                               const double *z, const double *qv, const double *cld,
                               const double *pblh, const double *dga, const double *dgw,
                               const double *wd, double *tend, double *dgi, double *dgc,
                               const double *dstate, const double *dT, const double *dp,
                               const double *dqv, double *dtend) {
  return cell_host_run3(cfgv, prec, ncol, nlev, dt, state, T, p, z, qv, cld, pblh, dga, dgw, wd,
                        tend, dgi, dgc, dstate, dT, dp, dqv, dtend, nullptr, nullptr);
}

extern "C" long cell_host_run(const int *cfgv, int prec, int ncol, int nlev, double dt,
                              const double *state, const double *T, const double *p,
                              const double *z, const double *qv, const double *cld,
                              const double *pblh, const double *dga, const double *dgw,
                              const double *wd, double *tend, double *dgi, double *dgc) {
  return cell_host_run2(cfgv, prec, ncol, nlev, dt, state, T, p, z, qv, cld, pblh, dga, dgw, wd,
                        tend, dgi, dgc, nullptr, nullptr, nullptr, nullptr, nullptr);
}

// Tangent with an optional pre-cloud-chemistry state (precld, NULL = none).
extern "C" long cell_host_tangent(const int *cfgv, int ncol, int nlev, double dt,
                                  const double *state, const double *T, const double *p,
                                  const double *z, const double *qv, const double *cld,
                                  const double *pblh, const double *dga, const double *dgw,
                                  const double *wd, const double *precld, const double *dstate,
                                  const double *dT, const double *dp, const double *dqv,
                                  double *tend, double *dtend) {
  static ChainConsts K;
  chain_consts_init(&K);
  const long n = (long)ncol * nlev;
  ChainArgs a;
  std::memset(&a, 0, sizeof(a));
  for (int f = 0; f < NSTATE; ++f) {
    a.state[f] = state + f * n;
    a.tend[f] = tend + f * n;
    if (dstate) a.dstate[f] = dstate + f * n;
    a.dtend[f] = dtend + f * n;
    if (precld) a.precld[f] = precld + f * n;
  }
  a.T = T; a.p = p; a.z = z; a.qv = qv; a.cld = cld; a.pblh = pblh;
  a.dT = dT; a.dp = dp; a.dqv = dqv;
  for (int m = 0; m < 4; ++m) {
    a.dga[m] = dga + m * n;
    a.dgw[m] = dgw + m * n;
    a.wd[m] = wd + m * n;
  }
  a.K = &K;
  a.ncol = ncol; a.nlev = nlev; a.ldc = ncol; a.layout = 0;
  a.cfg.do_cond = cfgv[0]; a.cfg.do_rename = cfgv[1]; a.cfg.do_newnuc = cfgv[2];
  a.cfg.do_coag = cfgv[3]; a.cfg.gaexch_h2so4_uptake_optaa = cfgv[4];
  a.cfg.nmono = (unsigned)cfgv[5];
  a.cfg.ntsubstep = g_ntsubstep;
  a.do_calcsize = cfgv[6]; a.do_xfer = cfgv[7]; a.do_adjust = cfgv[8];
  a.dt = dt;
  long iters = 0;
  for (int k = 0; k < nlev; ++k)
    for (int i = 0; i < ncol; ++i) {
      const long off = (long)k * ncol + i;
      iters += precld ? dev::chain_cell_step<true, Tangent>(a, off, i)
                      : dev::chain_cell_step<false, Tangent>(a, off, i);
    }
  return iters;
}

// Reverse sweep: tbar ([52][nlev][ncol]) is the cotangent of the tendencies;
// sbar ([52][nlev][ncol]), Tbar, pbar, qvbar receive the gradient of
// <tbar, tendencies> per cell.  The SOA loop's states are kept on a vector.
extern "C" long cell_host_adjoint(const int *cfgv, int ncol, int nlev, double dt,
                                  const double *state, const double *T, const double *p,
                                  const double *z, const double *qv, const double *cld,
                                  const double *pblh, const double *dga, const double *dgw,
                                  const double *wd, const double *precld, const double *tbar,
                                  double *sbar, double *Tbar, double *pbar, double *qvbar) {
  static ChainConsts K;
  chain_consts_init(&K);
  const long n = (long)ncol * nlev;
  ChainArgs a;
  std::memset(&a, 0, sizeof(a));
  for (int f = 0; f < NSTATE; ++f) {
    a.state[f] = state + f * n;
    if (precld) a.precld[f] = precld + f * n;
  }
  a.T = T; a.p = p; a.z = z; a.qv = qv; a.cld = cld; a.pblh = pblh;
  for (int m = 0; m < 4; ++m) {
    a.dga[m] = dga + m * n;
    a.dgw[m] = dgw + m * n;
    a.wd[m] = wd + m * n;
  }
  a.K = &K;
  a.ncol = ncol; a.nlev = nlev; a.ldc = ncol; a.layout = 0;
  a.cfg.do_cond = cfgv[0]; a.cfg.do_rename = cfgv[1]; a.cfg.do_newnuc = cfgv[2];
  a.cfg.do_coag = cfgv[3]; a.cfg.gaexch_h2so4_uptake_optaa = cfgv[4];
  a.cfg.nmono = (unsigned)cfgv[5];
  a.cfg.ntsubstep = g_ntsubstep;
  a.do_calcsize = cfgv[6]; a.do_xfer = cfgv[7]; a.do_adjust = cfgv[8];
  a.dt = dt;
  VecSoaTape vt;
  dev::rev::SoaTape2<32> t2;
  VecSubTape sub;
  long iters = 0;
  for (int k = 0; k < nlev; ++k)
    for (int i = 0; i < ncol; ++i) {
      const long off = (long)k * ncol + i;
      double tb[NSTATE], sb[NSTATE];
      for (int f = 0; f < NSTATE; ++f) {
        tb[f] = tbar[f * n + off];
        sb[f] = 0.0;
      }
      double tB = 0.0, pB = 0.0, qB = 0.0;
      if (g_tape == 1)
        iters += precld ? dev::rev::adjoint_cell<true>(a, off, i, tb, t2, sub, sb, tB, pB, qB)
                        : dev::rev::adjoint_cell<false>(a, off, i, tb, t2, sub, sb, tB, pB, qB);
      else
        iters += precld ? dev::rev::adjoint_cell<true>(a, off, i, tb, vt, sub, sb, tB, pB, qB)
                        : dev::rev::adjoint_cell<false>(a, off, i, tb, vt, sub, sb, tB, pB, qB);
      for (int f = 0; f < NSTATE; ++f) sbar[f * n + off] = sb[f];
      Tbar[off] = tB;
      pbar[off] = pB;
      qvbar[off] = qB;
    }
  return iters;
}
