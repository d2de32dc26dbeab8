// Stand-alone CPU driver for the sanitizer and floating-point-trap checks:
// the oracle (oracle/mam4_oracle.c) and the CPU build of the kernel's cell
// code (FP64, mixed, tangent, reverse sweep) on one scene, built with
// AddressSanitizer + UndefinedBehaviorSanitizer (tests/test_sanitizers.py),
// optionally under the reference test suite's trapped exceptions
// (FE_DIVBYZERO | FE_INVALID | FE_OVERFLOW, src/validation/validation.hpp:52,
// src/tests/testing.cpp:214-218).  Test infrastructure only.
//
// usage: sanitize_driver <scene prefix> <ncol> <nlev> <ntsubstep> <traps 0|1|2>
// (2: trap, then divide by zero -- the check that trapping works)
// The scene is read from <prefix>.<field>.bin (raw doubles, the layouts of
// mam4xx_amd.scenes); <prefix>.precld.bin is the optional pre-cloud-chemistry
// state.  Exit status 0 when the FP64 cell code equals the oracle bit for bit
// and every derivative is finite; a trapped exception ends it with SIGFPE.
#include <cfenv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../mam4xx_amd/csrc/chain_adjoint.hpp"
#include "../../oracle/mam4_oracle.h"

using namespace mam4x;

namespace {

std::vector<double> read_plane(const std::string &path, size_t n, bool optional) {
  std::vector<double> v;
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) {
    if (optional) return v;
    std::fprintf(stderr, "cannot open %s\n", path.c_str());
    std::exit(2);
  }
  v.resize(n);
  if (std::fread(v.data(), sizeof(double), n, f) != n) {
    std::fprintf(stderr, "short read %s\n", path.c_str());
    std::exit(2);
  }
  std::fclose(f);
  return v;
}

struct VecTape {
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

struct VecSub {
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

}  // namespace

int main(int argc, char **argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s prefix ncol nlev ntsubstep traps\n", argv[0]);
    return 2;
  }
  const std::string pre = argv[1];
  const int ncol = std::atoi(argv[2]), nlev = std::atoi(argv[3]), ns = std::atoi(argv[4]);
  const bool traps = std::atoi(argv[5]) != 0;
  const size_t n = (size_t)ncol * nlev;
  auto st = read_plane(pre + ".state.bin", NSTATE * n, false);
  auto T = read_plane(pre + ".temperature.bin", n, false);
  auto p = read_plane(pre + ".pressure.bin", n, false);
  auto z = read_plane(pre + ".height.bin", n, false);
  auto qv = read_plane(pre + ".qv.bin", n, false);
  auto cld = read_plane(pre + ".cloud_fraction.bin", n, false);
  auto pblh = read_plane(pre + ".pblh.bin", ncol, false);
  auto dga = read_plane(pre + ".dgncur_a.bin", 4 * n, false);
  auto dgw = read_plane(pre + ".dgncur_awet.bin", 4 * n, false);
  auto wd = read_plane(pre + ".wetdens.bin", 4 * n, false);
  auto pc = read_plane(pre + ".precld.bin", NSTATE * n, true);
  const double dt = 1800.0;
  if (traps) feenableexcept(FE_DIVBYZERO | FE_INVALID | FE_OVERFLOW);
  if (std::atoi(argv[5]) == 2) {  // the traps' own check: 1/0 must end the process
    volatile double zero = 0.0;
    std::printf("%g\n", 1.0 / zero);
    return 3;
  }

  // the oracle
  orc_config oc;
  std::memset(&oc, 0, sizeof(oc));
  oc.do_cond = oc.do_rename = oc.do_newnuc = oc.do_coag = 1;
  oc.gaexch_h2so4_uptake_optaa = oc.newnuc_h2so4_conc_optaa = 2;
  oc.n_so4_monolayers_pcage = 8;
  oc.do_calcsize = oc.calcsize_do_aitacc_transfer = oc.calcsize_do_adjust = 1;
  oc.ntsubstep = ns;
  std::vector<double> to(NSTATE * n), dio(4 * n), dco(4 * n);
  orc_chain_step_proc(&oc, ncol, nlev, dt, st.data(), T.data(), p.data(), z.data(), qv.data(),
                      cld.data(), pblh.data(), dga.data(), dgw.data(), wd.data(), nullptr,
                      pc.empty() ? nullptr : pc.data(), to.data(), dio.data(), dco.data(),
                      nullptr, 1);

  // the kernel's cell code
  static ChainConsts K;
  chain_consts_init(&K);
  std::vector<double> th(NSTATE * n), tm(NSTATE * n), tt(NSTATE * n), dtend(NSTATE * n),
      dir(NSTATE * n);
  ChainArgs a;
  std::memset(&a, 0, sizeof(a));
  for (int f = 0; f < NSTATE; ++f) {
    a.state[f] = st.data() + f * n;
    if (!pc.empty()) a.precld[f] = pc.data() + f * n;
  }
  a.T = T.data(); a.p = p.data(); a.z = z.data(); a.qv = qv.data(); a.cld = cld.data();
  a.pblh = pblh.data();
  for (int m = 0; m < 4; ++m) {
    a.dga[m] = dga.data() + m * n;
    a.dgw[m] = dgw.data() + m * n;
    a.wd[m] = wd.data() + m * n;
  }
  a.K = &K;
  a.ncol = ncol; a.nlev = nlev; a.ldc = ncol; a.layout = 0;
  a.cfg.do_cond = a.cfg.do_rename = a.cfg.do_newnuc = a.cfg.do_coag = 1;
  a.cfg.gaexch_h2so4_uptake_optaa = 2;
  a.cfg.nmono = 8;
  a.cfg.ntsubstep = ns;
  a.do_calcsize = a.do_xfer = a.do_adjust = 1;
  a.dt = dt;
  const bool has_pre = !pc.empty();
  auto run = [&](auto policy, std::vector<double> &out) {
    using P = decltype(policy);
    for (int f = 0; f < NSTATE; ++f) a.tend[f] = out.data() + f * n;
    for (int k = 0; k < nlev; ++k)
      for (int i = 0; i < ncol; ++i) {
        const long off = (long)k * ncol + i;
        if (has_pre)
          dev::chain_cell_step<true, P>(a, off, i);
        else
          dev::chain_cell_step<false, P>(a, off, i);
      }
  };
  run(Fp64{}, th);
  run(Mixed{}, tm);
  for (size_t q = 0; q < NSTATE * n; ++q) dir[q] = st[q] * (1e-3 * (double)(q % 7) - 3e-3);
  for (int f = 0; f < NSTATE; ++f) {
    a.dstate[f] = dir.data() + f * n;
    a.dtend[f] = dtend.data() + f * n;
  }
  run(Tangent{}, tt);

  long bad = 0, nonfinite = 0;
  for (size_t q = 0; q < NSTATE * n; ++q) {
    bad += th[q] != to[q];
    nonfinite += !std::isfinite(tm[q]) || !std::isfinite(dtend[q]) || tt[q] != th[q];
  }
  // the reverse sweep, every cell, both SOA storages
  VecTape vt;
  dev::rev::SoaTape2<32> t2;
  VecSub sub;
  for (int k = 0; k < nlev; ++k)
    for (int i = 0; i < ncol; ++i) {
      const long off = (long)k * ncol + i;
      double tb[NSTATE], sb1[NSTATE], sb2[NSTATE];
      for (int f = 0; f < NSTATE; ++f) {
        tb[f] = 1.0 / (std::fabs(st[f * n + off]) + 1e-30) * (f % 2 ? 1e-3 : -1e-3);
        sb1[f] = sb2[f] = 0.0;
      }
      double T1 = 0, p1 = 0, q1 = 0, T2 = 0, p2 = 0, q2 = 0;
      if (has_pre) {
        dev::rev::adjoint_cell<true>(a, off, i, tb, vt, sub, sb1, T1, p1, q1);
        dev::rev::adjoint_cell<true>(a, off, i, tb, t2, sub, sb2, T2, p2, q2);
      } else {
        dev::rev::adjoint_cell<false>(a, off, i, tb, vt, sub, sb1, T1, p1, q1);
        dev::rev::adjoint_cell<false>(a, off, i, tb, t2, sub, sb2, T2, p2, q2);
      }
      for (int f = 0; f < NSTATE; ++f) nonfinite += !std::isfinite(sb1[f]) || sb1[f] != sb2[f];
      nonfinite += !std::isfinite(T1) || !std::isfinite(p1) || !std::isfinite(q1);
    }
  std::printf("cells %zu ntsubstep %d traps %d precld %d: oracle mismatches %ld, "
              "non-finite or inconsistent %ld\n",
              n, ns, (int)traps, (int)has_pre, bad, nonfinite);
  return (bad == 0 && nonfinite == 0) ? 0 : 1;
}
