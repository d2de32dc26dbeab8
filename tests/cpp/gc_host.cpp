// CPU build of the HIP kernel's per-cell gas-phase chemistry
// (mam4xx_amd/csrc/gas_chem.hpp) for the test suite: built with g++ and
// glibc, the oracle's own libm, it must reproduce the oracle bit for bit.
// Test infrastructure only.
#include <cmath>

#include "../../mam4xx_amd/csrc/gas_chem.hpp"

using namespace mam4x;

// planes [.][ncell]: temp, photo; extfrc [9][ncell] (#/cm3/s); inv
// [8][ncell]; het [31][ncell]; qq [31][ncell] updated in place
extern "C" void gc_host_run(long ncell, double dt, const double *temp, const double *photo,
                            const double *extfrc, const double *inv, const double *het,
                            double *qq, int *steps) {
  for (long c = 0; c < ncell; ++c) {
    double v[gc::kInv], e[gc::kExt], h[31], y[31], rxt[gc::kRxn];
    for (int i = 0; i < gc::kInv; ++i) v[i] = inv[i * ncell + c];
    for (int i = 0; i < gc::kExt; ++i) e[i] = extfrc[i * ncell + c] / v[0];
    for (int i = 0; i < 31; ++i) {
      h[i] = het[i * ncell + c];
      y[i] = qq[i * ncell + c];
    }
    gc::rates(temp[c], photo[c], v, rxt);
    steps[c] = gc::imp_sol(y, rxt, h, e, dt);
    for (int i = 0; i < 31; ++i) qq[i * ncell + c] = y[i];
  }
}
