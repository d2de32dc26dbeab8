// CPU build of the HIP kernel's washout column (mam4xx_amd/csrc/sethet.hpp)
// for the test suite: built with g++ and glibc, the oracle's own libm, it
// must reproduce the oracle bit for bit.  Test infrastructure only.
#include <cmath>

#include "../../mam4xx_amd/csrc/sethet.hpp"

using namespace mam4x;

// planes [nlev][ncol] (column fastest): press, zmid, tfld, cmfdqr, nrain,
// nevapr, inv, h2o2; per column rlat, phis; het [31][nlev][ncol] out
extern "C" void sethet_host_run(int ncol, int nlev, double delt, double xkgm, const double *rlat,
                                const double *phis, const double *press, const double *zmid,
                                const double *tfld, const double *cmfdqr, const double *nrain,
                                const double *nevapr, const double *inv, const double *h2o2,
                                double *het) {
  const long plane = (long)nlev * ncol;
  for (long i = 0; i < 31 * plane; ++i) het[i] = 0.0;
  for (int j = 0; j < ncol; ++j)
    sh::column(nlev, ncol, rlat[j], phis[j], delt, xkgm, press + j, zmid + j, tfld + j,
               cmfdqr + j, nrain + j, nevapr + j, inv + j, h2o2 + j, het + plane + j,
               het + 2 * plane + j, het + 3 * plane + j);
}
