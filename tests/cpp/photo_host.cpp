// CPU build of the HIP kernels' photolysis passes (mam4xx_amd/csrc/photo.hpp)
// for the test suite: built with g++ and glibc, the oracle's own libm, it
// must reproduce the oracle's table_photo bit for bit.  The table argument has
// the oracle's orc_photo_table layout (pht::Table's).  Test infrastructure
// only.
#include <cmath>
#include <vector>

#include "../../mam4xx_amd/csrc/photo.hpp"

using namespace mam4x;

// planes [nlev][ncol] (column fastest): photo (in place), pmid, pdel, temp,
// colo3, lwc, clouds; per column zen, alb.  The column pass writes the
// effective albedo and multiplier into eff / mult [nlev][ncol] as the kernel
// writes its work planes.
extern "C" void photo_host_run(const pht::Table *t, int ncol, int nlev, double *photo,
                               const double *pmid, const double *pdel, const double *temp,
                               const double *colo3, const double *zen, const double *alb,
                               const double *lwc, const double *clouds, double esfact,
                               double *eff, double *mult) {
  if (t->ind < 0) return;
  for (int j = 0; j < ncol; ++j) {
    if (!pht::daylight(pht::sza_deg(zen[j]))) continue;
    pht::cloud_column(nlev, ncol, zen[j], clouds + j, lwc + j, pdel + j, alb[j], eff + j,
                      mult + j);
  }
  for (int k = 0; k < nlev; ++k)
    for (int j = 0; j < ncol; ++j) {
      const double sza = pht::sza_deg(zen[j]);
      if (!pht::daylight(sza)) continue;
      const long i = (long)k * ncol + j;
      const pht::SzaPos s = pht::sza_pos(*t, sza);
      const double jr = pht::level_rate(*t, s, eff[i], pmid[i] * pht::kPaToMb, temp[i], colo3[i]);
      photo[i] = pht::update(*t, photo[i], mult[i], esfact, jr);
    }
}
