// Gas-phase chemistry of one cell: gas_phase_chemistry
// (gas_phase_chemistry.hpp:37-116) with the pp_linoz_mam4_resus_mom_soag
// mechanism (gas_chem_mechanism.hpp) and its implicit Euler solver imp_sol
// / newton_raphson_iter (gas_chem.hpp:73-425).  FP64; one lane per cell.
//
// The mechanism's implicit class is gas_pcnst species 1..30 in order (its
// permutation is the identity) and its sparse LU holds 32 entries: the
// diagonal plus the SO2 -> H2SO4 and DMS -> SO2 couplings.  The reactions
// are linear in the implicit species, so the first Newton step lands on the
// backward-Euler solution and the second confirms convergence.
#pragma once

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#else
#include <cmath>
// plain C++ compilers see the device annotations as no-ops (the test suite
// builds this file for the CPU: tests/cpp/gc_host.cpp)
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#ifndef __forceinline__
#define __forceinline__ inline
#endif
#endif

namespace mam4x {
namespace gc {

constexpr int kRxn = 7, kCls = 30, kExt = 9, kInv = 8;

// reaction rates: setrxt, usrrxt and adjrxt (gas_chem_mechanism.hpp:40-71,
// gas_chem.hpp:20-63, indices of gas_phase_chemistry.hpp:28-29); rxt[0] is
// the photolysis rate.  inv: M, N2, O2, H2O, OH, NO3, HO2, ... [#/cm3]
__host__ __device__ __forceinline__ void rates(double temp, double photo, const double inv[kInv],
                                               double rxt[kRxn]) {
  const double mtot = inv[0];
  rxt[0] = photo;
  rxt[2] = 2.9000000000e-12 * exp(-160.000000 / temp);
  rxt[4] = 9.6000000000e-12 * exp(-234.000000 / temp);
  rxt[6] = 1.9000000000e-13 * exp(520.000000 / temp);
  {  // HO2 + HO2 -> H2O2, enhanced by water vapour
    const double ko = 3.5e-13 * exp(430.0 / temp);
    const double kinf = 1.7e-33 * mtot * exp(1000. / temp);
    const double fc = 1.0 + 1.4e-21 * inv[3] * exp(2200. / temp);
    rxt[1] = (ko + kinf) * fc;
  }
  {  // DMS + OH -> 0.5 SO2 (addition channel)
    const double ko = 1.0 + 5.5e-31 * exp(7460. / temp) * mtot * 0.21;
    rxt[5] = 1.7e-42 * exp(7810. / temp) * mtot * 0.21 / ko;
  }
  {  // SO2 + OH -> H2SO4 (Troe fall-off)
    const double fc = 3.0e-31 * pow(300. / temp, 3.3);
    const double ko = fc * mtot / (1.0 + fc * mtot / 1.5e-12);
    const double lg = log10(fc * mtot / 1.5e-12);
    rxt[3] = ko * pow(0.6, 1.0 / (1.0 + lg * lg));
  }
  rxt[2] *= inv[4];
  rxt[3] *= inv[4];
  rxt[4] *= inv[4];
  rxt[5] *= inv[4];
  rxt[6] *= inv[5];
  rxt[1] *= inv[6] * inv[6] / mtot;
}

// class-independent production (indprd, class 4, gas_chem_mechanism.hpp:
// 100-143): entry k of the class is gas_pcnst species k + 1
__host__ __device__ __forceinline__ double ind_prd(int k, const double rxt[kRxn],
                                                   const double ext[kExt]) {
  switch (k) {
    case 0: return +rxt[1];
    case 2: return +ext[0];
    case 4: return +ext[8];
    case 5: return +ext[1];
    case 12: return +ext[5];
    case 13: return +ext[2];
    case 17: return +ext[6];
    case 26: return +ext[3];
    case 27: return +ext[4];
    case 29: return +ext[7];
    default: return 0.0;
  }
}

// imp_sol over delt: y[31] (gas_pcnst, mol/mol) advanced in place; het[31]
// washout rates [1/s]; ext[9] forcing already divided by M.  Returns the
// time steps taken.
__host__ __device__ __forceinline__ int imp_sol(double y[31], const double rxt[kRxn],
                                                const double het[31], const double ext[kExt],
                                                double delt) {
  // the coupled block (class 0..3: H2O2, H2SO4, SO2, DMS): diagonal of the
  // linear matrix and the two couplings (linmat, :229-268)
  const double m0 = -(+rxt[0] + rxt[2] + het[1]);
  const double m1 = -(+het[2]);
  const double c2 = +rxt[3];
  const double m3 = -(+rxt[3] + het[3]);
  const double c4 = +rxt[4] + 0.500000 * rxt[5] + rxt[6];
  const double m5 = -(+rxt[4] + rxt[5] + rxt[6] + het[4]);
  double sol[kCls], inv_part[kCls];
  double dt = delt, done = 0.0;
  int cut_cnt = 0, stp_con_cnt = 0, nstep = 0;
  for (int step = 0; step < 1000; ++step) {
    ++nstep;
    const double dti = 1.0 / dt;
#pragma unroll
    for (int k = 0; k < kCls; ++k) {
      sol[k] = y[k + 1];
      inv_part[k] = dti * sol[k] + ind_prd(k, rxt, ext);
    }
    // nlnmat + lu_fac: the factored diagonal (entries 2 and 4 are not
    // inverted)
    const double l0 = 1.0 / (m0 - dti), l1 = 1.0 / (m1 - dti), l3 = 1.0 / (m3 - dti),
                 l5 = 1.0 / (m5 - dti);
    bool convergence = false;
    for (int it = 0; it < 11; ++it) {
      // f(y) with prod / loss of the current iterate (imp_prod_loss)
      double f[kCls];
      f[0] = sol[0] * dti - (inv_part[0] + 0.0 - (het[1] + rxt[0] + rxt[2]) * y[1]);
      f[1] = sol[1] * dti - (inv_part[1] + rxt[3] * y[3] - het[2] * y[2]);
      f[2] = sol[2] * dti -
             (inv_part[2] + (rxt[4] + 0.500000 * rxt[5] + rxt[6]) * y[4] - (het[3] + rxt[3]) * y[3]);
      f[3] = sol[3] * dti - (inv_part[3] + 0.0 - (het[4] + rxt[4] + rxt[5] + rxt[6]) * y[4]);
#pragma unroll
      for (int k = 4; k < kCls; ++k)
        f[k] = sol[k] * dti - (inv_part[k] + 0.0 - het[k + 1] * y[k + 1]);
      // lu_slv (:183-228)
#pragma unroll
      for (int k = 4; k < kCls; ++k) f[k] *= 1.0 / (-(+het[k + 1]) - dti);
      f[3] *= l5;
      f[2] -= c4 * f[3];
      f[2] *= l3;
      f[1] -= c2 * f[2];
      f[1] *= l1;
      f[0] *= l0;
      // update, limit, transfer, convergence (newton_raphson_iter, :143-224)
      bool conv = true;
#pragma unroll
      for (int k = 0; k < kCls; ++k) {
        sol[k] += f[k];
        if (sol[k] < 0.0) sol[k] = 0.0;
        y[k + 1] = sol[k];
        if (fabs(f[k]) > 1.0e-40 && !(fabs(f[k]) <= 1.0e-3 * fabs(sol[k]))) conv = false;
      }
      if (it > 0 && conv) {
        convergence = true;
        break;
      }
    }
    if (!convergence) {
      stp_con_cnt = 0;
      if (cut_cnt < 5) {
        cut_cnt += 1;
        dt *= cut_cnt < 5 ? 0.5 : 0.1;
      }
    }
    done += dt;
    if (fabs(delt - done) <= 0.0001) break;
    if (convergence) stp_con_cnt += 1;
    if (stp_con_cnt >= 2) {
      dt *= 2.0;
      stp_con_cnt = 0;
    }
    dt = (delt - done) < dt ? (delt - done) : dt;
  }
#pragma unroll
  for (int k = 0; k < kCls; ++k) y[k + 1] = sol[k];
  return nstep;
}

}  // namespace gc
}  // namespace mam4x
