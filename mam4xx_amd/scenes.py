"""Seeded synthetic scenes for the MAM4 microphysics chain.

A scene is a set of columns on ``nlev`` hybrid levels with the prognostic,
meteorological and modal-diameter inputs of one microphysics step, laid out
``[field][level][column]`` (column index fastest), float64.  The value
distributions follow the benchmark configurations of the reference study:

* surface pressure U(95000, 103000) Pa; ``nlev`` mid-levels log-spaced from the
  model top (100 Pa at 72 levels, 10 Pa for the 128-level large scene) to the
  surface, ``p_k = ptop * (ps/ptop)**((k + 1/2)/nlev)``;
* surface T U(250, 305) K with a 6.5 K/km lapse rate below 11 km and 216 K
  above (the discontinuity at 11 km is kept); heights from hypsometric
  integration (R_d = 287.04, g = 9.80616);
* RH = 0.05 + 0.93 Beta(2,2) per level, qv = RH * qv_sat(T, p) using the
  chain's own Magnus qv_sat (mam4_amicphys.hpp:2199-2205), so the chain's RH
  is the drawn RH;
* cloud fraction 0 on a fraction ``clear_frac`` (0.7) of the levels, else
  U(0.1, 1); PBL height U(200, 2000) m;
* modes acc/ait/coarse/pca with dry geometric-mean diameters 110 nm, 26 nm,
  2 um, 50 nm (wet diameter = dry, no water uptake), wet density =
  max(1000, mass-weighted dry density);
* interstitial and cloud-borne species mass mixing ratios log-uniform in
  [1e-12, 1e-8] kg/kg; numbers N = sum(q_s/rho_s) / (pi/6 D^3 e^{4.5 ln^2 s})
  in #/kg;
* H2SO4 log-uniform [1e-14, 1e-10] mol/mol, SOAG [1e-12, 1e-9] mol/mol, stored
  as kg/kg with the E3SM molecular weights (gas_chem_mechanism.hpp:31-37).

Generation uses ``numpy.random.default_rng(seed)``; the same seed gives the
same scene on any machine.
"""
import math

import numpy as np

from . import layout as L

RD = 287.04
GRAV = 9.80616
MW_H2O = 0.018016
MW_AIR = 0.028966


def qv_sat(T, p):
    """Saturation specific humidity with the chain's improved-Magnus formula
    (conversions.hpp:148-180, mam4_amicphys.hpp:2199-2205)."""
    epsqs = MW_H2O / MW_AIR
    celsius = T - 273.15
    ew = 610.94 * np.exp(17.625 * celsius / (234.04 + celsius))
    ev = 1.00071 * np.exp(4.5e-8 * p) * ew
    return epsqs * ev / (p - (1 - epsqs) * ev)


def _loguniform(rng, lo, hi, size):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size))


def generate(ncol, nlev=72, seed=0, ptop=None, clear_frac=0.7,
             h2so4_range=(1e-14, 1e-10), soag_range=(1e-12, 1e-9)):
    """Return a dict of float64 arrays describing one scene.

    Keys: ``state`` [NSTATE, nlev, ncol] (kg/kg, #/kg), ``temperature``,
    ``pressure``, ``height``, ``qv``, ``cloud_fraction``, ``hydrostatic_dp``
    [nlev, ncol], ``pblh`` [ncol], ``dgncur_a``, ``dgncur_awet``, ``wetdens``
    [4, nlev, ncol].  Mid levels sit at ptop (ps/ptop)^((k+1/2)/nlev) and the
    layer interfaces at ptop (ps/ptop)^(k/nlev), whose differences are
    hydrostatic_dp.
    """
    if ptop is None:
        ptop = 100.0 if nlev <= 72 else 10.0
    rng = np.random.default_rng(seed)
    shape = (nlev, ncol)
    ps = rng.uniform(95000.0, 103000.0, ncol)
    frac = (np.arange(nlev) + 0.5) / nlev
    p = ptop * (ps[None, :] / ptop) ** frac[:, None]
    pint = ptop * (ps[None, :] / ptop) ** (np.arange(nlev + 1) / nlev)[:, None]
    dp = np.diff(pint, axis=0)
    tsfc = rng.uniform(250.0, 305.0, ncol)

    def temp_of_z(z):
        return np.where(z < 11000.0, tsfc - 6.5e-3 * z, 216.0)

    z = np.empty(shape)
    zc = RD * tsfc / GRAV * np.log(ps / p[nlev - 1])
    z[nlev - 1] = zc
    for k in range(nlev - 2, -1, -1):
        tk = temp_of_z(z[k + 1])
        z[k] = z[k + 1] + RD * tk / GRAV * np.log(p[k + 1] / p[k])
    T = temp_of_z(z)
    rh = 0.05 + 0.93 * rng.beta(2.0, 2.0, shape)
    qv = rh * qv_sat(T, p)
    cld = np.where(rng.uniform(0.0, 1.0, shape) < clear_frac, 0.0,
                   rng.uniform(0.1, 1.0, shape))
    pblh = rng.uniform(200.0, 2000.0, ncol)

    state = np.zeros((L.NSTATE,) + shape)
    dgn = np.empty((4,) + shape)
    wetdens = np.empty((4,) + shape)
    for m in range(4):
        dgn[m] = L.MODE_DGN_INIT[m]
        vol_i = np.zeros(shape)
        vol_c = np.zeros(shape)
        mass_i = np.zeros(shape)
        for s in range(L.NSPEC_MODE[m]):
            qi = _loguniform(rng, 1e-12, 1e-8, shape)
            qc = _loguniform(rng, 1e-12, 1e-8, shape)
            state[L.f_qi(m, s)] = qi
            state[L.f_qc(m, s)] = qc
            rho = L.SPECIES_DENSITY[L.MODE_SPECIES[m][s]]
            vol_i += qi / rho
            vol_c += qc / rho
            mass_i += qi
        v1 = math.pi / 6.0 * L.MODE_DGN_INIT[m] ** 3 * math.exp(4.5 * math.log(L.MODE_SIGMA[m]) ** 2)
        state[L.f_ni(m)] = vol_i / v1
        state[L.f_nc(m)] = vol_c / v1
        wetdens[m] = np.maximum(1000.0, mass_i / vol_i)
    state[L.F_H2SO4] = _loguniform(rng, h2so4_range[0], h2so4_range[1], shape) * (98.0784 / 28.966)
    state[L.F_SOAG] = _loguniform(rng, soag_range[0], soag_range[1], shape) * (12.011 / 28.966)
    return {
        "state": state,
        "temperature": T,
        "pressure": p,
        "height": z,
        "qv": qv,
        "cloud_fraction": cld,
        "hydrostatic_dp": dp,
        "pblh": pblh,
        "dgncur_a": dgn,
        "dgncur_awet": dgn.copy(),
        "wetdens": wetdens,
    }


def uneven_batch(nscenes=32, nlev=72, seed=0, min_cols=96, max_cols=21600,
                 total_cols=None):
    """Scenes of uneven column counts (log-uniform in [min_cols, max_cols]),
    each with its own cloudy-level fraction U(0, 0.6) and an H2SO4 range
    shifted so the nucleation-active fraction varies between scenes.

    ``total_cols`` rescales the drawn counts to that total (for small tests).
    Returns a list of scene dicts."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(min_cols), math.log(max_cols), nscenes))
    if total_cols is not None:
        counts = counts * (total_cols / counts.sum())
    counts = np.maximum(1, np.round(counts)).astype(int)
    scenes = []
    for i, n in enumerate(counts):
        cloudy = rng.uniform(0.0, 0.6)
        shift = 10.0 ** rng.uniform(-1.5, 1.0)
        scenes.append(generate(int(n), nlev=nlev, seed=seed * 1000 + i + 1,
                               clear_frac=1.0 - cloudy,
                               h2so4_range=(1e-14 * shift, 1e-10 * shift)))
    return scenes
