"""State layout shared by the Python host code (mirrors include/mam4x.h).

The 52 prognostic fields of one cell, in AeroProcess units (kg/kg, #/kg):

====== ============================================================
index  field
====== ============================================================
0-3    interstitial number, modes acc, ait, coarse, pca   (n_mode_i)
4-7    cloud-borne number                                 (n_mode_c)
8-28   interstitial species mass, mode-major, species in
       mode_aero_species order (aero_modes.hpp:214-243)   (q_aero_i)
29-49  cloud-borne species mass, same order               (q_aero_c)
50     H2SO4 gas                                          (q_gas)
51     SOAG gas                                           (q_gas)
====== ============================================================

Every field is a [level][column] plane; a full state is
``[NSTATE][nlev][ncol]`` with the column index fastest.
"""

NSTATE = 52
NMODE = 4
MODE_NAMES = ("acc", "ait", "coarse", "pca")
SPECIES_NAMES = ("soa", "so4", "pom", "bc", "ncl", "dst", "mom")  # AeroId order
SPECIES_DENSITY = (1000.0, 1770.0, 1000.0, 1700.0, 1900.0, 2600.0, 1601.0)
NSPEC_MODE = (7, 4, 7, 3)
# species of each mode, AeroId indices (aero_modes.hpp:214-243)
MODE_SPECIES = ((1, 2, 0, 3, 5, 4, 6), (1, 0, 4, 6), (5, 4, 1, 3, 2, 0, 6), (2, 3, 6))
MODE_OFFSET = (0, 7, 11, 18)
MODE_SIGMA = (1.8, 1.6, 1.8, 1.60000002384186)
MODE_DMIN = (5.35e-8, 8.7e-9, 1e-6, 1e-8)
MODE_DNOM = (1.1e-7, 2.6e-8, 2e-6, 5e-8)
MODE_DMAX = (4.4e-7, 5.2e-8, 4e-6, 1e-7)
MODE_DGN_INIT = (110e-9, 26e-9, 2e-6, 50e-9)
F_H2SO4 = 50
F_SOAG = 51


def f_ni(m):
    return m


def f_nc(m):
    return 4 + m


def f_qi(m, s):
    return 8 + MODE_OFFSET[m] + s


def f_qc(m, s):
    return 29 + MODE_OFFSET[m] + s


def field_names():
    names = []
    for m in range(4):
        names.append("num_%s_i" % MODE_NAMES[m])
    for m in range(4):
        names.append("num_%s_c" % MODE_NAMES[m])
    for tag in ("i", "c"):
        for m in range(4):
            for s in range(NSPEC_MODE[m]):
                names.append("%s_%s_%s" % (SPECIES_NAMES[MODE_SPECIES[m][s]], MODE_NAMES[m], tag))
    names += ["h2so4", "soag"]
    return names


def is_number(f):
    return f < 8


def is_mass(f):
    return 8 <= f < 50


def is_gas(f):
    return f >= 50


def is_cloud_borne(f):
    return 4 <= f < 8 or 29 <= f < 50


# ---- per-process tendency diagnostics (include/mam4x.h, process_tendencies)
# Blocks of the gas_aero_exchange_* arrays (diagnostic_arrays.hpp:13-18),
# each over the 31 gas_pcnst tracers, in mol/mol/s (#/kmol/s for numbers).
GAS_PCNST = 31
PROCESS_NAMES = ("condensation", "renaming", "nucleation", "coagulation",
                 "renaming_cloud_borne")
NPROCESS = 5
# E3SM adv_mass of the gas_pcnst tracers [g/mol] (gas_chem_mechanism.hpp:31-37)
ADV_MASS = (47.998200, 34.013600, 98.078400, 64.064800, 62.132400, 12.011000,
            115.107340, 12.011000, 12.011000, 12.011000, 135.064039, 58.442468,
            250092.672000, 1.007400, 115.107340, 12.011000, 58.442468, 250092.672000,
            1.007400, 135.064039, 58.442468, 115.107340, 12.011000, 12.011000,
            12.011000, 250092.672000, 1.007400, 12.011000, 12.011000, 250092.672000,
            1.007400)
MWAIR = 28.966  # g/mol (mam4_constants.hpp:60)
_NUMPTR = (13, 18, 26, 30)          # mam4_amicphys.hpp:35-63
_MASS_SLOT0 = (6, 14, 19, 27)


def pcnst_slot(f):
    """gas_pcnst index of state field f (interstitial and cloud-borne
    copies of a species share it)"""
    if f < 8:
        return _NUMPTR[f % 4]
    if f == F_H2SO4:
        return 2
    if f == F_SOAG:
        return 5
    r = (f - 8) % 21
    m = 3 if r >= MODE_OFFSET[3] else 2 if r >= MODE_OFFSET[2] else 1 if r >= MODE_OFFSET[1] else 0
    return _MASS_SLOT0[m] + r - MODE_OFFSET[m]


def to_vmr(f):
    """kg/kg -> mol/mol (#/kg -> #/kmol for numbers) for state field f"""
    return MWAIR if f < 8 else MWAIR / ADV_MASS[pcnst_slot(f)]
