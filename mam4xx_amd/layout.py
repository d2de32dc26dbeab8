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
