"""The oracle's aerosol water uptake (modal_aero_water_uptake_dr,
water_uptake.hpp), pinned on the CPU:

* the closed-form quartic (makoh_quartic) against numpy's companion-matrix
  roots, and its roots satisfy the polynomial;
* the Kohler solve against the reference's own Mathematica roots of the
  Kohler polynomial (tests/golden/kohler_roots.json, 8,000 grid points);
* the Goff & Gratch saturation pressure at its defining point and against
  textbook values;
* the composed cell: the wet radius is a root of the Kohler polynomial above
  the deliquescence RH, dry below the crystallisation RH, the hysteresis
  branch interpolates the water volume linearly, and the wet density closes
  the mass balance.
"""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
AA = 2.0e4 * 18.0 * 76.0 / (8.3e7 * 273.0 * 1.0)  # water_uptake.hpp:193-203
SIGMA = [1.8, 1.6, 1.8, 1.60000002384186]
NSPEC = [7, 4, 7, 3]
# (density, hygroscopicity) of the species of each mode, mode_aero_species
# order: the E3SM types of water_uptake.hpp:312-360
SO4, POM, SOA, BC, NCL, DST, MOM = ((1770.0, 0.507), (1000.0, 0.1000000083e-9),
                                    (1000.0, 0.14), (1700.0, 0.1000000013e-9),
                                    (1900.0, 1.16), (2600.0, 0.068), (1601.0, 0.1000000015))
SPECIES = [[SO4, POM, SOA, BC, DST, NCL, MOM], [SO4, SOA, NCL, MOM],
           [DST, NCL, SO4, BC, POM, SOA, MOM], [POM, BC, MOM]]


@pytest.fixture(scope="module")
def orc():
    import oracle.pyoracle as P
    return P.Oracle()


def test_quartic_matches_companion_roots(orc):
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(400):
        # Kohler-shaped coefficients (p2 = 0, as the reference calls it) and generic ones
        if rng.random() < 0.5:
            rd, hyg, rh = 10 ** rng.uniform(-2, 1), rng.uniform(0.05, 1.2), rng.uniform(0.3, 0.99)
            vol, sl = rd ** 3, math.log(rh)
            p = (-AA / sl, 0.0, vol * hyg / sl - vol, AA * vol / sl)
        else:
            p = tuple(rng.standard_normal(4))
        got = np.sort_complex(orc.makoh_quartic(*p))
        ref = np.sort_complex(np.roots([1.0, *p]))
        scale = max(1.0, np.abs(ref).max())
        # pair each root with its nearest companion root
        d = max(np.abs(ref - g).min() for g in got) / scale
        worst = max(worst, d)
        # residual relative to the polynomial's term magnitudes
        for x in got:
            terms = np.abs([x ** 4, p[0] * x ** 3, p[1] * x ** 2, p[2] * x, p[3]])
            res = abs(x ** 4 + p[0] * x ** 3 + p[1] * x ** 2 + p[2] * x + p[3])
            assert res <= 1e-7 * terms.sum(), (p, x, res, terms.sum())
    assert worst < 1e-6, worst


def test_kohler_solve_matches_reference_mathematica_roots(orc):
    g = json.load(open(os.path.join(HERE, "golden", "kohler_roots.json")))
    roots = np.array(g["roots"])
    # Kelvin coefficient of the data set: kohler.hpp:35-66 with the
    # mam4_constants.hpp values (IAPWS surface tension at the triple point)
    Tt, Tc = 273.16, 647.096
    tau = 1 - Tt / Tc
    sigma = 0.2358 * tau ** 1.256 * (1 + -0.625 * tau)
    Rv = 6.022214076e23 * 1.380649e-23 / 0.018016
    A = 2 * sigma / (Rv * Tt * 1.0e3) * 1e6
    n = 20
    rh = 0.05 + np.arange(n) * (0.98 - 0.05) / (n - 1)
    hy = 1e-6 + np.arange(n) * (1.3 - 1e-6) / (n - 1)
    rd = 1e-3 + np.arange(n) * (30 - 1e-3) / (n - 1)
    err = np.empty(n ** 3)
    for i in range(n):
        for j in range(n):
            for k in range(n):
                idx = n * n * i + n * j + k
                r = orc.kohler_root(A, rd[k], hy[j], rh[i])
                err[idx] = abs(r - roots[idx]) / roots[idx]
    # the closed form is exact to round-off except where the Kohler quartic
    # is ill-conditioned (the smallest dry radii: 16 of 8,000 points above
    # 1e-13, one at 1e-5 -- a property of the reference's closed form)
    assert np.median(err) < 1e-15, np.median(err)
    assert (err > 1e-13).mean() < 0.005, np.sort(err)[-20:]
    assert (err > 1e-9).sum() <= 1, np.sort(err)[-5:]
    assert err.max() < 2e-5, err.max()


def test_goff_gratch_saturation(orc):
    # at the steam point every correction term of Goff & Gratch vanishes and
    # the pressure is exactly the steam-point pressure 1013.246 hPa
    es, qs = orc.qsat_water(373.16, 2.0e5)
    assert es == pytest.approx(101324.6, rel=1e-14)
    eps = 0.018016 / 0.028966
    assert qs == pytest.approx(eps * es / (2.0e5 - (1 - eps) * es), rel=1e-15)
    # textbook saturation pressures over water (Goff & Gratch is within 0.1 %)
    for T, e_ref in ((273.16, 611.66), (283.15, 1228.1), (293.15, 2339.3), (303.15, 4246.0)):
        assert orc.qsat_water(T, 1e5)[0] == pytest.approx(e_ref, rel=2e-3), T
    # p <= es: qs = 1 and es limited to p
    es, qs = orc.qsat_water(373.16, 5.0e4)
    assert qs == 1.0 and es == 5.0e4


def _cell_inputs(rng):
    q = np.zeros((4, 7))
    for m in range(4):
        q[m, :NSPEC[m]] = 10 ** rng.uniform(-12, -8, NSPEC[m])
    dga = np.array([1.1e-7, 2.6e-8, 2e-6, 5e-8]) * 10 ** rng.uniform(-0.3, 0.3, 4)
    return q, dga


def _dry(q, dga):
    """dryaer restated independently: volume-mean hygroscopicity, dry radius,
    dry volume and mass of one particle, number."""
    out = []
    for m in range(4):
        vols = [q[m, s] / SPECIES[m][s][0] for s in range(NSPEC[m])]
        dv = sum(vols)
        hyg = sum(v * SPECIES[m][s][1] for s, v in enumerate(vols)) / dv
        vol1 = math.pi / 6 * dga[m] ** 3 * math.exp(4.5 * math.log(SIGMA[m]) ** 2)
        out.append(dict(hyg=hyg, vol1=vol1, rdry=(vol1 / (4 / 3 * math.pi)) ** (1 / 3),
                        mass1=q[m, :NSPEC[m]].sum() / dv * vol1, naer=dv / vol1))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_water_uptake_cell_branches(orc, seed):
    rng = np.random.default_rng(seed)
    T, p = 280.0, 9.0e4
    _, qs = orc.qsat_water(T, p)  # pinned by test_goff_gratch_saturation
    for _ in range(20):
        q, dga = _cell_inputs(rng)
        dry = _dry(q, dga)
        # below crystallisation: dry
        dgw, qw, wd = orc.water_uptake_cell(q, T, p, 0.2 * qs, 0.0, dga)
        assert np.array_equal(dgw, dga) and np.all(qw == 0.0)
        for m in range(4):
            assert wd[m] == pytest.approx(dry[m]["mass1"] / dry[m]["vol1"], rel=1e-13)
        # above deliquescence: the wet radius is the Kohler root (microns)
        rh = 0.9
        dgw, qw, wd = orc.water_uptake_cell(q, T, p, rh * qs, 0.0, dga)
        for m in range(4):
            rdry_um = dry[m]["rdry"] * 1e6
            r = rdry_um * dgw[m] / dga[m]
            vol, sl, B = rdry_um ** 3, math.log(rh), dry[m]["hyg"]
            if r < 30.0 and B * vol / AA / rdry_um ** 2 >= 1e-4:
                terms = [abs(sl) * r ** 4, AA * r ** 3, abs(B - sl) * vol * r, AA * vol]
                res = sl * r ** 4 - AA * r ** 3 + (B - sl) * vol * r + AA * vol
                assert abs(res) <= 1e-9 * sum(terms), (m, res, terms)
            # water and wet density close the mass balance
            wtr1 = 4 / 3 * math.pi * (r * 1e-6) ** 3 - dry[m]["vol1"]
            assert qw[m] == pytest.approx(1e3 * dry[m]["naer"] * wtr1, rel=1e-9)
            wet1 = dry[m]["vol1"] + wtr1
            assert wd[m] == pytest.approx((dry[m]["mass1"] + 1e3 * wtr1) / wet1, rel=1e-9)
        # hysteresis branch: water volume linear in RH between 0.35 and 0.8
        # (relative to the upper curve at the same RH, which is not formed
        # here): water at RH and at the same RH with cloud fraction scaling
        # the clear-air RH to the same value agree
        a = orc.water_uptake_cell(q, T, p, 0.6 * qs, 0.0, dga)
        cld = 0.25
        b = orc.water_uptake_cell(q, T, p, (0.6 * (1 - cld) + cld) * qs, cld, dga)
        for x, y in zip(a, b):
            np.testing.assert_allclose(x, y, rtol=1e-12)
        assert np.all(a[1] > 0) and np.all(a[1] < qw)
