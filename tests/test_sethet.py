"""Washout rates (mo_sethet.hpp) of the oracle, pinned by what the reference
only follows: the precipitation rescale preserves the column's net total
and removes negative (evaporating) levels; the tropopause proxy is the
lowest level above 100 hPa (300 hPa poleward of 60 degrees); washout never
creates gas, stops where the rain is saturated and stays within what the
rain can carry; the in-cloud rate is the closed form of a drop in Henry's-
law equilibrium; levels without rain or above the tropopause proxy get no
rate.  Then the kernel's code built for the CPU against the oracle, bit for
bit."""
import ctypes
import os
import subprocess

import numpy as np

KB, NA = 1.380649e-23, 6.022214076e23


def columns(rng, nlev, ncol):
    ps = rng.uniform(95000, 103000, ncol)
    edges = 10.0 * (ps / 10.0) ** (np.arange(nlev + 1)[:, None] / nlev)
    press = np.sqrt(edges[1:] * edges[:-1])
    T = np.where(press < 2e4, rng.uniform(200, 220, (nlev, ncol)),
                 rng.uniform(220, 300, (nlev, ncol)))
    # the reference's zmid argument: heights [m] (sethet passes
    # Atmosphere::height; its formulas read km)
    zmid = 7000.0 * np.log(ps / press)
    cloudy = rng.uniform(size=(nlev, ncol)) < 0.35
    cmfdqr = np.where(cloudy, 10 ** rng.uniform(-10, -7, (nlev, ncol)), 0.0)
    nrain = np.where(cloudy, 10 ** rng.uniform(-10, -7, (nlev, ncol)), 0.0)
    nevapr = np.where(rng.uniform(size=(nlev, ncol)) < 0.3,
                      10 ** rng.uniform(-10, -7.5, (nlev, ncol)), 0.0)
    dry = rng.uniform(size=ncol) < 0.15
    cmfdqr[:, dry] = nrain[:, dry] = 0.0
    return dict(press=press, zmid=zmid, tfld=T, cmfdqr=cmfdqr, nrain=nrain, nevapr=nevapr,
                inv_m=press / (KB * T) * 1e-6, h2o2=10 ** rng.uniform(-10, -9, (nlev, ncol)),
                so2=10 ** rng.uniform(-11, -9, (nlev, ncol)),
                rlat=rng.uniform(-1.5, 1.5, ncol), phis=rng.uniform(0, 3e4, ncol))


def test_precip_rescale(orc):
    rng = np.random.default_rng(1)
    for _ in range(50):
        a, b, c = (10 ** rng.uniform(-10, -7, 72) * (rng.uniform(size=72) < 0.5)
                   for _ in range(3))
        p = orc.precip_rescale(a, b, c)
        net = np.sum(a + b - c)
        assert (p >= 0).all()
        if net <= 0:
            assert not p.any()
        else:
            assert abs(p.sum() - net) <= 1e-13 * net
            raw = np.maximum(a + b - c, 0)
            np.testing.assert_allclose(p, raw * net / raw.sum(), rtol=1e-14)


def test_find_ktop(orc):
    p = 10.0 * (1e5 / 10.0) ** ((np.arange(72) + 0.5) / 72)
    for lat, lim in ((0.3, 1e4), (-0.9, 1e4), (1.1, 3e4), (-1.2, 3e4)):
        k = orc.find_ktop(lat, p)
        assert p[k] < lim <= p[k + 1]
    assert orc.find_ktop(0.0, np.full(10, 2e4)) == 0  # no level above the limit


def test_mass_transfer_coefficient(orc):
    """Frossling (1938): k = D/d (2 + 0.6 Re^(1/2) Sc^(1/3)), Re = d u / nu,
    Sc = nu / D, for a 1.89 mm drop at 7.48 m/s"""
    D, d, u, nu = .112, .189, 748.0, 6.18e-2
    want = D / d * (2 + 0.6 * np.sqrt(d * u / nu) * (nu / D) ** (1 / 3))
    assert abs(orc.sethet_xkgm() / want - 1) < 1e-14


def test_washout_removes_and_saturates(orc):
    rng = np.random.default_rng(2)
    nlev = 40
    xkgm = orc.sethet_xkgm()
    T = rng.uniform(250, 295, nlev)
    delz = rng.uniform(2e4, 8e4, nlev)
    xhen = 7.45e4 * np.exp(6620.0 * (298.0 - T) / (298.0 * T))
    x0 = 10 ** rng.uniform(9, 11, nlev)
    rain = np.where(rng.uniform(size=nlev) < 0.5, 1e6, 0.0)
    xliq = np.where(rain > 0, 10 ** rng.uniform(-2, 0, nlev), 0.0)
    g = orc.gas_washout(5, xkgm, xliq, rain, xhen, T, delz, x0)
    assert (g >= 0).all() and (g <= x0).all()
    np.testing.assert_array_equal(g[:5], x0[:5])  # above ktop
    first = np.argmax(rain[5:] > 0) + 5
    np.testing.assert_array_equal(g[:first], x0[:first])  # above the first raining level
    assert (g[first:] < x0[first:]).any()
    # no rain anywhere: nothing changes
    np.testing.assert_array_equal(orc.gas_washout(0, xkgm, xliq, 0 * rain, xhen, T, delz, x0), x0)
    # a saturated column (equilibrium concentration below the first removal)
    # keeps its gas: tiny Henry constant
    np.testing.assert_array_equal(
        orc.gas_washout(0, xkgm, xliq, rain, 1e-30 * xhen, T, delz, x0), x0)


def test_rates_closed_forms(orc):
    """above ktop and where it does not rain: 0; elsewhere the H2SO4 rate
    is the in-cloud rate of a drop in Henry's-law equilibrium with the
    reference's HNO3 constant, sat * rain / (h2o_mol (L + 1/(H k T))), and
    SO2 takes the H2O2 rate"""
    rng = np.random.default_rng(3)
    cols = columns(rng, 72, 64)
    het = orc.sethet(1800.0, cols)
    assert not het[[0] + list(range(4, 31))].any()
    np.testing.assert_array_equal(het[3], het[1])
    for j in range(64):
        k = orc.find_ktop(cols["rlat"][j], cols["press"][:, j])
        assert not het[1:4, :k + 1, j].any()
        pr = orc.precip_rescale(cols["cmfdqr"][:, j], cols["nrain"][:, j], cols["nevapr"][:, j])
        inv = cols["inv_m"][:, j]
        rain = 29.0 * pr * inv / 18.016
        assert not het[1:4, k + 1:, j][:, rain[k + 1:] <= 0].any()
        T = cols["tfld"][:, j]
        xliq = pr * 1800.0 * inv / NA * 29.0 * 1e6
        hen = 2.1e5 * np.exp(8700.0 * (298.0 - T) / (298.0 * T)) * (1 + 15.4 / 1e-5)
        want = 0.016 * rain / (1e3 / 18.016 * (NA * 1e-9 * xliq + 1 / (hen * KB * 10 * T)))
        sel = np.arange(72) > k
        sel &= rain > 0
        np.testing.assert_allclose(het[2, sel, j], want[sel], rtol=1e-12)
        assert (het[1, sel, j] >= 0).all()
    assert (het[1] > 0).sum() > 200 and (het[2] > 0).sum() > 200


def test_kernel_code_bitwise_equals_oracle(orc, tmp_path):
    here = os.path.dirname(os.path.abspath(__file__))
    so = str(tmp_path / "libsethethost.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o",
                    so, os.path.join(here, "cpp", "sethet_host.cpp")], check=True)
    lib = ctypes.CDLL(so)
    rng = np.random.default_rng(4)
    nlev, ncol = 72, 300
    cols = columns(rng, nlev, ncol)
    ref = orc.sethet(1800.0, cols)
    dp = ctypes.POINTER(ctypes.c_double)
    p = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(dp)
    het = np.full((31, nlev, ncol), np.nan)
    args = [np.ascontiguousarray(cols[k]) for k in ("rlat", "phis", "press", "zmid", "tfld",
                                                    "cmfdqr", "nrain", "nevapr", "inv_m",
                                                    "h2o2")]
    lib.sethet_host_run.restype = None
    lib.sethet_host_run(ncol, nlev, ctypes.c_double(1800.0), ctypes.c_double(orc.sethet_xkgm()),
                        *[p(a) for a in args], het.ctypes.data_as(dp))
    assert np.array_equal(het, ref), np.argwhere(het != ref)[:3]
    assert (ref[1] > 0).sum() > 1000
