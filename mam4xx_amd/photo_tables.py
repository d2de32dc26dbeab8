"""Synthetic photolysis tables and columns (the tests' and the benchmark's
inputs of mam4x_compute_photolysis).  The reference ships no table (E3SM
reads them from its input files), so the values are synthetic; the axes
follow the conventions mo_photo.hpp relies on: zenith angles, albedos and
ozone ratios increasing, pressures decreasing from the surface, del_* the
reciprocal spacings."""
import numpy as np

# the axes of the reference test's table (src/tests/mam4_photo_setcol.cpp:143-248)
REF_AXES = dict(sza=[0.0, 45.0, 88.0], alb=[0.0, 0.3, 0.8],
                press=[1000.0, 500.0, 100.0, 50.0, 10.0],
                colo3=[5e17, 4e17, 3e17, 2e17, 1e17], o3rat=[0.5, 1.0, 1.5, 2.0],
                prs=[500.0, 100.0, 10.0, 1.0])


def wide_axes(nump=51, numsza=10, numcolo3=10, numalb=6, np_xs=15):
    """axes spanning a 72-level column (10 Pa to the surface)"""
    press = 1000.0 * np.logspace(0.0, -4.0, nump)            # [hPa], surface first
    return dict(sza=np.linspace(0.0, 88.0, numsza),
                alb=np.linspace(0.0, 1.0, numalb),
                press=press,
                colo3=1e19 * (press / 1000.0) ** 0.6,        # ozone column above each
                o3rat=np.linspace(0.1, 5.0, numcolo3),
                prs=1000.0 * np.logspace(0.0, -5.0, np_xs))


def _dels(ax):
    """the del_* arrays: reciprocals of each axis's spacings"""
    d = lambda v, dec: 1.0 / ((v[:-1] - v[1:]) if dec else (v[1:] - v[:-1]))
    a = {k: np.asarray(v, dtype=np.float64) for k, v in ax.items()}
    a["del_sza"] = d(a["sza"], False)
    a["del_alb"] = d(a["alb"], False)
    a["del_p"] = d(a["press"], True)
    a["del_o3rat"] = d(a["o3rat"], False)
    a["dprs"] = d(a["prs"], True)
    return a


def random_table(rng, axes, nw=8, numj=2, nt=201):
    """log-uniform positive source function and cross sections"""
    t = _dels(axes)
    nump, numsza, numcolo3, numalb = (len(t[k]) for k in ("press", "sza", "o3rat", "alb"))
    np_xs = len(t["prs"])
    t["rsf_tab"] = 10 ** rng.uniform(-1, 1, (nw, nump, numsza, numcolo3, numalb))
    t["xsqy"] = 10 ** rng.uniform(-21, -18, (numj, nw, nt, np_xs))
    t["etfphot"] = 1e13 * rng.uniform(0.8, 1.8, nw)
    return t


def linear_table(rng, axes, coef, nw=5, numj=1, nt=201, xs_e=0.0):
    """rsf_tab(w, iz, is, iv, ial) = E_w (1 + a press) (1 + b sza) (1 + c o3rat)
    (1 + d alb) for coef = (a, b, c, d); xsqy(j, w, it, ip) = X_jw (it + 1)
    (1 + xs_e prs).  Returns (table, E, X)."""
    t = _dels(axes)
    a, b, c, d = coef
    E = rng.uniform(0.5, 2.0, nw)
    f = lambda v, k: 1.0 + k * np.asarray(v)
    t["rsf_tab"] = (E[:, None, None, None, None] * f(t["press"], a)[None, :, None, None, None]
                    * f(t["sza"], b)[None, None, :, None, None]
                    * f(t["o3rat"], c)[None, None, None, :, None]
                    * f(t["alb"], d)[None, None, None, None, :])
    X = rng.uniform(0.5, 2.0, (numj, nw)) * 1e-20
    it = np.arange(nt, dtype=np.float64)
    t["xsqy"] = (X[:, :, None, None] * (it + 1.0)[None, None, :, None]
                 * f(t["prs"], xs_e)[None, None, None, :])
    t["etfphot"] = rng.uniform(0.8, 1.8, nw) * 1e13
    return t, E, X


def columns(rng, nlev, ncol, axes, night_frac=0.2):
    """nlev levels from 10 Pa to the surface: photo (in), pmid, pdel, temp,
    colo3, lwc, clouds [nlev, ncol]; zen_angle [rad], srf_alb [ncol]"""
    ps = rng.uniform(95000, 103000, ncol)
    edges = 10.0 * (ps / 10.0) ** (np.arange(nlev + 1)[:, None] / nlev)
    pmid = np.sqrt(edges[1:] * edges[:-1])
    pdel = edges[1:] - edges[:-1]
    temp = np.where(pmid < 2e4, rng.uniform(185, 230, (nlev, ncol)),
                    rng.uniform(220, 305, (nlev, ncol)))
    press, col = np.asarray(axes["press"]), np.asarray(axes["colo3"])
    ref = np.interp(pmid / 100.0, press[::-1], col[::-1])
    colo3 = ref * 10 ** rng.uniform(-0.4, 0.6, (nlev, ncol))
    cloudy = rng.uniform(size=(nlev, ncol)) < 0.3
    clouds = np.where(cloudy, rng.uniform(0.02, 1.0, (nlev, ncol)), 0.0)
    lwc = np.where(cloudy, 10 ** rng.uniform(-6, -3, (nlev, ncol)), 0.0)
    zen = np.deg2rad(rng.uniform(0.0, 87.0, ncol))
    night = rng.uniform(size=ncol) < night_frac
    zen[night] = np.deg2rad(rng.uniform(88.9, 160.0, night.sum()))
    return dict(photo=10 ** rng.uniform(-7, -5, (nlev, ncol)), pmid=pmid, pdel=pdel, temp=temp,
                colo3=colo3, lwc=lwc, clouds=clouds, zen_angle=zen,
                srf_alb=rng.uniform(0.03, 0.9, ncol))
