"""The wet-size producer's oracle on the CPU: the E3SM-layout calcsize
(modal_aero_calcsize_sub) and its composition with water uptake
(compute_calcsize_and_water_uptake_dr).

modal_aero_calcsize_sub runs CalcSize's algorithm (pinned by the chain's
tests) with two changes; each is pinned by a case that isolates it:
* inside the size bounds and without a transfer that involves
  non-transferable species, both calcsizes agree bit for bit;
* szadj_block_fac: an Aitken (accumulation) population above (below) the
  size bounds keeps its own diameter -- the lognormal closed form -- where
  CalcSize clips it, as long as no inter-mode transfer resizes it;
* the accumulation -> Aitken transfer counts the non-transferable species'
  number with num2vol_ratio_max: the number tendency matches the closed form
  of the transfer fraction with that ratio, and mass and number are
  conserved by the transfer.
"""
import math

import numpy as np
import pytest

SIGMA = [1.8, 1.6, 1.8, 1.60000002384186]
DNOM = [1.1e-7, 2.6e-8, 2e-6, 5e-8]
DMIN = [5.35e-8, 8.7e-9, 1e-6, 1e-8]
DMAX = [4.4e-7, 5.2e-8, 4e-6, 1e-7]
NSPEC = [7, 4, 7, 3]
# densities in mode_aero_species order (aero_modes.hpp:154-243)
DENS = [[1770.0, 1000.0, 1000.0, 1700.0, 2600.0, 1900.0, 1601.0],
        [1770.0, 1000.0, 1900.0, 1601.0],
        [2600.0, 1900.0, 1770.0, 1700.0, 1000.0, 1000.0, 1601.0],
        [1000.0, 1700.0, 1601.0]]
ACC, AIT = 0, 1
DT = 1800.0


@pytest.fixture(scope="module")
def orc():
    import oracle.pyoracle as P
    return P.Oracle()


def v2n(m, d):
    """number per dry volume of a lognormal mode of median diameter d"""
    return 1.0 / (math.pi / 6 * d ** 3 * math.exp(4.5 * math.log(SIGMA[m]) ** 2))


def dgn(m, vol_per_num):
    return (vol_per_num / (math.pi / 6)) ** (1 / 3) * math.exp(-1.5 * math.log(SIGMA[m]) ** 2)


def _masses(rng, scale=1e-9):
    q = np.zeros((4, 7))
    for m in range(4):
        q[m, :NSPEC[m]] = scale * rng.uniform(0.2, 1.0, NSPEC[m])
    return q


def _drv(q, m):
    return sum(q[m, s] / DENS[m][s] for s in range(NSPEC[m]))


def test_calcsizes_agree_inside_bounds(orc):
    rng = np.random.default_rng(11)
    for _ in range(50):
        qi, qc = _masses(rng), _masses(rng)
        # numbers at sizes inside each mode's window (no inter-mode transfer)
        f = rng.uniform(0.85, 1.15, (2, 4))
        ni = np.array([_drv(qi, m) * v2n(m, DNOM[m] * f[0, m]) for m in range(4)])
        nc = np.array([_drv(qc, m) * v2n(m, DNOM[m] * f[1, m]) for m in range(4)])
        a = orc.calcsize_level(DT, ni, nc, qi, qc, e3sm=False)
        b = orc.calcsize_level(DT, ni, nc, qi, qc, e3sm=True)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("mode", [AIT, ACC])
def test_szadj_block_keeps_the_unbounded_size(orc, mode):
    rng = np.random.default_rng(12 + mode)
    for _ in range(20):
        qi, qc = _masses(rng), _masses(rng)
        ni = np.array([_drv(qi, m) * v2n(m, DNOM[m]) for m in range(4)])
        nc = np.array([_drv(qc, m) * v2n(m, DNOM[m]) for m in range(4)])
        if mode == AIT:
            # interstitial Aitken particles 1.5x the largest Aitken size, but
            # so many small cloud-borne ones that the pair stays below the
            # Aitken -> accumulation transfer threshold
            d_i = 1.5 * DMAX[AIT]
            qc[AIT] *= 20
            nc[AIT] = _drv(qc, AIT) * v2n(AIT, 0.7 * DNOM[AIT])
        else:
            # interstitial accumulation particles 0.6x the smallest size, but
            # so few large cloud-borne ones that no accumulation -> Aitken
            # transfer happens
            d_i = 0.6 * DMIN[ACC]
            qc[ACC] *= 50
            nc[ACC] = _drv(qc, ACC) * v2n(ACC, 0.9 * DMAX[ACC])
        ni[mode] = _drv(qi, mode) * v2n(mode, d_i)
        a = orc.calcsize_level(DT, ni, nc, qi, qc, e3sm=False)
        b = orc.calcsize_level(DT, ni, nc, qi, qc, e3sm=True)
        # CalcSize clips to the mode's bound; the E3SM calcsize keeps the size
        assert a[4][mode] == (DMAX[AIT] if mode == AIT else DMIN[ACC])
        assert b[4][mode] == pytest.approx(dgn(mode, _drv(qi, mode) / ni[mode]), rel=1e-13)
        assert b[4][mode] == pytest.approx(d_i, rel=1e-12)
        # the other outputs are unaffected
        for k in (0, 1, 2, 3, 5):
            np.testing.assert_array_equal(a[k], b[k])


def test_noxf_number_uses_num2vol_ratio_max(orc):
    rng = np.random.default_rng(14)
    n_checked = 0
    for _ in range(30):
        qi, qc = _masses(rng), _masses(rng)
        ni = np.array([_drv(qi, m) * v2n(m, DNOM[m]) for m in range(4)])
        nc = np.array([_drv(qc, m) * v2n(m, DNOM[m]) for m in range(4)])
        # small accumulation particles: an accumulation -> Aitken transfer
        d = rng.uniform(0.3, 0.42) * DNOM[ACC]
        ni[ACC] = _drv(qi, ACC) * v2n(ACC, d)
        nc[ACC] = _drv(qc, ACC) * v2n(ACC, d)
        dn_i, dn_c, dq_i, dq_c, _, _ = orc.calcsize_level(DT, ni, nc, qi, qc, e3sm=True)
        # closed form of compute_coef_acc_ait_transfer with num2vol_ratio_max
        noxf = [1, 3, 4]  # POM, BC, DST: not in the Aitken mode
        drv_noxf = sum((qi[ACC, s] + qc[ACC, s]) / DENS[ACC][s] for s in noxf)
        v2n_ait, v2n_acc = v2n(AIT, DNOM[AIT]), v2n(ACC, DNOM[ACC])
        gm = math.sqrt(v2n_ait) * math.sqrt(v2n_acc)
        num_t0 = ni[ACC] + nc[ACC]
        drv_t0 = _drv(qi, ACC) + _drv(qc, ACC)
        if not num_t0 > drv_t0 * gm:
            continue
        num_t = max(0.0, num_t0 - drv_noxf * v2n(ACC, DMIN[ACC]))
        drv_t = drv_t0 - drv_noxf
        if not num_t > drv_t * gm:
            continue
        fv = 1.0 if num_t > drv_t * v2n_ait else (num_t / drv_t - gm) / (v2n_ait - gm)
        fn = 1.0 if fv == 1.0 else fv * (drv_t * v2n_ait / num_t)
        fn, fv = (1.0, 1.0) if fn >= 1 or fv >= 1 else (fn, fv)
        fn = fn * num_t / num_t0
        tsinv = 1.0 / 86400.0
        # the accumulation number loses its share to the Aitken mode
        assert -dn_i[ACC] == pytest.approx(ni[ACC] * fn * tsinv, rel=1e-12)
        assert dn_i[AIT] == pytest.approx(ni[ACC] * fn * tsinv, rel=1e-12)
        assert -dn_c[ACC] == pytest.approx(nc[ACC] * fn * tsinv, rel=1e-12)
        # transferred species move mass between the modes, conserved
        for sa, sc in ((0, 0), (1, 2), (2, 5), (3, 6)):
            assert dq_i[AIT, sa] == pytest.approx(-dq_i[ACC, sc], rel=1e-14)
            assert -dq_i[ACC, sc] == pytest.approx(qi[ACC, sc] * fv * tsinv, rel=1e-12)
        for s in noxf:
            assert dq_i[ACC, s] == 0.0
        n_checked += 1
    assert n_checked >= 10


def test_wet_size_composes_calcsize_and_water_uptake(orc):
    from mam4xx_amd import scenes
    sc = scenes.generate(6, 72, seed=15)
    dgi, dgc, dgw, qw, wd, tend = orc.wet_size(sc, DT)
    st = sc["state"]
    for k in range(0, 72, 7):
        for i in range(6):
            ni, nc = st[0:4, k, i], st[4:8, k, i]
            qi, qc = np.zeros((4, 7)), np.zeros((4, 7))
            off = [0, 7, 11, 18]
            for m in range(4):
                qi[m, :NSPEC[m]] = st[8 + off[m]:8 + off[m] + NSPEC[m], k, i]
                qc[m, :NSPEC[m]] = st[29 + off[m]:29 + off[m] + NSPEC[m], k, i]
            dn_i, dn_c, dq_i, dq_c, di, dc = orc.calcsize_level(DT, ni, nc, qi, qc, e3sm=True)
            w = orc.water_uptake_cell(qi, sc["temperature"][k, i], sc["pressure"][k, i],
                                      sc["qv"][k, i], sc["cloud_fraction"][k, i], di)
            assert np.array_equal(dgi[:, k, i], di) and np.array_equal(dgc[:, k, i], dc)
            for x, y in zip((dgw, qw, wd), w):
                assert np.array_equal(x[:, k, i], y)
            assert np.array_equal(tend[0:4, k, i], dn_i) and np.array_equal(tend[4:8, k, i], dn_c)
            for m in range(4):
                s0 = off[m]
                assert np.array_equal(tend[8 + s0:8 + s0 + NSPEC[m], k, i], dq_i[m, :NSPEC[m]])
                assert np.array_equal(tend[29 + s0:29 + s0 + NSPEC[m], k, i], dq_c[m, :NSPEC[m]])
            assert tend[50, k, i] == 0.0 and tend[51, k, i] == 0.0
