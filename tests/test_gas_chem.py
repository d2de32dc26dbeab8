"""The oracle's gas-phase chemistry (gas_phase_chemistry.hpp with the
pp_linoz_mam4_resus_mom_soag mechanism of gas_chem_mechanism.hpp and the
implicit solver of gas_chem.hpp), pinned on the CPU:

* the rate constants at 298 K and 1 atm against the JPL evaluation's
  recommended values (OH + H2O2, OH + SO2 in its termolecular form, OH + DMS
  abstraction, NO3 + DMS, HO2 + HO2 with and without water);
* the solver: the mechanism is linear in the implicit species, so one
  Newton step is the backward-Euler step; the result equals numpy's dense
  solve of (I/dt - A) y = y0/dt + P built independently from the
  mechanism's reactions, and an isolated washed-out tracer follows the
  closed form (y0 + P dt) / (1 + k dt);
* species outside the mechanism's coupling keep their mass without
  washout or forcing.
"""
import math

import numpy as np
import pytest

M298 = 101325.0 / (1.380649e-23 * 298.0) * 1e-6  # air number density at 1 atm [#/cm3]


def _inv(M, h2o=0.0, oh=1e6, no3=1e7, ho2=1e8):
    return np.array([M, 0.78 * M, 0.21 * M, h2o, oh, no3, ho2, 0.0])


def test_rate_constants_against_jpl(orc):
    T, M = 298.0, M298
    r = orc.gc_rates(T, _inv(M))
    k = lambda i, c: r[i] / c
    # JPL 15-10 recommended 298 K values (cm3/molecule/s)
    assert k(2, 1e6) == pytest.approx(1.7e-12, rel=0.1)         # OH + H2O2
    k0, kinf = 3.3e-31 * (T / 300) ** -4.3 * M, 1.6e-12       # OH + SO2 (+M)
    troe = k0 / (1 + k0 / kinf) * 0.6 ** (1 / (1 + math.log10(k0 / kinf) ** 2))
    assert k(3, 1e6) == pytest.approx(troe, rel=0.15)
    assert k(4, 1e6) == pytest.approx(1.2e-11 * math.exp(-280 / T), rel=0.1)  # OH + DMS
    assert k(6, 1e7) == pytest.approx(1.9e-13 * math.exp(530 / T), rel=0.05)  # NO3 + DMS
    # HO2 + HO2: bimolecular + termolecular parts, enhanced by water vapour
    dry = 3.0e-13 * math.exp(460 / T) + 2.1e-33 * M * math.exp(920 / T)
    assert r[1] / (1e8 ** 2 / M) == pytest.approx(dry, rel=0.2)
    wet = orc.gc_rates(T, _inv(M, h2o=0.01 * M))[1] / r[1]
    assert wet == pytest.approx(1 + 1.4e-21 * 0.01 * M * math.exp(2200 / T), rel=1e-12)


def _case(rng):
    M = M298 * rng.uniform(0.3, 1.0)
    inv = _inv(M, h2o=rng.uniform(0, 0.02) * M, oh=10 ** rng.uniform(5, 7),
               no3=10 ** rng.uniform(5, 8), ho2=10 ** rng.uniform(7, 9))
    het = np.zeros(31)
    het[1:] = 10 ** rng.uniform(-7, -3, 30) * (rng.uniform(size=30) < 0.7)
    ext = 10 ** rng.uniform(-4, 0, 9) / M  # [1/s] after scaling by M
    y0 = np.zeros(31)
    y0[1:] = 10 ** rng.uniform(-12, -8, 30)
    return M, inv, het, ext, y0


# (species, external forcing entry): SO2, SOAG, so4_a1, num_a1, so4_a2,
# num_a2, pom_a4, bc_a4, num_a4 (indprd, gas_chem_mechanism.hpp:100-143)
FORCED = ((3, 0), (5, 8), (6, 1), (13, 5), (14, 2), (18, 6), (27, 3), (28, 4), (30, 7))


def _backward_euler(rxt, het, ext, y0, dt):
    """(I/dt - A) y = y0/dt + P over species 1..30, from the reactions:
    H2O2 lost by photolysis (rxt0) and OH (rxt2), made by HO2 + HO2 (rxt1);
    SO2 -> H2SO4 by OH (rxt3); DMS -> SO2 by OH (rxt4, rxt5 at half yield)
    and NO3 (rxt6); every species washed out at its het rate; external
    forcing into SO2, DMS, SOAG and six aerosol tracers."""
    n = 30
    A = np.zeros((n, n))
    P = np.zeros(n)
    s = lambda j: j - 1  # gas_pcnst index -> class index
    for j in range(1, 31):
        A[s(j), s(j)] -= het[j]
    A[s(1), s(1)] -= rxt[0] + rxt[2]
    P[s(1)] += rxt[1]
    A[s(3), s(3)] -= rxt[3]
    A[s(2), s(3)] += rxt[3]
    A[s(4), s(4)] -= rxt[4] + rxt[5] + rxt[6]
    A[s(3), s(4)] += rxt[4] + 0.5 * rxt[5] + rxt[6]
    for j, e in FORCED:
        P[s(j)] += ext[e]
    y = np.linalg.solve(np.eye(n) / dt - A, y0[1:] / dt + P)
    return y


@pytest.mark.parametrize("dt", [60.0, 1800.0, 3600.0])
def test_solver_is_the_backward_euler_step(orc, dt):
    rng = np.random.default_rng(int(dt))
    for _ in range(40):
        M, inv, het, ext, y0 = _case(rng)
        rxt = orc.gc_rates(rng.uniform(220, 300), inv)
        rxt[0] = 10 ** rng.uniform(-7, -5)  # photolysis of H2O2
        y, nstep = orc.gc_imp_sol(y0, rxt, het, ext, dt)
        assert nstep == 1
        ref = _backward_euler(rxt, het, ext, y0, dt)
        np.testing.assert_allclose(y[1:], ref, rtol=1e-12, atol=0)
        assert y[0] == y0[0]  # O3 is not in the implicit class


def test_isolated_tracer_closed_form(orc):
    rng = np.random.default_rng(9)
    M, inv, het, ext, y0 = _case(rng)
    rxt = orc.gc_rates(280.0, inv)
    rxt[0] = 1e-6
    dt = 1800.0
    y, _ = orc.gc_imp_sol(y0, rxt, het, ext, dt)
    for j, e in FORCED[1:]:
        assert y[j] == pytest.approx((y0[j] + ext[e] * dt) / (1 + het[j] * dt), rel=1e-13)
    forced = {j for j, _ in FORCED}
    for j in range(5, 31):
        if j not in forced:
            assert y[j] == pytest.approx(y0[j] / (1 + het[j] * dt), rel=1e-13)
    # no washout, no forcing: aerosol tracers keep their mass exactly
    y2, _ = orc.gc_imp_sol(y0, rxt, np.zeros(31), np.zeros(9), dt)
    assert np.array_equal(y2[6:], y0[6:])


def test_domain_driver_scales_forcing_by_density(orc):
    rng = np.random.default_rng(10)
    n = 50
    temp = rng.uniform(220, 300, n)
    photo = 10 ** rng.uniform(-7, -5, n)
    inv = np.stack([_case(rng)[1] for _ in range(n)], axis=1)
    extfrc = 10 ** rng.uniform(1, 5, (9, n))  # [1/cm3/s]
    het = 10 ** rng.uniform(-7, -4, (31, n))
    qq = 10 ** rng.uniform(-12, -8, (31, n))
    out, steps = orc.gas_phase_chemistry(1800.0, temp, photo, extfrc, inv, het, qq)
    for c in range(0, n, 7):
        rxt = orc.gc_rates(temp[c], inv[:, c])
        rxt[0] = photo[c]
        y, _ = orc.gc_imp_sol(qq[:, c], rxt, het[:, c], extfrc[:, c] / inv[0, c], 1800.0)
        assert np.array_equal(out[:, c], y)
    assert (steps == 1).all()


def _domain(rng, n):
    temp = rng.uniform(200, 310, n)
    photo = 10 ** rng.uniform(-8, -4, n)
    inv = np.stack([_case(rng)[1] for _ in range(n)], axis=1)
    extfrc = 10 ** rng.uniform(0, 6, (9, n)) * (rng.uniform(size=(9, n)) < 0.8)
    het = 10 ** rng.uniform(-8, -2, (31, n)) * (rng.uniform(size=(31, n)) < 0.7)
    qq = 10 ** rng.uniform(-14, -7, (31, n))
    qq[:, :10] = 0.0  # empty cells
    return temp, photo, extfrc, inv, het, qq


def test_kernel_code_bitwise_equals_oracle(orc, tmp_path):
    """mam4xx_amd/csrc/gas_chem.hpp built for the CPU (g++, glibc) against
    the oracle: bit for bit, steps included."""
    import ctypes
    import os
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    so = str(tmp_path / "libgchost.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o",
                    so, os.path.join(here, "cpp", "gc_host.cpp")], check=True)
    lib = ctypes.CDLL(so)
    rng = np.random.default_rng(12)
    n = 3000
    temp, photo, extfrc, inv, het, qq = _domain(rng, n)
    for dt in (1.0, 1800.0):
        ref, steps = orc.gas_phase_chemistry(dt, temp, photo, extfrc, inv, het, qq)
        dp = ctypes.POINTER(ctypes.c_double)
        p = lambda a: a.ctypes.data_as(dp)
        q = qq.copy()
        st = np.zeros(n, dtype=np.int32)
        lib.gc_host_run.restype = None
        lib.gc_host_run(ctypes.c_long(n), ctypes.c_double(dt), p(temp), p(photo), p(extfrc),
                        p(inv), p(het), p(q), st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        assert np.array_equal(st, steps)
        assert np.array_equal(q, ref), np.argwhere(q != ref)[:5]
        assert (ref[1:] != qq[1:]).sum() > n
