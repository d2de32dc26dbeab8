"""Column budgets on the CPU: the restatement of the library's order against
the exact sums, and the two closures on the oracle's step.

* the ordered sum (8 strided partials, fixed tree) stays within a few ulps
  of the column's gross of the exact sum, and differs from a plain ascending
  sum where the order matters (so a test comparing bits would notice an
  ascending-sum implementation);
* the sulfur and organic closures of the oracle's step vanish to 1e-12 of
  their gross: the chain moves H2SO4 into SO4 and SOAG into SOA mole for
  mole (in the exchange's basis) with no source or sink, and renaming,
  coagulation, aging and calcsize move mass between modes only.  With gas
  chemistry upstream too: its production is already in the current state.
  A closure with a wrong molecular weight or a dropped field fails this."""
import math

import numpy as np
import pytest

from mam4xx_amd import layout as L
from mam4xx_amd import scenes

from . import budget_ref as B


def test_closure_fields():
    sul, org = B.closures()
    names = L.field_names()
    assert [names[f] for f, _ in sul] == ["so4_acc_i", "so4_ait_i", "so4_coarse_i", "so4_acc_c",
                                         "so4_ait_c", "so4_coarse_c", "h2so4"]
    assert [names[f] for f, _ in org] == ["soa_acc_i", "soa_ait_i", "soa_coarse_i", "soa_acc_c",
                                         "soa_ait_c", "soa_coarse_c", "soag"]
    # SOAG and SOA are exchanged at 12.011/150 and 0.08 per 12.011 g/mol
    assert math.isclose(org[-1][1], 1000.0 / 150.0, rel_tol=1e-15)
    assert math.isclose(org[0][1], 80.0 / 12.011, rel_tol=1e-15)


def test_ordered_sum_is_the_exact_sum_to_rounding():
    rng = np.random.default_rng(0)
    nlev, ncol = 72, 6
    tend = rng.standard_normal((L.NSTATE, nlev, ncol)) * 10.0 ** rng.uniform(-20, -5,
                                                                              (L.NSTATE, 1, 1))
    dp = rng.uniform(10.0, 2000.0, (nlev, ncol))
    ex, gross = B.exact(tend, dp)
    got = B.ordered(tend, dp, range(ncol))
    assert (np.abs(got - ex) <= 1e-14 * gross).all()
    # ascending-k sums differ from the tree in the last bits somewhere
    asc = np.array([[sum(float(tend[b, k, c]) * float(dp[k, c]) * B.INV_GRAVITY
                         for k in range(nlev)) for c in range(ncol)] for b in range(L.NSTATE)])
    assert not np.array_equal(asc, got[:L.NSTATE])


@pytest.mark.parametrize("seed,pregas", [(3, False), (8, False), (5, True)])
def test_closures_vanish_on_the_oracle_step(orc, seed, pregas):
    sc = scenes.generate(48, 72, seed=seed)
    pre = None
    if pregas:  # gas chemistry upstream: H2SO4 and SOAG before it, 10-90 % of now
        rng = np.random.default_rng(seed)
        frac = rng.uniform(0.1, 0.9, (2,) + sc["temperature"].shape)
        pre = sc["state"][[L.F_H2SO4, L.F_SOAG]] * frac
    tend = orc.chain_step(sc, 1800.0, pregaschem_gas=pre)[0]
    ex, gross = B.exact(tend, sc["hydrostatic_dp"])
    for b in (B.SULFUR, B.ORGANIC):
        assert (gross[b] > 0).all()
        rel = np.abs(ex[b]) / gross[b]
        assert rel.max() < 1e-12, (b, rel.max())
    # a closure built with H2SO4 at SO4's molecular weight does not vanish
    bad = sum(tend[f] * c for f, c in B.closures()[0][:-1]) + tend[L.F_H2SO4] * 1000.0 / B.MW_SO4
    badcol = (bad * sc["hydrostatic_dp"]).sum(0)
    assert (np.abs(badcol) / gross[B.SULFUR]).max() > 1e-3


def test_generated_layer_thicknesses_span_the_column():
    sc = scenes.generate(5, 72, seed=1)
    dp, p = sc["hydrostatic_dp"], sc["pressure"]
    assert (dp > 0).all()
    # interfaces run from ptop = 100 Pa to the surface pressure, which the
    # lowest mid level fixes: p[-1] = ptop (ps/ptop)^((nlev - 1/2)/nlev)
    ps = 100.0 * (p[-1] / 100.0) ** (72 / 71.5)
    assert np.allclose(100.0 + dp.sum(0), ps, rtol=1e-12, atol=0)
    bottom = 100.0 + np.cumsum(dp, axis=0)  # lower interface of each layer
    assert (p < bottom).all() and (p[1:] > bottom[:-1]).all()
