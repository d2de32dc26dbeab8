"""Column budgets of a step's tendencies (mam4x_outputs.column_budgets),
restated on the CPU for the tests.

* ``exact``: the mathematically exact column integrals, sum_k x(k) dp(k)/g
  formed with math.fsum, and the sum of |terms| (the scale of a column's
  rounding error and of a closure's cancellation);
* ``ordered``: the library's documented operation order, with fused
  multiply-adds emulated exactly (Fraction arithmetic, one rounding), so a
  GPU result can be compared bit for bit.

The closures follow include/mam4x.h: sulfur = H2SO4 gas and the SO4 fields
at 1000 / molecular weight [mol/kg]; organic = SOAG and the SOA fields in the
gas-aerosol exchange's basis (mam4_amicphys.hpp:131-154: SOAG * 12.011/150,
SOA * 0.08, both per 12.011 g/mol)."""
import math
from fractions import Fraction

import numpy as np

from mam4xx_amd import layout as L

GRAVITY = 9.80616  # mam4_constants.hpp:48
INV_GRAVITY = 1.0 / GRAVITY
NBUDGET = L.NSTATE + 2
SULFUR, ORGANIC = L.NSTATE, L.NSTATE + 1
MW_SO4, MW_SOA, MW_H2SO4, MW_SOAG = 115.10734, 12.011, 98.0784, 12.011  # adv_mass


def closures():
    """[(fields, mol/kg factors)] for the sulfur and organic closures, the
    fields in state order."""
    sul, org = [], []
    for f in range(L.NSTATE):
        if not L.is_mass(f):
            continue
        q = f - (29 if f >= 29 else 8)
        m = max(i for i in range(4) if L.MODE_OFFSET[i] <= q)
        aid = L.MODE_SPECIES[m][q - L.MODE_OFFSET[m]]
        if aid == 1:  # SO4
            sul.append((f, 1000.0 / MW_SO4))
        elif aid == 0:  # SOA
            org.append((f, 1000.0 * 0.08 / MW_SOA))
    sul.append((L.F_H2SO4, 1000.0 / MW_H2SO4))
    org.append((L.F_SOAG, 1000.0 * (12.011 / 150.0) / MW_SOAG))
    return sul, org


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def exact(tend, dp):
    """(budgets [NBUDGET, ncol], gross [NBUDGET, ncol]) of tendencies
    [52, nlev, ncol] and layer thicknesses [nlev, ncol], summed exactly."""
    nlev, ncol = dp.shape
    w = dp * INV_GRAVITY
    out = np.zeros((NBUDGET, ncol))
    gross = np.zeros((NBUDGET, ncol))
    terms = [tend[f] for f in range(L.NSTATE)]
    for cl in closures():
        terms.append(sum(tend[f] * c for f, c in cl))
    for b, x in enumerate(terms):
        t = x * w
        for c in range(ncol):
            out[b, c] = math.fsum(t[:, c])
            gross[b, c] = math.fsum(np.abs(t[:, c]))
    # a closure's scale is its fields' gross, not that of the summed term
    for b, cl in ((SULFUR, closures()[0]), (ORGANIC, closures()[1])):
        g = sum(np.abs(tend[f] * c) for f, c in cl) * w
        gross[b] = [math.fsum(g[:, c]) for c in range(ncol)]
    return out, gross


def ordered(tend, dp, cols):
    """The library's result for columns `cols`, bit for bit: partial l of
    the levels k = l, l + 8, ... by fused multiply-adds in ascending k (weight
    dp * (1/g)), partials combined as ((p0+p4)+(p2+p6))+((p1+p5)+(p3+p7));
    closure terms formed over their fields in state order by fused
    multiply-adds before the weight."""
    nlev = dp.shape[0]
    cl = closures()
    out = np.zeros((NBUDGET, len(cols)))
    for j, c in enumerate(cols):
        for b in range(NBUDGET):
            p = [0.0] * 8
            for k in range(nlev):
                if b < L.NSTATE:
                    x = float(tend[b, k, c])
                else:
                    x = 0.0
                    for f, coef in cl[b - L.NSTATE]:
                        x = _fma(float(tend[f, k, c]), coef, x)
                p[k % 8] = _fma(x, float(dp[k, c]) * INV_GRAVITY, p[k % 8])
            out[b, j] = ((p[0] + p[4]) + (p[2] + p[6])) + ((p[1] + p[5]) + (p[3] + p[7]))
    return out
