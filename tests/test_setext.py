"""External forcing (mo_setext.hpp) of the oracle: each forced species'
rate is the sum of its sectors' data, in sector order, with bottom-up data
read flipped, written to its 1-based destination entry."""
import numpy as np


def sectors(rng, nlev, ncol):
    return [[10 ** rng.uniform(-2, 4, (nlev, ncol)) for _ in range(rng.integers(0, 5))]
            for _ in range(9)]


def test_sum_of_sectors(orc):
    rng = np.random.default_rng(1)
    nlev, ncol = 30, 7
    sec = sectors(rng, nlev, ncol)
    sec[0] = [10 ** rng.uniform(-2, 4, (nlev, ncol))] * 2  # at least one non-empty
    alt = [m % 3 == 0 for m in range(9)]
    ndx = list(rng.permutation(9) + 1)
    got = orc.setext(sec, ndx, alt)
    for m in range(9):
        want = np.zeros((nlev, ncol))
        for x in sec[m]:
            want = want + (x[::-1] if alt[m] else x)
        np.testing.assert_array_equal(got[ndx[m] - 1], want)
