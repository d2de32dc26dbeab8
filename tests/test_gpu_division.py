"""The SOA sub-step's fast division (scalar.hpp, num::div_fast) against IEEE
a / b on the GPU, bit for bit (tests/cpp/div_check.hip): 2 x 2^24 operand
pairs spanning the ranges the sub-step feeds it, and 2 x 2^24 pairs whose
numerators lie in 2^-1074..2^-960, either side of the 2^-968 bound.  Inside
the range the division is exact on (numerator 0 or >= 2^-968, quotient
normal) no pair may differ; outside it some do, which is why the sub-step
checks that its operands stay inside (chain_device.hpp, soa_fast_exact).
Numerator zeros are +0: a -0 numerator gives +0 where IEEE gives -0, and the
sub-step's numerators are sums and products of non-negative values clipped
with max(x, 0), never -0."""
import ctypes
import os

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libdivcheck.so")


@pytest.mark.parametrize("tiny", [0, 1], ids=["sub-step-range", "tiny-numerators"])
def test_division_is_correctly_rounded(tiny):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    torch.cuda.init()
    if not os.path.exists(LIB):
        pytest.fail("%s not built: run __graft_entry__.build()" % LIB)
    lib = ctypes.CDLL(LIB)
    lib.div_check.restype = ctypes.c_longlong
    lib.div_check.argtypes = [ctypes.c_long, ctypes.c_ulonglong, ctypes.c_int,
                              ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_ulonglong)]
    first = (ctypes.c_double * 2)()
    outside = (ctypes.c_ulonglong * 2)()
    for seed in (1, 0x5eed):
        bad = lib.div_check(1 << 24, seed, tiny, first, outside)
        print("tiny %d seed %d: inside mismatches %d, outside %d of %d pairs"
              % (tiny, seed, bad, outside[0], outside[1]))
        assert bad == 0, (bad, first[0], first[1])
        if tiny:
            # the bound is where exactness ends: a few percent of the pairs
            # below it round differently
            assert outside[1] > (1 << 22) and outside[0] > 1000
        else:
            assert outside[1] == 0
