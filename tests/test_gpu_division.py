"""The SOA sub-step's division (scalar.hpp, num::div) against IEEE a / b on
the GPU, bit for bit, over 2 x 2^24 operand pairs spanning the ranges the
sub-step feeds it (tests/cpp/div_check.hip).  Numerator zeros are +0: a -0
numerator gives +0 where IEEE gives -0, and the sub-step's numerators are
sums and products of non-negative values clipped with max(x, 0), never -0."""
import ctypes
import os

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libdivcheck.so")


def test_division_is_correctly_rounded():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    torch.cuda.init()
    if not os.path.exists(LIB):
        pytest.fail("%s not built: run __graft_entry__.build()" % LIB)
    lib = ctypes.CDLL(LIB)
    lib.div_check.restype = ctypes.c_longlong
    lib.div_check.argtypes = [ctypes.c_long, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_double)]
    first = (ctypes.c_double * 2)()
    for seed in (1, 0x5eed):
        bad = lib.div_check(1 << 24, seed, first)
        assert bad == 0, (bad, first[0], first[1])
