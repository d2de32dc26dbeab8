"""The allocation-free step (mam4x_compute_tendencies_noalloc) against
mam4x_compute_tendencies on the same handle, through the C ABI.

The work-queue schedule is forced, so the step's workspace lives in the
library's fixed device array.  Two calls in a row, the second on a smaller
domain than the first (the array keeps the first call's contents), must give
the existing entry point's tendencies, diameters and updated state bit for
bit: the two differ only in where the workspace lives.  A domain above the
limit is rejected before anything is launched.
"""
import numpy as np
import pytest

from mam4xx_amd import layout as L
from mam4xx_amd import scenes

pytestmark = pytest.mark.gpu

DT = 1800.0


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


def _run(torch, ch, sc, noalloc):
    from mam4xx_amd.chain import scene_to_device
    nlev, ncol = sc["temperature"].shape
    d = scene_to_device(sc)
    tend = torch.full((L.NSTATE, nlev, ncol), np.nan, dtype=torch.float64, device="cuda")
    dgi = torch.full((4, nlev, ncol), np.nan, dtype=torch.float64, device="cuda")
    dgc = torch.full((4, nlev, ncol), np.nan, dtype=torch.float64, device="cuda")
    post = torch.full((L.NSTATE, nlev, ncol), np.nan, dtype=torch.float64, device="cuda")
    ch.compute_tendencies(d, DT, tend, dgn_i=dgi, dgn_c=dgc, state_out=post, noalloc=noalloc)
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in (tend, dgi, dgc, post)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_equals_compute_tendencies(torch_cuda, precision):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    prec = {"fp64": _abi.PREC_FP64, "mixed": _abi.PREC_MIXED}[precision]
    ch = Chain(precision=prec, schedule=_abi.SCHED_QUEUE)
    big = scenes.generate(7200, 72, seed=21)     # 518,400 cells, just under the limit
    small = scenes.generate(700, 72, seed=22)    # 50,400 cells
    assert 7200 * 72 <= _abi.NOALLOC_MAX_CELLS
    for name, sc in (("7200x72", big), ("700x72", small)):
        print("noalloc vs compute_tendencies, %s, %s" % (precision, name), flush=True)
        na = _run(torch_cuda, ch, sc, noalloc=True)
        ref = _run(torch_cuda, ch, sc, noalloc=False)
        for a, b in zip(na, ref):
            assert np.all(np.isfinite(a))
            assert np.array_equal(a, b)
    ch.close()


def test_noalloc_rejects_domains_above_the_limit(torch_cuda):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, ChainError
    ch = Chain(schedule=_abi.SCHED_QUEUE)
    sc = scenes.generate(_abi.NOALLOC_MAX_CELLS // 72 + 1, 72, seed=23)
    with pytest.raises(ChainError, match="allocation-free limit"):
        _run(torch_cuda, ch, sc, noalloc=True)
    ch.close()
