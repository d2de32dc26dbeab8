"""mam4x_compute_washout on the GPU against the oracle's orc_sethet_domain,
pinned in test_sethet.py (whose host build of the kernel's column code
equals the oracle bit for bit).

Bar: every rate within 1e-12 of its value, or within 10x the oracle's own
spread under single-ulp perturbations of its exp / sqrt / pow where that
spread is larger (a level whose running washout sum lands within rounding
of its saturation concentration is washed in one and not the other).
Layouts agree bit for bit; planes other than H2O2, H2SO4 and SO2 are
zeroed."""
import numpy as np
import pytest

from .test_sethet import columns

pytestmark = pytest.mark.gpu

DT = 1800.0
PLANES = {"pressure": "press", "zmid": "zmid", "temperature": "tfld", "cmfdqr": "cmfdqr",
          "nrain": "nrain", "nevapr": "nevapr", "air_density": "inv_m", "h2o2": "h2o2"}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


def _gpu(torch, cols, col_lev=False):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    tr = (lambda a: np.ascontiguousarray(np.swapaxes(a, -1, -2))) if col_lev else (lambda a: a)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).cuda()
    nlev, ncol = cols["press"].shape
    het = dev(tr(np.full((31, nlev, ncol), np.nan)))
    ch = Chain()
    ch.compute_washout(DT, {k: dev(tr(cols[v])) for k, v in PLANES.items()}, dev(cols["rlat"]),
                       dev(cols["phis"]), het,
                       layout=_abi.LAYOUT_COL_LEV if col_lev else _abi.LAYOUT_LEV_COL)
    torch.cuda.synchronize()
    ch.close()
    return tr(het.cpu().numpy())


def test_washout_matches_oracle(orc, torch_cuda):
    rng = np.random.default_rng(31)
    cols = columns(rng, 72, 400)
    g = _gpu(torch_cuda, cols)
    ref = orc.sethet(DT, cols)
    env = np.zeros_like(ref)
    try:
        for s in range(1, 7):
            orc.set_libm_noise(s)
            env = np.maximum(env, np.abs(orc.sethet(DT, cols) - ref))
    finally:
        orc.set_libm_noise(0)
    assert np.isfinite(g).all()
    assert not g[[0] + list(range(4, 31))].any()
    bound = np.maximum(1e-12 * np.abs(ref), 10 * env)
    bad = np.argwhere(np.abs(g - ref) > bound)
    assert len(bad) == 0, (len(bad), bad[:3])
    assert (ref[1] > 0).sum() > 1000


def test_washout_layouts(torch_cuda):
    rng = np.random.default_rng(32)
    cols = columns(rng, 72, 100)
    assert np.array_equal(_gpu(torch_cuda, cols), _gpu(torch_cuda, cols, col_lev=True))
