"""mam4x_compute_setsox (aqueous sulfur chemistry, mo_setsox.hpp) on the GPU
against the oracle's orc_setsox, pinned in test_setsox.py (whose host build
of the kernel's cell code equals the oracle bit for bit).

Bar: each updated tracer within 1e-12 of its magnitude in the cell (the
larger of its value before and after), the tendencies within 1e-12 of the
cell's cloud-borne sulfate over dt -- or within 10x the oracle's own spread
where single-ulp libm perturbations move the oracle by more than 1e-13 (a
pH bisection step that lands within rounding of the root flips the bracket,
a 0.5 % move of [H+]).  Layouts agree bit for bit."""
import numpy as np
import pytest

from .test_setsox import SO4_CW, _domain

pytestmark = pytest.mark.gpu

DT = 1800.0
TOL = 1e-12


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


MET = {"pressure": "press", "temperature": "tfld", "liquid_water": "lwc",
       "cloud_fraction": "cldfrc", "droplet_number": "cldnum", "air_number_density": "xhnm"}


def _gpu(torch, met, qcw, qin, nlev, ncol, layout=None, outputs=True):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    lay = _abi.LAYOUT_LEV_COL if layout is None else layout
    tr = (lambda a: np.ascontiguousarray(np.swapaxes(a, -1, -2))) if lay == _abi.LAYOUT_COL_LEV \
        else (lambda a: a)
    sh = lambda a: tr(np.asarray(a).reshape(a.shape[:-1] + (nlev, ncol)))
    dm = {k: torch.as_tensor(sh(met[v])).cuda() for k, v in MET.items()}
    qc, qi = torch.as_tensor(sh(qcw)).cuda(), torch.as_tensor(sh(qin)).cuda()
    d1 = torch.full(qc.shape, np.nan, dtype=torch.float64, device="cuda") if outputs else None
    d2 = torch.full(qc.shape, np.nan, dtype=torch.float64, device="cuda") if outputs else None
    ch = Chain()
    ch.compute_setsox(DT, dm, qc, qi, d1, d2, layout=lay)
    torch.cuda.synchronize()
    ch.close()
    back = lambda t: tr(t.cpu().numpy()).reshape(31, -1) if t is not None else None
    return back(qc), back(qi), back(d1), back(d2)


def test_setsox_matches_oracle(orc, torch_cuda):
    rng = np.random.default_rng(11)
    nlev, ncol = 40, 50
    met, qcw, qin = _domain(rng, nlev * ncol)
    g = _gpu(torch_cuda, met, qcw, qin, nlev, ncol)
    ref = orc.setsox(DT, met, qcw, qin)
    env = [np.zeros_like(r) for r in ref]
    try:
        for s in range(1, 7):
            orc.set_libm_noise(s)
            for e, x, r in zip(env, orc.setsox(DT, met, qcw, qin), ref):
                np.maximum(e, np.abs(x - r), out=e)
    finally:
        orc.set_libm_noise(0)
    # magnitudes: per tracer the larger of before / after; the tendencies
    # against the cell's cloud-borne sulfate over dt
    so4 = sum(np.maximum(qcw[i], ref[0][i]) for i in SO4_CW)
    scales = [np.maximum(np.abs(qcw), np.abs(ref[0])), np.maximum(np.abs(qin), np.abs(ref[1])),
              np.broadcast_to(so4 / DT, qcw.shape), np.broadcast_to(so4 / DT, qcw.shape)]
    cell_env = np.max([(e / np.maximum(sc, 1e-300)).max(axis=0)
                       for e, sc in zip(env, scales)], axis=0)
    bound = np.where(cell_env > 1e-13, np.maximum(TOL, 10 * cell_env), TOL)
    for name, x, r, sc in zip(("qcw", "qin", "dqdt_aqso4", "dqdt_aqh2so4"), g, ref, scales):
        assert np.isfinite(x).all(), name
        err = np.abs(x - r) / np.maximum(sc, 1e-300)
        bad = np.argwhere(err > bound[None])
        assert len(bad) == 0, (name, len(bad), bad[0], err[tuple(bad[0])])
    assert (cell_env > 1e-13).mean() < 0.05
    assert np.count_nonzero(ref[2]) > 1000  # chemistry happened


def test_layouts_and_optional_outputs(torch_cuda):
    from mam4xx_amd import _abi
    rng = np.random.default_rng(12)
    nlev, ncol = 24, 70
    met, qcw, qin = _domain(rng, nlev * ncol)
    a = _gpu(torch_cuda, met, qcw, qin, nlev, ncol)
    b = _gpu(torch_cuda, met, qcw, qin, nlev, ncol, layout=_abi.LAYOUT_COL_LEV)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    c = _gpu(torch_cuda, met, qcw, qin, nlev, ncol, outputs=False)
    assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])
    # untouched tracers keep their values; the tendency planes given are
    # zero outside the sulfate slots
    keep_c = [i for i in range(31) if i not in SO4_CW]
    assert np.array_equal(a[0][keep_c], qcw[keep_c])
    assert np.array_equal(a[1][[0] + list(range(4, 31))], qin[[0] + list(range(4, 31))])
    for d in a[2:]:
        assert not d[keep_c].any()


def test_missing_plane_is_an_error(torch_cuda):
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, ChainError
    ch = Chain()
    a = _abi.SetsoxArgs()
    a.ncol, a.nlev = 4, 3
    with pytest.raises(ChainError, match="NULL"):
        ch._check(ch.lib.mam4x_compute_setsox(ch.h, DT, a, None), "mam4x_compute_setsox")
    ch.close()
