"""Padded planes through the chemistry entry points: mam4x_compute_linoz (the
LDS block kernel up to 128 levels and the one for deeper columns),
mam4x_compute_washout, mam4x_compute_gas_chemistry and mam4x_compute_setsox,
each in both layouts with the leading dimension ldc larger than the fastest
extent (ncol for LAYOUT_LEV_COL, nlev for LAYOUT_COL_LEV), as a host passes
arrays dimensioned for its largest chunk.  The padding holds NaN on input
and output planes alike.

* the live entries equal those of the same call on packed planes bit for
  bit (the packed calls are held to the oracle in test_gpu_linoz.py,
  test_gpu_sethet.py, test_gpu_gas_chem.py and test_gpu_setsox.py);
* linoz and gas chemistry are also checked against the oracle directly, at
  those files' bars;
* no padding entry is written (it stays NaN), and no NaN reaches a live one.
(setinv and setext: test_gpu_chem_edges.py.)"""
import numpy as np
import pytest

from .test_gas_chem import _domain as gc_domain
from .test_gpu_gas_chem import _gpu as gc_packed
from .test_gpu_linoz import _gpu_linoz as linoz_packed
from .test_gpu_sethet import PLANES as SETHET_PLANES
from .test_gpu_sethet import _gpu as sethet_packed
from .test_gpu_setsox import MET as SOX_MET
from .test_gpu_setsox import _gpu as sox_packed
from .test_linoz import CONF, columns as linoz_columns
from .test_sethet import columns as sethet_columns
from .test_setsox import _domain as sox_domain

pytestmark = pytest.mark.gpu
DT = 1800.0
PAD = 5
LAYOUTS = ["lev_col", "col_lev"]


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


class Pad:
    """[..., nlev, ncol] arrays to padded CUDA planes of a layout and back."""

    def __init__(self, torch, layout):
        from mam4xx_amd import _abi
        self.torch = torch
        self.col_lev = layout == "col_lev"
        self.layout = _abi.LAYOUT_COL_LEV if self.col_lev else _abi.LAYOUT_LEV_COL

    def _tr(self, a):
        return np.ascontiguousarray(np.swapaxes(a, -1, -2)) if self.col_lev else a

    def live(self, nlev, ncol):
        return nlev if self.col_lev else ncol

    def dev(self, a):
        """padded CUDA tensor of the array in this layout (NaN padding)"""
        a = self._tr(np.asarray(a, dtype=np.float64))
        t = self.torch.full(a.shape[:-1] + (a.shape[-1] + PAD,), float("nan"),
                            dtype=self.torch.float64, device="cuda")
        t[..., :a.shape[-1]] = self.torch.as_tensor(a)
        return t

    def back(self, t):
        """the live entries as [..., nlev, ncol]; asserts the padding is NaN"""
        x = t.cpu().numpy()
        assert np.isnan(x[..., -PAD:]).all(), "padding written"
        return self._tr(x[..., :-PAD])


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("nlev", [72, 150])  # LDS block kernel / deeper columns
def test_linoz(orc, torch_cuda, layout, nlev):
    from mam4xx_amd.chain import Chain
    torch = torch_cuda
    P = Pad(torch, layout)
    rng = np.random.default_rng(60 + nlev)
    ncol = 77
    cols = linoz_columns(rng, nlev, ncol)
    col = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).cuda()
    o3 = P.dev(cols["o3"])
    outs = [P.dev(np.full((nlev, ncol), np.nan)) for _ in range(4)]
    sink = col(np.full(ncol, np.nan))
    ch = Chain()
    ch.compute_linoz(DT, P.dev(cols["temp"]), P.dev(cols["pmid"]), P.dev(cols["pdel"]),
                     P.dev(cols["clim"]), col(cols["sza"]), col(cols["rlats"]),
                     col(cols["ltrop"], torch.int32), o3, CONF,
                     column_above_top=col(cols["col0"]), o3_column=outs[0], do3_linoz=outs[1],
                     do3_linoz_psc=outs[2], ss_o3=outs[3], o3_sfcsink=sink, layout=P.layout,
                     live=P.live(nlev, ncol))
    torch.cuda.synchronize()
    ch.close()
    got = [P.back(o3)] + [P.back(t) for t in outs] + [sink.cpu().numpy()]
    ref = linoz_packed(torch, cols, col_lev=P.col_lev)
    for x, y in zip(got, ref):
        assert np.array_equal(x, y)
    r = orc.linoz(DT, cols, CONF)
    scale = np.maximum(np.abs(cols["o3"]), np.abs(r[0]))
    assert (np.abs(got[0] - r[0]) <= 1e-12 * scale).all()
    assert np.array_equal(got[1], r[1])


@pytest.mark.parametrize("layout", LAYOUTS)
def test_washout(torch_cuda, layout):
    from mam4xx_amd.chain import Chain
    torch = torch_cuda
    P = Pad(torch, layout)
    rng = np.random.default_rng(61)
    nlev, ncol = 72, 83
    cols = sethet_columns(rng, nlev, ncol)
    col = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).cuda()
    het = P.dev(np.full((31, nlev, ncol), np.nan))
    ch = Chain()
    ch.compute_washout(DT, {k: P.dev(cols[v]) for k, v in SETHET_PLANES.items()},
                       col(cols["rlat"]), col(cols["phis"]), het, layout=P.layout,
                       live=P.live(nlev, ncol))
    torch.cuda.synchronize()
    ch.close()
    got = P.back(het)
    assert np.isfinite(got).all()
    assert np.array_equal(got, sethet_packed(torch, cols, col_lev=P.col_lev))
    assert (got[1] > 0).sum() > 100


@pytest.mark.parametrize("layout", LAYOUTS)
def test_gas_chemistry(orc, torch_cuda, layout):
    from mam4xx_amd.chain import Chain
    torch = torch_cuda
    P = Pad(torch, layout)
    rng = np.random.default_rng(62)
    nlev, ncol = 24, 70
    f = gc_domain(rng, nlev * ncol)
    sh = lambda a: np.asarray(a).reshape(a.shape[:-1] + (nlev, ncol))
    temp, photo, extfrc, inv, het, qq = (sh(x) for x in f)
    q = P.dev(qq)
    ch = Chain()
    ch.compute_gas_chemistry(1800.0, P.dev(temp), P.dev(photo), P.dev(extfrc), P.dev(inv),
                             P.dev(het), q, layout=P.layout, live=P.live(nlev, ncol))
    torch.cuda.synchronize()
    ch.close()
    got = P.back(q).reshape(31, -1)
    assert np.array_equal(got, gc_packed(torch, 1800.0, f, nlev, ncol, col_lev=P.col_lev))
    ref, _ = orc.gas_phase_chemistry(1800.0, *f)
    scale = np.maximum(np.abs(f[5]), np.abs(ref))
    err = np.where(scale > 0, np.abs(got - ref) / np.where(scale > 0, scale, 1.0), 0.0)
    assert err[1:].max() < 1e-12


@pytest.mark.parametrize("layout", LAYOUTS)
def test_setsox(torch_cuda, layout):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    torch = torch_cuda
    P = Pad(torch, layout)
    rng = np.random.default_rng(63)
    nlev, ncol = 30, 66
    met, qcw, qin = sox_domain(rng, nlev * ncol)
    sh = lambda a: np.asarray(a).reshape(a.shape[:-1] + (nlev, ncol))
    qc, qi = P.dev(sh(qcw)), P.dev(sh(qin))
    d1, d2 = (P.dev(np.full((31, nlev, ncol), np.nan)) for _ in range(2))
    ch = Chain()
    ch.compute_setsox(DT, {k: P.dev(sh(met[v])) for k, v in SOX_MET.items()}, qc, qi, d1, d2,
                      layout=P.layout, live=P.live(nlev, ncol))
    torch.cuda.synchronize()
    ch.close()
    got = [P.back(t).reshape(31, -1) for t in (qc, qi, d1, d2)]
    lay = _abi.LAYOUT_COL_LEV if P.col_lev else _abi.LAYOUT_LEV_COL
    ref = sox_packed(torch, met, qcw, qin, nlev, ncol, layout=lay)
    for x, y in zip(got, ref):
        assert np.isfinite(x).all()
        assert np.array_equal(x, y)
    assert np.count_nonzero(got[2]) > 100
