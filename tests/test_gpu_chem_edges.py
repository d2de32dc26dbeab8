"""Edge cases of the chemistry-feeder entry points through the C ABI: empty
domains launch nothing and succeed; planes with a padded leading dimension
(ldc larger than the fastest extent) give the same values as packed planes
and leave the padding untouched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


def _padded(torch, a, pad):
    """[..., nlev, ncol] array -> CUDA tensor [..., nlev, ncol + pad] (NaN
    padding) and its view of the live columns"""
    sh = a.shape[:-1] + (a.shape[-1] + pad,)
    t = torch.full(sh, float("nan"), dtype=torch.float64, device="cuda")
    t[..., :a.shape[-1]] = torch.as_tensor(a)
    return t


def test_empty_domains_succeed(torch_cuda):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    ch = Chain()
    s = ctypes.c_void_p(torch_cuda.cuda.current_stream().cuda_stream)
    one = torch_cuda.zeros(1, dtype=torch_cuda.float64, device="cuda")
    p = one.data_ptr()
    a = _abi.SetinvArgs()
    a.temperature = a.pressure = a.vapor_mixing_ratio = p
    assert ch.lib.mam4x_compute_invariants(ch.h, ctypes.byref(a), s) == _abi.OK
    b = _abi.SetextArgs()
    for m in range(9):
        b.frc_ndx[m] = m + 1
    assert ch.lib.mam4x_compute_forcing(ch.h, ctypes.byref(b), s) == _abi.OK
    c = _abi.LinozArgs()
    c.temperature = c.pressure = c.pdel = c.o3 = c.zenith_angle = c.latitude = c.ltrop = p
    for i in range(8):
        c.climatology[i] = p
    c.nlev = 72
    assert ch.lib.mam4x_compute_linoz(ch.h, 1800.0, ctypes.byref(c), s) == _abi.OK
    d = _abi.SethetArgs()
    for f in ("pressure", "zmid", "temperature", "cmfdqr", "nrain", "nevapr", "air_density",
              "h2o2", "latitude", "phis"):
        setattr(d, f, p)
    for i in (1, 2, 3):
        d.het_rates[i] = p
    d.nlev = 72
    assert ch.lib.mam4x_compute_washout(ch.h, 1800.0, ctypes.byref(d), s) == _abi.OK
    torch_cuda.cuda.synchronize()
    assert one.item() == 0.0  # nothing written
    ch.close()


def test_padded_planes(orc, torch_cuda):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    torch = torch_cuda
    rng = np.random.default_rng(51)
    nlev, ncol, pad = 72, 90, 7
    ld = ncol + pad
    T, p = rng.uniform(180, 320, (nlev, ncol)), 10 ** rng.uniform(0, 5.01, (nlev, ncol))
    qv, co = rng.uniform(0, .03, (nlev, ncol)), 10 ** rng.uniform(-12, -6, (4, nlev, ncol))
    ref = orc.setinv(T, p, qv, co)
    ch = Chain()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    dT, dp, dq, dco = (_padded(torch, x, pad) for x in (T, p, qv, co))
    inv = _padded(torch, np.full((8, nlev, ncol), np.nan), pad)
    a = _abi.SetinvArgs()
    a.ncol, a.nlev, a.ldc, a.layout = ncol, nlev, ld, _abi.LAYOUT_LEV_COL
    a.temperature, a.pressure, a.vapor_mixing_ratio = dT.data_ptr(), dp.data_ptr(), dq.data_ptr()
    for i in range(4):
        a.cnst_offline[i] = dco[i].data_ptr()
    for i in range(8):
        a.invariants[i] = inv[i].data_ptr()
    assert ch.lib.mam4x_compute_invariants(ch.h, ctypes.byref(a), s) == _abi.OK
    # forcing: species m has m % 3 sectors
    sec = [[10 ** rng.uniform(-2, 4, (nlev, ncol)) for _ in range(m % 3)] for m in range(9)]
    fref = orc.setext(sec)
    dsec = [[_padded(torch, x, pad) for x in ss] for ss in sec]
    fr = _padded(torch, np.full((9, nlev, ncol), np.nan), pad)
    b = _abi.SetextArgs()
    b.ncol, b.nlev, b.ldc, b.layout = ncol, nlev, ld, _abi.LAYOUT_LEV_COL
    for m in range(9):
        b.frc_ndx[m] = m + 1
        b.nsectors[m] = len(dsec[m])
        for i, x in enumerate(dsec[m]):
            b.sectors[m][i] = x.data_ptr()
        b.extfrc[m] = fr[m].data_ptr()
    assert ch.lib.mam4x_compute_forcing(ch.h, ctypes.byref(b), s) == _abi.OK
    torch.cuda.synchronize()
    got_inv, got_fr = inv.cpu().numpy(), fr.cpu().numpy()
    assert np.array_equal(got_inv[..., :ncol], ref)
    assert np.isnan(got_inv[..., ncol:]).all()  # padding untouched
    assert np.array_equal(got_fr[..., :ncol], fref)
    assert np.isnan(got_fr[..., ncol:]).all()
    # an ldc below the fastest extent is an error
    a.ldc = ncol - 1
    assert ch.lib.mam4x_compute_invariants(ch.h, ctypes.byref(a), s) == _abi.ERR_ARG
    ch.close()
