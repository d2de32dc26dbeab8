"""mam4x_compute_photolysis on the GPU against the oracle's orc_table_photo,
pinned in test_photo.py (whose host build of the kernels' code equals the
oracle bit for bit).

Bar: every rate within 1e-12 of its value (the cloud optical depths use the
device's pow and cos; the interpolation is products and sums in the
reference's order).  Columns in darkness keep their input rates bit for bit;
layouts and padded planes agree bit for bit; table and argument errors are
reported, not launched."""
import numpy as np
import pytest

import oracle.pyoracle as P

from mam4xx_amd import photo_tables as PT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


def _dev_table(torch, t, **kw):
    d = {k: torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64)).cuda()
         for k, v in t.items()}
    d.update(kw)
    return d


def _gpu(torch, t, cols, col_lev=False, pad=0, esfact=1.02, **kw):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    tr = (lambda a: np.ascontiguousarray(np.swapaxes(a, -1, -2))) if col_lev else (lambda a: a)

    def plane(a):
        a = tr(np.asarray(a, dtype=np.float64))
        x = torch.full(a.shape[:-1] + (a.shape[-1] + pad,), float("nan"), dtype=torch.float64,
                       device="cuda")
        x[..., :a.shape[-1]] = torch.as_tensor(a)
        return x

    nlev, ncol = cols["photo"].shape
    photo = plane(cols["photo"])
    names = dict(pressure="pmid", pdel="pdel", temperature="temp", o3_column="colo3",
                 liquid_water="lwc", cloud_fraction="clouds")
    planes = {k: plane(cols[v]) for k, v in names.items()}
    col = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    ch = Chain()
    live = (nlev if col_lev else ncol) if pad else None
    ch.compute_photolysis(_dev_table(torch, t, **kw), planes, col(cols["zen_angle"]),
                          col(cols["srf_alb"]), photo, esfact=esfact,
                          layout=_abi.LAYOUT_COL_LEV if col_lev else _abi.LAYOUT_LEV_COL,
                          live=live)
    torch.cuda.synchronize()
    ch.close()
    out = photo.cpu().numpy()
    if pad:
        assert np.isnan(out[..., -pad:]).all(), "padding written"
        out = out[..., :-pad]
    return tr(out)


def _case(seed, nlev=72, ncol=500, nw=20, numj=2):
    rng = np.random.default_rng(seed)
    ax = PT.wide_axes()
    t = PT.random_table(rng, ax, nw=nw, numj=numj)
    cols = PT.columns(rng, nlev, ncol, ax)
    return t, cols


# the blocked column pass with 32 and 16 columns per block; 400 levels take
# the lane-per-column walk
@pytest.mark.parametrize("nlev", [72, 128, 400])
def test_photolysis_matches_oracle(orc, torch_cuda, nlev):
    t, cols = _case(40 + nlev, nlev=nlev, ncol=500 if nlev < 400 else 120)
    g = _gpu(torch_cuda, t, cols, pht_alias_mult=0.9, lng_index=1)
    ref = orc.table_photo(P.photo_table(t, pht_alias_mult=0.9, lng_index=1), cols, 1.02)
    assert np.isfinite(g).all()
    err = np.abs(g - ref) / np.abs(ref)
    assert err.max() < 1e-12, (err.max(), np.unravel_index(err.argmax(), err.shape))
    night = np.rad2deg(cols["zen_angle"]) >= 88.85
    assert night.any() and (~night).any()
    assert np.array_equal(g[:, night], cols["photo"][:, night])
    assert (g[:, ~night] != cols["photo"][:, ~night]).all()


def test_layouts_padding_and_reaction_index(torch_cuda):
    t, cols = _case(50, ncol=130)
    a = _gpu(torch_cuda, t, cols, lng_index=0)
    b = _gpu(torch_cuda, t, cols, col_lev=True, lng_index=0)
    assert np.array_equal(a, b)
    for col_lev in (False, True):
        assert np.array_equal(_gpu(torch_cuda, t, cols, col_lev=col_lev, pad=5, lng_index=0), a)
    # lng_index = -1: every rate as it was
    assert np.array_equal(_gpu(torch_cuda, t, cols, lng_index=-1), cols["photo"])
    # the reaction's row of the cross sections is the one used
    c = _gpu(torch_cuda, t, cols, lng_index=1)
    assert not np.array_equal(a, c)


def test_bad_tables_and_arguments_are_errors(torch_cuda):
    from mam4xx_amd.chain import ChainError
    torch = torch_cuda
    t, cols = _case(51, ncol=8, nw=4, numj=1)
    bad_nt = dict(t, xsqy=t["xsqy"][:, :, :150, :])  # the temperature index runs to 200
    with pytest.raises(ChainError, match="nt >= 201"):
        _gpu(torch, bad_nt, cols)
    with pytest.raises(ChainError, match="lng_index"):
        _gpu(torch, t, cols, lng_index=1)
