"""mam4x_compute_linoz and mam4x_compute_invariants on the GPU against the
oracle's orc_linoz_column and orc_setinv, pinned in test_linoz.py (whose
host build of the kernels' code equals the oracle bit for bit).

Bar: the overhead ozone column and the invariants bit for bit (products and
sums only); the advanced ozone within 1e-12 of its magnitude, the
tendencies within 1e-12 of the ozone over dt, the steady state and the
surface sink within 1e-12 relative (device exp / log10 / sqrt).  Layouts
agree bit for bit."""
import numpy as np
import pytest

from .test_linoz import CONF, columns

pytestmark = pytest.mark.gpu

DT = 1800.0


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


def _gpu_linoz(torch, cols, col_lev=False, conf=CONF, outputs=True):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    tr = (lambda a: np.ascontiguousarray(np.swapaxes(a, -1, -2))) if col_lev else (lambda a: a)
    dev = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).cuda()
    pl = lambda k: dev(tr(cols[k]))
    o3 = pl("o3")
    nlev, ncol = cols["o3"].shape
    new = lambda: dev(tr(np.full((nlev, ncol), np.nan))) if outputs else None
    o3col, do3, dpsc, ss = new(), new(), new(), new()
    sink = dev(np.full(ncol, np.nan)) if outputs else None
    ch = Chain()
    ch.compute_linoz(DT, pl("temp"), pl("pmid"), pl("pdel"), dev(tr(cols["clim"])),
                     dev(cols["sza"]), dev(cols["rlats"]), dev(cols["ltrop"], torch.int32), o3,
                     conf, column_above_top=dev(cols["col0"]), o3_column=o3col, do3_linoz=do3,
                     do3_linoz_psc=dpsc, ss_o3=ss, o3_sfcsink=sink,
                     layout=_abi.LAYOUT_COL_LEV if col_lev else _abi.LAYOUT_LEV_COL)
    torch.cuda.synchronize()
    ch.close()
    back = lambda t: tr(t.cpu().numpy()) if t is not None else None
    return back(o3), back(o3col), back(do3), back(dpsc), back(ss), \
        (sink.cpu().numpy() if sink is not None else None)


@pytest.mark.parametrize("nlev", [72, 128, 150])  # LDS block kernel up to 128 levels (64 KiB)
def test_linoz_matches_oracle(orc, torch_cuda, nlev):
    rng = np.random.default_rng(21 + nlev)
    cols = columns(rng, nlev, 500)
    cols["ltrop"][:5] = (0, 1, nlev, nlev - 1, nlev + 5)
    g = _gpu_linoz(torch_cuda, cols)
    r = orc.linoz(DT, cols, CONF)
    o3_0 = cols["o3"]
    scale = np.maximum(np.abs(o3_0), np.abs(r[0]))
    assert np.isfinite(g[0]).all()
    assert (np.abs(g[0] - r[0]) <= 1e-12 * scale).all(), np.abs(g[0] - r[0]).max()
    assert np.array_equal(g[1], r[1])  # overhead column
    for x, y in zip(g[2:4], r[2:4]):  # tendencies
        assert (np.abs(x - y) <= 1e-12 * scale / DT).all()
    assert np.array_equal(g[3] != 0, r[3] != 0) and (r[3] != 0).sum() > 100
    assert (np.abs(g[4] - r[4]) <= 1e-12 * np.abs(r[4])).all()
    assert (np.abs(g[5] - r[5]) <= 1e-12 * np.abs(r[5])).all()


def test_linoz_layouts_and_optional_outputs(orc, torch_cuda):
    rng = np.random.default_rng(22)
    cols = columns(rng, 72, 96)
    a = _gpu_linoz(torch_cuda, cols)
    b = _gpu_linoz(torch_cuda, cols, col_lev=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    c = _gpu_linoz(torch_cuda, cols, outputs=False)
    assert np.array_equal(c[0], a[0])
    # no surface sink: the lowest levels keep the relaxation alone
    d = _gpu_linoz(torch_cuda, cols, conf=dict(CONF, o3_lbl=0))
    assert np.array_equal(d[0][:-4], a[0][:-4]) and not d[5].any()


def test_linoz_rejects_bad_arguments(torch_cuda):
    from mam4xx_amd.chain import Chain, ChainError
    torch = torch_cuda
    rng = np.random.default_rng(23)
    cols = columns(rng, 8, 4)
    dev = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).cuda()
    ch = Chain()
    args = (DT, dev(cols["temp"]), dev(cols["pmid"]), dev(cols["pdel"]), dev(cols["clim"]),
            dev(cols["sza"]), dev(cols["rlats"]), dev(cols["ltrop"], torch.int32),
            dev(cols["o3"]))
    for conf in (dict(CONF, o3_lbl=9), dict(CONF, o3_lbl=-1), dict(CONF, o3_tau=0.0)):
        with pytest.raises(ChainError):
            ch.compute_linoz(*args, conf)
    ch.close()


def test_invariants_match_oracle(orc, torch_cuda):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    torch = torch_cuda
    rng = np.random.default_rng(24)
    nlev, ncol = 72, 300
    sh = (nlev, ncol)
    T, p, qv = rng.uniform(180, 320, sh), 10 ** rng.uniform(0, 5.01, sh), rng.uniform(0, .03, sh)
    co = 10 ** rng.uniform(-12, -6, (4,) + sh)
    ref = orc.setinv(T, p, qv, co)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a)).cuda()
    ch = Chain()
    for col_lev in (False, True):
        tr = (lambda a: np.swapaxes(a, -1, -2)) if col_lev else (lambda a: a)
        inv = dev(np.full((8,) + tr(T).shape, np.nan))
        ch.compute_invariants(dev(tr(T)), dev(tr(p)), dev(tr(qv)), dev(tr(co)), inv,
                              layout=_abi.LAYOUT_COL_LEV if col_lev else _abi.LAYOUT_LEV_COL)
        torch.cuda.synchronize()
        assert np.array_equal(tr(inv.cpu().numpy()), ref)
    # absent offline species read as 0
    inv = dev(np.full((8,) + sh, np.nan))
    ch.compute_invariants(dev(T), dev(p), dev(qv), None, inv)
    torch.cuda.synchronize()
    got = inv.cpu().numpy()
    assert np.array_equal(got[:4], ref[:4]) and not got[4:].any()
    ch.close()
