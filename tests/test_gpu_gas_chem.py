"""mam4x_compute_gas_chemistry on the GPU against the oracle's
orc_gas_phase_chemistry_domain, pinned in test_gas_chem.py (whose host build
of the kernel's cell code equals the oracle bit for bit).

Bar: every advanced mixing ratio within 1e-12 of the larger of its values
before and after the step (rate constants carry device-libm rounding; the
solve itself is a few well-conditioned divisions).  Layouts agree bit for
bit."""
import numpy as np
import pytest

from .test_gas_chem import _domain

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


def _gpu(torch, dt, fields, nlev, ncol, col_lev=False, drop=()):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    temp, photo, extfrc, inv, het, qq = fields
    tr = (lambda a: np.ascontiguousarray(np.swapaxes(a, -1, -2))) if col_lev else (lambda a: a)
    sh = lambda a: torch.as_tensor(tr(np.asarray(a).reshape(a.shape[:-1] + (nlev, ncol)))).cuda()
    q = sh(qq)
    ex = sh(extfrc)
    het_t = sh(het)
    ch = Chain()
    ch.compute_gas_chemistry(dt, sh(temp), sh(photo), ex if "extfrc" not in drop else None,
                             sh(inv), het_t if "het" not in drop else None, q,
                             layout=_abi.LAYOUT_COL_LEV if col_lev else _abi.LAYOUT_LEV_COL)
    torch.cuda.synchronize()
    ch.close()
    return tr(q.cpu().numpy()).reshape(31, -1)


@pytest.mark.parametrize("dt", [1.0, 600.0, 1800.0])
def test_gas_chemistry_matches_oracle(orc, torch_cuda, dt):
    rng = np.random.default_rng(int(dt) + 20)
    nlev, ncol = 36, 80
    f = _domain(rng, nlev * ncol)
    g = _gpu(torch_cuda, dt, f, nlev, ncol)
    ref, steps = orc.gas_phase_chemistry(dt, *f)
    qq = f[5]
    scale = np.maximum(np.abs(qq), np.abs(ref))
    err = np.where(scale > 0, np.abs(g - ref) / np.where(scale > 0, scale, 1.0), 0.0)
    assert np.isfinite(g).all()
    assert err[1:].max() < 1e-12, (err.max(), np.unravel_index(err.argmax(), err.shape))
    assert np.array_equal(g[0], qq[0])  # O3 untouched
    assert (ref[1:] != qq[1:]).sum() > qq.shape[1]


def test_layouts_and_absent_rates(orc, torch_cuda):
    rng = np.random.default_rng(30)
    nlev, ncol = 20, 64
    f = _domain(rng, nlev * ncol)
    a = _gpu(torch_cuda, 1800.0, f, nlev, ncol)
    b = _gpu(torch_cuda, 1800.0, f, nlev, ncol, col_lev=True)
    assert np.array_equal(a, b)
    # NULL extfrc / het_rates planes are zero forcing / no washout
    c = _gpu(torch_cuda, 1800.0, f, nlev, ncol, drop=("extfrc", "het"))
    ref, _ = orc.gas_phase_chemistry(1800.0, f[0], f[1], np.zeros_like(f[2]), f[3],
                                     np.zeros_like(f[4]), f[5])
    scale = np.maximum(np.abs(f[5]), np.abs(ref))
    err = np.where(scale > 0, np.abs(c - ref) / np.where(scale > 0, scale, 1.0), 0.0)
    assert err.max() < 1e-12
