"""Tangent entry point and torch.autograd.Function on the GPU.

* mam4x_compute_tangent's primal outputs equal the FP64 fused kernel's bit
  for bit (the dual numbers' value part performs the same operations);
* its derivative matches central finite differences of the CPU oracle on the
  smooth elements (see test_tangent_host.py for the criterion);
* autograd: reverse mode is the adjoint of forward mode, <g, J v> = <J^T g, v>
  (forward mode through torch.autograd.forward_ad).
"""
import numpy as np
import pytest

from mam4xx_amd import layout as L
from mam4xx_amd import scenes

from .test_tangent_host import central_fd

pytestmark = pytest.mark.gpu
DT = 1800.0


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


def _direction(sc, seed):
    rng = np.random.default_rng(seed)
    return (sc["state"] * rng.standard_normal(sc["state"].shape),
            rng.standard_normal(sc["temperature"].shape),
            sc["qv"] * rng.standard_normal(sc["qv"].shape))


def test_tangent_primal_and_derivative(orc, torch_cuda):
    import oracle.pyoracle as P
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    sc = scenes.generate(48, 72, seed=21)
    v_state, v_T, v_qv = _direction(sc, 22)
    d = scene_to_device(sc)
    ch = Chain(schedule=_abi.SCHED_FUSED)
    shape = (L.NSTATE, 72, 48)
    tend = torch.empty(shape, dtype=torch.float64, device="cuda")
    dtend = torch.empty_like(tend)
    t64 = torch.empty_like(tend)
    ch.compute_tangent(d, DT, tend, dtend, dstate=torch.as_tensor(v_state).cuda(),
                       dtemperature=torch.as_tensor(v_T).cuda(), dqv=torch.as_tensor(v_qv).cuda())
    ch.compute_tendencies(d, DT, t64)
    torch.cuda.synchronize()
    assert torch.equal(tend, t64)
    oc = P.default_config()
    f1 = central_fd(orc, sc, oc, v_state, v_T, v_qv, 1e-6)
    f2 = central_fd(orc, sc, oc, v_state, v_T, v_qv, 5e-7)
    scale = np.abs(sc["state"]) / DT + 1e-300
    smooth = np.abs(f1 - f2) <= 1e-4 * (np.abs(f1) + scale)
    err = np.abs(dtend.cpu().numpy() - f1) / (np.abs(f1) + scale)
    assert smooth.mean() > 0.999
    assert np.median(err[smooth]) < 1e-8
    assert (err[smooth] > 1e-4).mean() < 5e-4, np.sort(err[smooth])[-5:]


def test_autograd_reverse_is_adjoint_of_forward(torch_cuda):
    import torch.autograd.forward_ad as fwAD
    from mam4xx_amd.autograd import tendencies
    from mam4xx_amd.chain import scene_to_device
    torch = torch_cuda
    sc = scenes.generate(16, 72, seed=23)
    d = scene_to_device(sc)
    v_state, v_T, v_qv = (torch.as_tensor(x).cuda() for x in _direction(sc, 24))
    v_p = torch.zeros_like(d["pressure"])
    x = [d["state"], d["temperature"], d["pressure"], d["qv"]]
    f = lambda s, T, p, q: tendencies(s, T, p, q, d, DT)
    with fwAD.dual_level():
        duals = [fwAD.make_dual(a, b) for a, b in zip(x, (v_state, v_T, v_p, v_qv))]
        y, jv = fwAD.unpack_dual(f(*duals))
    g = torch.randn_like(y)
    xs = [t.clone().requires_grad_(True) for t in x]
    y2 = f(*xs)
    assert torch.equal(y, y2.detach())
    gx = torch.autograd.grad(y2, xs, grad_outputs=g)
    lhs = (g * jv).sum().item()
    rhs = sum((gi * vi).sum().item() for gi, vi in zip(gx, (v_state, v_T, v_p, v_qv)))
    assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + abs(rhs)), (lhs, rhs)
