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


def test_adjoint_entry_point_per_cell(torch_cuda):
    """mam4x_compute_adjoint against the tangent, cell by cell:
    grad_j = sum_f w_f d tend_f / d x_j for unit seeds of single fields, and
    <w, J v> = <J^T w, v> in every cell for a random direction v."""
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    nlev, ncol = 72, 24
    sc = scenes.generate(ncol, nlev, seed=25)
    d = scene_to_device(sc)
    ch = Chain(schedule=_abi.SCHED_FUSED)
    g = torch.Generator(device="cuda").manual_seed(26)
    w = torch.randn((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda", generator=g)
    w[5] = 0.0
    wp = [None if f == 7 else w[f] for f in range(L.NSTATE)]  # a NULL cotangent plane
    w[7] = 0.0
    gs = torch.full((L.NSTATE, nlev, ncol), float("nan"), dtype=torch.float64, device="cuda")
    gT, gp, gq = (torch.full((nlev, ncol), float("nan"), dtype=torch.float64, device="cuda")
                  for _ in range(3))
    ch.compute_adjoint(d, DT, wp, state_bar=gs, temperature_bar=gT, pressure_bar=gp, qv_bar=gq)
    # requesting one plane gives the same bits as requesting all
    gT1 = torch.empty_like(gT)
    ch.compute_adjoint(d, DT, w, temperature_bar=gT1)
    assert torch.equal(gT1, gT)
    tend = torch.empty((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda")
    dtend = torch.empty_like(tend)
    one = torch.ones((nlev, ncol), dtype=torch.float64, device="cuda")
    for j in (0, 4, 9, 13, 30, 45, 50, 51, "T", "p", "qv"):
        unit = torch.zeros((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda")
        kw = {}
        if j == "T":
            kw["dtemperature"] = one
            got = gT
        elif j == "p":
            kw["dpressure"] = one
            got = gp
        elif j == "qv":
            kw["dqv"] = one
            got = gq
        else:
            unit[j] = 1.0
            kw["dstate"] = unit
            got = gs[j]
        ch.compute_tangent(d, DT, tend, dtend, **kw)
        ref = (w * dtend).sum(dim=0)
        mag = (w * dtend).abs().sum(dim=0)
        assert torch.isfinite(got).all(), j
        assert ((got - ref).abs() <= 1e-13 * mag + 1e-300).all(), (j, (got - ref).abs().max())
    # per-cell dot-product test along a random direction
    v_state, v_T, v_qv = (torch.as_tensor(x).cuda() for x in _direction(sc, 27))
    v_p = d["pressure"] * torch.randn((nlev, ncol), dtype=torch.float64, device="cuda",
                                      generator=g) * 1e-3
    ch.compute_tangent(d, DT, tend, dtend, dstate=v_state, dtemperature=v_T, dpressure=v_p,
                       dqv=v_qv)
    lhs = (w * dtend).sum(dim=0)
    rhs = (gs * v_state).sum(dim=0) + gT * v_T + gp * v_p + gq * v_qv
    mag = (w * dtend).abs().sum(dim=0) + (gs * v_state).abs().sum(dim=0) + (gT * v_T).abs() \
        + (gp * v_p).abs() + (gq * v_qv).abs()
    assert ((lhs - rhs).abs() <= 1e-11 * mag + 1e-300).all(), ((lhs - rhs).abs() / mag).max()
