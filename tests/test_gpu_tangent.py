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


def _gpu_jacobians(ch, d, nlev, ncol, pre=None):
    """Every cell's 52 x 55 Jacobian block twice: columns from the tangent
    (a unit seed per input), rows from the reverse sweep (a unit cotangent
    per output)."""
    import torch
    nf = L.NSTATE
    Jt = torch.empty((nf, nf + 3, nlev, ncol), dtype=torch.float64, device="cuda")
    Ja = torch.full_like(Jt, float("nan"))
    tend = torch.empty((nf, nlev, ncol), dtype=torch.float64, device="cuda")
    dtend = torch.empty_like(tend)
    one = torch.ones((nlev, ncol), dtype=torch.float64, device="cuda")
    for j in range(nf + 3):
        kw = {}
        if j < nf:
            unit = torch.zeros((nf, nlev, ncol), dtype=torch.float64, device="cuda")
            unit[j] = 1.0
            kw["dstate"] = unit
        else:
            kw[("dtemperature", "dpressure", "dqv")[j - nf]] = one
        ch.compute_tangent(d, DT, tend, dtend, precldchem=pre, **kw)
        Jt[:, j] = dtend
    for f in range(nf):
        tb = [None] * nf
        tb[f] = one
        gs = torch.empty((nf, nlev, ncol), dtype=torch.float64, device="cuda")
        gT, gp, gq = (torch.empty((nlev, ncol), dtype=torch.float64, device="cuda")
                      for _ in range(3))
        ch.compute_adjoint(d, DT, tb, state_bar=gs, temperature_bar=gT, pressure_bar=gp,
                           qv_bar=gq, precldchem=pre)
        Ja[f, :nf] = gs
        Ja[f, nf], Ja[f, nf + 1], Ja[f, nf + 2] = gT, gp, gq
    torch.cuda.synchronize()
    return Jt.cpu().numpy(), Ja.cpu().numpy()


@pytest.mark.parametrize("precld", [False, True])
def test_adjoint_jacobian_equals_tangent_jacobian(torch_cuda, precld):
    """mam4x_compute_adjoint (a reverse sweep) against mam4x_compute_tangent
    (forward mode) element by element, in units of the largest contribution
    |J_fj' x_j'| to the element's row (or |x_f| / dt): the two modes
    accumulate the chain rule in opposite orders, so cancellation in the
    SOA loop leaves a few elements at 1e-9 of their row rather than at
    round-off.  With and without a pre-cloud-chemistry state; a quarter of
    the columns overcast, a quarter partly cloudy."""
    from mam4xx_amd.chain import Chain, scene_to_device
    from .test_adjoint_host import row_scaled_error
    sc = scenes.generate(8, 72, seed=25)
    rng = np.random.default_rng(25)
    sc["cloud_fraction"][:, :2] = 1.0
    sc["cloud_fraction"][:, 2:4] = rng.uniform(0.05, 0.95, (72, 2))
    d = scene_to_device(sc)
    pre = None
    if precld:
        import torch
        pre = torch.as_tensor(sc["state"] * rng.uniform(0.5, 1.0, sc["state"].shape)).cuda()
    ch = Chain()
    Jt, Ja = _gpu_jacobians(ch, d, 72, 8, pre)
    assert np.isfinite(Ja).all()
    err = row_scaled_error(sc, Jt, Ja)
    print("row-scaled |J_rev - J_tan|: max %.2e, 99.99 %% %.2e, median %.2e"
          % (err.max(), np.quantile(err, 0.9999), np.median(err)))
    assert np.quantile(err, 0.9999) < 1e-11
    assert err.max() < 1e-8


def test_adjoint_entry_point_planes_and_dot_product(torch_cuda):
    """NULL cotangent planes count as zero, requesting one gradient plane
    gives the same bits as requesting all, and <w, J v> = <J^T w, v> in every
    cell for a random direction v (the tolerance as on the CPU build)."""
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    nlev, ncol = 72, 24
    sc = scenes.generate(ncol, nlev, seed=25)
    d = scene_to_device(sc)
    ch = Chain(schedule=_abi.SCHED_FUSED)
    st = torch.as_tensor(sc["state"]).cuda()
    g = torch.Generator(device="cuda").manual_seed(26)
    w = torch.randn((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda", generator=g)
    w = w / st.abs().clamp_min(1e-30) * 1e-3
    w[5] = 0.0
    wp = [None if f == 7 else w[f] for f in range(L.NSTATE)]  # a NULL cotangent plane
    w[7] = 0.0
    gs = torch.full((L.NSTATE, nlev, ncol), float("nan"), dtype=torch.float64, device="cuda")
    gT, gp, gq = (torch.full((nlev, ncol), float("nan"), dtype=torch.float64, device="cuda")
                  for _ in range(3))
    ch.compute_adjoint(d, DT, wp, state_bar=gs, temperature_bar=gT, pressure_bar=gp, qv_bar=gq)
    gT1 = torch.empty_like(gT)
    ch.compute_adjoint(d, DT, w, temperature_bar=gT1)
    assert torch.equal(gT1, gT)
    for x in (gs, gT, gp, gq):
        assert torch.isfinite(x).all()
    tend = torch.empty((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda")
    dtend = torch.empty_like(tend)
    v_state, v_T, v_qv = (torch.as_tensor(x).cuda() for x in _direction(sc, 27))
    v_state = v_state * 1e-3
    v_p = d["pressure"] * torch.randn((nlev, ncol), dtype=torch.float64, device="cuda",
                                      generator=g) * 1e-4
    ch.compute_tangent(d, DT, tend, dtend, dstate=v_state, dtemperature=v_T, dpressure=v_p,
                       dqv=v_qv)
    lhs = (w * dtend).sum(dim=0)
    rhs = (gs * v_state).sum(dim=0) + gT * v_T + gp * v_p + gq * v_qv
    mag = (w * dtend).abs().sum(dim=0) + (gs * v_state).abs().sum(dim=0) + (gT * v_T).abs() \
        + (gp * v_p).abs() + (gq * v_qv).abs() + (w.abs() * v_state.abs()).sum(dim=0) / DT
    rel = ((lhs - rhs).abs() / (mag + 1e-300)).cpu().numpy()
    print("per-cell <w, Jv> vs <J^T w, v>: median %.2e, max %.2e, cells > 1e-8: %d of %d"
          % (np.median(rel), rel.max(), (rel > 1e-8).sum(), rel.size))
    assert np.median(rel) < 1e-14
    assert (rel > 1e-8).mean() <= 2e-3
    assert rel.max() < 1e-5


def test_adjoint_against_finite_differences_full_domain(torch_cuda):
    """The reverse sweep at the study's size, 21,600 x 72, against central
    differences of the FP64 step itself (the product's forward kernel) along
    a random direction, on 64 columns spread over the domain: per cell,
    <w, (f(x + h v) - f(x - h v)) / 2h> against <J^T w, v>.  A cell counts
    as smooth when the differences at h and h/2 agree to 1e-5 of the cell's
    scale (clips and regime switches are not differentiable); there the two
    must agree to 1e-5."""
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    ncol, nlev = 21600, 72
    sc = scenes.generate(ncol, nlev, seed=0)
    cols = np.linspace(0, ncol - 1, 64).astype(np.int64)
    sub = {k: (v[..., cols] if k != "pblh" else v[cols]) for k, v in sc.items()}
    d = scene_to_device(sc)
    ch = Chain()
    rng = np.random.default_rng(31)
    st = sc["state"]
    w_np = np.zeros_like(st)
    w_np[..., cols] = (rng.standard_normal((L.NSTATE, nlev, 64))
                       / np.maximum(np.abs(sub["state"]), 1e-30) * 1e-3)
    w = torch.as_tensor(w_np).cuda()
    gs = torch.empty_like(w)
    gT, gp, gq = (torch.empty((nlev, ncol), dtype=torch.float64, device="cuda")
                  for _ in range(3))
    ch.compute_adjoint(d, DT, w, state_bar=gs, temperature_bar=gT, pressure_bar=gp, qv_bar=gq)
    torch.cuda.synchronize()
    for x in (gs, gT, gp, gq):
        assert torch.isfinite(x).all()
    v = np.zeros_like(st)
    v[..., cols] = sub["state"] * rng.standard_normal((L.NSTATE, nlev, 64))
    vT = np.zeros((nlev, ncol))
    vT[:, cols] = rng.standard_normal((nlev, 64))
    vq = np.zeros((nlev, ncol))
    vq[:, cols] = sub["qv"] * rng.standard_normal((nlev, 64))
    tend = torch.empty_like(w)

    def fd(h):
        out = []
        for sgn in (1.0, -1.0):
            dd = dict(d)
            dd["state"] = d["state"] + sgn * h * torch.as_tensor(v).cuda()
            dd["temperature"] = d["temperature"] + sgn * h * torch.as_tensor(vT).cuda()
            dd["qv"] = d["qv"] + sgn * h * torch.as_tensor(vq).cuda()
            ch.compute_tendencies(dd, DT, tend)
            out.append(tend[..., cols].cpu().numpy())
        return (out[0] - out[1]) / (2 * h)

    f1, f2 = fd(1e-6), fd(5e-7)
    wc = w_np[..., cols]
    lhs1, lhs2 = (wc * f1).sum(axis=0), (wc * f2).sum(axis=0)
    gsc = gs[..., cols].cpu().numpy()
    rhs = ((gsc * v[..., cols]).sum(axis=0) + gT[:, cols].cpu().numpy() * vT[:, cols]
           + gq[:, cols].cpu().numpy() * vq[:, cols])
    scale = (np.abs(wc * f1).sum(axis=0) + np.abs(gsc * v[..., cols]).sum(axis=0)
             + (np.abs(wc) * np.abs(v[..., cols])).sum(axis=0) / DT + 1e-300)
    smooth = np.abs(lhs1 - lhs2) <= 1e-5 * scale
    rel = np.abs(lhs1 - rhs) / scale
    print("64 columns x 72 levels: smooth cells %.2f %%, |FD - adjoint| / scale on them: "
          "median %.2e, 99.9 %% %.2e, max %.2e"
          % (100 * smooth.mean(), np.median(rel[smooth]), np.quantile(rel[smooth], 0.999),
             rel[smooth].max()))
    assert smooth.mean() > 0.99
    assert np.median(rel[smooth]) < 1e-8
    assert rel[smooth].max() < 1e-5
