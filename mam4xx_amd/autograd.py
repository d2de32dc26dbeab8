"""torch.autograd.Function over the microphysics step.

``tendencies(state, temperature, pressure, qv, scene, dt)`` returns the
[52, nlev, ncol] tendencies of one step (mam4x_compute_tendencies) and is
differentiable in the prognostic state, temperature, pressure and water
vapour.

* Forward mode (``torch.autograd.forward_ad``): one call of the tangent
  kernel (mam4x_compute_tangent) per direction.  The functorch transforms
  (``torch.func``) are not supported: the kernels need plain device tensors.
* Reverse mode (``backward``): one call of mam4x_compute_adjoint, a
  hand-written reverse sweep of the step (one lane per cell, the SOA
  sub-step states kept as checkpoints), which returns every requested
  gradient plane at once.  Cells are independent, so each cell's Jacobian is
  a 52 x 55 block and the product needs no communication between lanes.

Derivatives are those of the branch each cell takes (include/mam4x.h,
mam4x_compute_tangent).
"""
import torch

from . import layout as L
from .chain import Chain

_chains = {}


def _chain():
    dev = torch.cuda.current_device()
    if dev not in _chains:
        _chains[dev] = Chain()
    return _chains[dev]


def _scene(scene, state, temperature, pressure, qv):
    d = dict(scene)
    d["state"] = state.detach().contiguous()
    d["temperature"] = temperature.detach().contiguous()
    d["pressure"] = pressure.detach().contiguous()
    d["qv"] = qv.detach().contiguous()
    return d


def _tangent(d, dt, dstate, dT, dp, dqv):
    nlev, ncol = d["temperature"].shape
    tend = torch.empty((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda")
    dtend = torch.empty_like(tend)
    _chain().compute_tangent(d, dt, tend, dtend, dstate=dstate, dtemperature=dT, dpressure=dp,
                             dqv=dqv)
    return tend, dtend


class _Tendencies(torch.autograd.Function):
    @staticmethod
    def forward(state, temperature, pressure, qv, scene, dt):
        d = _scene(scene, state, temperature, pressure, qv)
        nlev, ncol = d["temperature"].shape
        tend = torch.empty((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda")
        _chain().compute_tendencies(d, dt, tend)
        return tend

    @staticmethod
    def setup_context(ctx, inputs, output):
        state, temperature, pressure, qv, scene, dt = inputs
        ctx.save_for_forward(state, temperature, pressure, qv)
        ctx.save_for_backward(state, temperature, pressure, qv)
        ctx.rest, ctx.dt = scene, dt

    @staticmethod
    def jvp(ctx, dstate, dT, dp, dqv, _dscene, _ddt):
        c = lambda x: None if x is None else x.contiguous()
        d = _scene(ctx.rest, *ctx.saved_tensors)
        _, dtend = _tangent(d, ctx.dt, c(dstate), c(dT), c(dp), c(dqv))
        return dtend

    @staticmethod
    def backward(ctx, g):
        d, dt = _scene(ctx.rest, *ctx.saved_tensors), ctx.dt
        g = g.contiguous()
        nlev, ncol = d["temperature"].shape
        new = lambda *shape: torch.empty(shape, dtype=torch.float64, device="cuda")
        grads = [None, None, None, None]
        if ctx.needs_input_grad[0]:
            grads[0] = new(L.NSTATE, nlev, ncol)
        for k in (1, 2, 3):
            if ctx.needs_input_grad[k]:
                grads[k] = new(nlev, ncol)
        _chain().compute_adjoint(d, dt, g, state_bar=grads[0], temperature_bar=grads[1],
                                 pressure_bar=grads[2], qv_bar=grads[3])
        return grads[0], grads[1], grads[2], grads[3], None, None


def tendencies(state, temperature, pressure, qv, scene, dt):
    """Differentiable tendencies of one step; scene supplies the other
    inputs (height, cloud fraction, pblh, diameters, wet density) as CUDA
    float64 tensors in the [field][level][column] layout."""
    return _Tendencies.apply(state, temperature, pressure, qv, scene, dt)
