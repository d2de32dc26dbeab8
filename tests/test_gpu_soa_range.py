"""The SOA loop's fast quotients on the GPU.

The FP64 step divides in its SOA sub-steps with num::div_fast (scalar.hpp),
which is IEEE's a / b while numerators are 0 or >= 2^-968 and quotients are
normal; chain_device.hpp (soa_fast_exact) states the input ranges and the
per-sub-step condition (every SOA mass 0 or >= 2^-400) under which that holds
for every quotient of a cell's loop.  A cell outside them is rerun with IEEE
quotients: by soa_ieee_kernel after the work queue, in-lane by the fused
kernel.  These scenes put cells outside on purpose -- SOA masses and SOAG at
1e-300, in the subnormals, or above the input range -- next to ordinary
cells, and both schedules must equal, bit for bit, mam4x_compute_tangent's
primal outputs, whose dual numbers divide with IEEE a / b throughout.
"""
import numpy as np
import pytest

from mam4xx_amd import layout as L
from mam4xx_amd import scenes

pytestmark = pytest.mark.gpu
DT = 1800.0


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


def soa_fields():
    """State indices of SOAG and of SOA in each mode that holds it."""
    out = [L.F_SOAG]
    for m in range(4):
        for s, sp in enumerate(L.MODE_SPECIES[m]):
            if L.SPECIES_NAMES[sp] == "soa":
                out.append(L.f_qi(m, s))
    return out


def tiny_soa_scene(ncol=256, nlev=72, seed=90):
    """Groups of 32 columns: every SOA mass tiny (1e-300), every one subnormal
    (log-uniform over 2^-1074..2^-1022), tiny SOAG only, one tiny mode and
    the rest 0, masses above the input range, ordinary cells, and cloudy
    columns (clear subarea split off) with tiny and with subnormal masses."""
    sc = scenes.generate(ncol, nlev, seed=seed)
    st = sc["state"]
    rng = np.random.default_rng(seed)
    g = [slice(32 * i, 32 * i + 32) for i in range(8)]
    sc["cloud_fraction"][:, g[6]] = rng.uniform(0.1, 0.9, (nlev, 32))
    sc["cloud_fraction"][:, g[7]] = rng.uniform(0.1, 0.9, (nlev, 32))

    def tiny():
        return 1e-300 * rng.uniform(0.5, 2.0, (nlev, 32))

    def subnormal():
        return 2.0 ** (-1074.0 + 52.0 * rng.uniform(0.0, 1.0, (nlev, 32)))

    for k, fi in enumerate(soa_fields()):
        st[fi, :, g[0]] = tiny()
        st[fi, :, g[1]] = subnormal()
        if k == 0:
            st[fi, :, g[2]] = tiny()
        st[fi, :, g[3]] = tiny() if k == 1 else 0.0
        st[fi, :, g[4]] = rng.uniform(1e3, 1e4, (nlev, 32))
        # g[5]: ordinary cells
        st[fi, :, g[6]] = tiny()
        st[fi, :, g[7]] = subnormal()
    return sc


@pytest.mark.parametrize("schedule", ["queue", "fused"])
def test_schedules_equal_ieee_tangent_primal(torch_cuda, schedule):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    sc = tiny_soa_scene()
    d = scene_to_device(sc)
    nlev, ncol = sc["temperature"].shape
    shape = (L.NSTATE, nlev, ncol)
    ref = torch.empty(shape, dtype=torch.float64, device="cuda")
    dref = torch.empty_like(ref)
    Chain(schedule=_abi.SCHED_FUSED).compute_tangent(d, DT, ref, dref)
    sched = {"queue": _abi.SCHED_QUEUE, "fused": _abi.SCHED_FUSED}[schedule]
    ch = Chain(schedule=sched)
    tend = torch.empty_like(ref)
    ch.compute_tendencies(d, DT, tend)
    torch.cuda.synchronize()
    r, t = ref.cpu().numpy(), tend.cpu().numpy()
    assert np.isfinite(t).all()
    # the scene reaches the tiny range in the outputs it checks
    tiny = (np.abs(t[soa_fields()]) < 1e-290) & (t[soa_fields()] != 0)
    assert tiny.any()
    bad = np.argwhere(r != t)
    assert bad.size == 0, (len(bad), bad[:5])
