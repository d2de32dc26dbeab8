"""mam4x_init (AeroProcess::init, aero_process.hpp:76-80) on the GPU: a
re-initialized handle steps exactly as a handle created with the new
configuration, a rejected configuration leaves the handle as it was, and a
handle with a live batch refuses to change."""
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


def _run(torch, ch, d):
    nlev, ncol = d["temperature"].shape
    tend = torch.empty((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda")
    ch.compute_tendencies(d, DT, tend)
    torch.cuda.synchronize()
    return tend.cpu().numpy()


def test_reinit_equals_a_fresh_handle(torch_cuda):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, ChainError, default_config, scene_to_device
    d = scene_to_device(scenes.generate(400, 72, seed=71))
    ch = Chain(schedule=_abi.SCHED_QUEUE)
    base = _run(torch_cuda, ch, d)
    for kw in (dict(do_coag=0), dict(do_newnuc=0, ntsubstep=2), dict(precision=_abi.PREC_MIXED),
               dict(soa_blocks_per_cu=1, soa_chunk=16)):
        cfg = default_config(schedule=_abi.SCHED_QUEUE, **kw)
        ch.init(cfg)
        got = _run(torch_cuda, ch, d)
        fresh = Chain(config=default_config(schedule=_abi.SCHED_QUEUE, **kw),
                      precision=cfg.precision)
        assert np.array_equal(got, _run(torch_cuda, fresh, d)), kw
        fresh.close()
        if "soa_chunk" in kw:  # performance-only settings
            assert np.array_equal(got, base)
        else:
            assert not np.array_equal(got, base), kw
    # a rejected configuration keeps the last one
    with pytest.raises(ChainError, match="invalid configuration"):
        ch.init(default_config(levels_per_wave=3))
    ch.init(default_config(schedule=_abi.SCHED_QUEUE))
    assert np.array_equal(_run(torch_cuda, ch, d), base)
    ch.close()


def test_reinit_refused_while_a_batch_lives(torch_cuda):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, ChainError, default_config, scene_to_device
    torch = torch_cuda
    ds = [scene_to_device(s) for s in scenes.uneven_batch(nscenes=2, seed=7, total_cols=128)]
    ch = Chain(schedule=_abi.SCHED_QUEUE)
    outs = [torch.empty((L.NSTATE,) + tuple(d["temperature"].shape), dtype=torch.float64,
                        device="cuda") for d in ds]
    b = ch.batch(ds, DT, outs)
    with pytest.raises(ChainError, match="live batch"):
        ch.init(default_config(do_coag=0))
    b.close()
    ch.init(default_config(do_coag=0))
    ch.close()
