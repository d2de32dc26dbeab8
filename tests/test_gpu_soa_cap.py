"""The SOA-cap flag plane (mam4x_outputs.soa_capped) on the GPU: the
work-queue schedule (the flag travels in the cell's bucket byte), the fused
kernel and three sub-steps against the kernel code's CPU build.  The device
libm can move a loop that ends within a sub-step or two of the cap across
it, so a few of the capped cells may differ."""
import numpy as np
import pytest

from .test_soa_cap import capped_scene, host_capped
from .test_cell_host import cellhost  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu
DT = 1800.0


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


@pytest.mark.parametrize("schedule,ns", [("queue", 1), ("fused", 1), ("queue", 3)])
def test_flag_plane_matches_cpu_build(torch_cuda, cellhost, schedule, ns):
    from mam4xx_amd import _abi
    from mam4xx_amd import layout as L
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    sc = capped_scene(ncol=16, seed=1)
    ref = host_capped(cellhost, sc, ns)
    ch = Chain(schedule=_abi.SCHED_QUEUE if schedule == "queue" else _abi.SCHED_FUSED,
               ntsubstep=ns)
    d = scene_to_device(sc)
    tend = torch.empty((L.NSTATE, 72, 16), dtype=torch.float64, device="cuda")
    flags = torch.full((72, 16), -1, dtype=torch.int32, device="cuda")
    ch.compute_tendencies(d, DT, tend, soa_capped=flags)
    torch.cuda.synchronize()
    got = flags.cpu().numpy()
    assert set(np.unique(got)) <= {0, 1}
    print("%s n=%d: capped cells GPU %d, CPU build %d, differing %d"
          % (schedule, ns, got.sum(), ref.sum(), (got != ref).sum()))
    assert ref.sum() > 0
    assert (got != ref).sum() <= max(2, 0.02 * ref.sum())
    ch.close()
