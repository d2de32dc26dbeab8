"""Batches of uneven scenes (mam4x_batch_*) on the GPU.

* one batched launch sequence gives, scene by scene, bitwise the results of
  stepping each scene on its own (same per-cell operations, one shared work
  queue), column budgets included;
* a batch run captured in a CUDA graph and replayed over chained steps
  matches an eager batch stepping an identical copy, replay by replay, with
  eager work and synchronisation between the replays;
* a batch owns its workspace: a captured run stays valid after the same
  handle grows its own workspace with a larger single-domain step.
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


def _outputs(torch, sc):
    nlev, ncol = sc["temperature"].shape
    return (torch.empty((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda"),
            torch.empty((4, nlev, ncol), dtype=torch.float64, device="cuda"))


@pytest.mark.parametrize("sched", ["queue", "fused"])
def test_batch_equals_scene_by_scene(torch_cuda, sched):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    code = _abi.SCHED_QUEUE if sched == "queue" else _abi.SCHED_FUSED
    scs = scenes.uneven_batch(nscenes=7, seed=3, total_cols=700)
    scs.append(scenes.generate(1, 72, seed=9))  # a one-column scene
    ds = [scene_to_device(s) for s in scs]
    ch = Chain(schedule=code)
    outs = [_outputs(torch, s) for s in scs]
    # column budgets for every other scene (None elsewhere)
    buds = [torch.full((_abi.NBUDGET, s["temperature"].shape[1]), float("nan"),
                       dtype=torch.float64, device="cuda") if i % 2 == 0 else None
            for i, s in enumerate(scs)]
    b = ch.batch(ds, DT, [o[0] for o in outs], dgn_i=[o[1] for o in outs], budgets=buds)
    b.run()
    torch.cuda.synchronize()
    for d, (tb, gb), bb in zip(ds, outs, buds):
        t1, g1 = _outputs(torch, {"temperature": d["temperature"]})
        b1 = None if bb is None else torch.empty_like(bb)
        ch.compute_tendencies(d, DT, t1, dgn_i=g1, budgets=b1)
        torch.cuda.synchronize()
        assert torch.equal(tb, t1) and torch.equal(gb, g1)
        if bb is not None:
            assert torch.equal(bb, b1)
    b.close()
    ch.close()


def test_batch_graph_replays_match_eager(torch_cuda):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    scs = scenes.uneven_batch(nscenes=5, seed=4, total_cols=600)
    copies = []
    for _ in range(2):
        ds = [scene_to_device(s) for s in scs]
        for d in ds:
            d["dgncur_awet"] = d["dgncur_a"]
        copies.append(ds)
    chans = [Chain(schedule=_abi.SCHED_QUEUE), Chain(schedule=_abi.SCHED_QUEUE)]
    batches = []
    buds = []
    for ch, ds in zip(chans, copies):
        outs = [_outputs(torch, s) for s in scs]
        bud = [torch.empty((_abi.NBUDGET, s["temperature"].shape[1]), dtype=torch.float64,
                           device="cuda") for s in scs]
        buds.append(bud)
        batches.append(ch.batch(ds, DT, [o[0] for o in outs], dgn_i=[d["dgncur_a"] for d in ds],
                                state_out=[d["state"] for d in ds], budgets=bud))
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        batches[0].run()  # warm-up outside capture (first use of the workspace)
    torch.cuda.current_stream().wait_stream(side)
    batches[1].run()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        batches[0].run()
    for step in range(4):
        for x in buds[0]:
            x.fill_(float("nan"))  # each replay writes them anew
        graph.replay()
        batches[1].run()
        torch.cuda.synchronize()
        _ = torch.ones(1024, device="cuda").sum().item()  # eager work between replays
        for da, db in zip(copies[0], copies[1]):
            assert torch.equal(da["state"], db["state"]), step
            assert torch.equal(da["dgncur_a"], db["dgncur_a"]), step
        for xa, xb in zip(buds[0], buds[1]):
            assert torch.equal(xa, xb), step
    for b in batches:
        b.close()


def test_batch_graph_survives_handle_workspace_growth(torch_cuda):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    scs = scenes.uneven_batch(nscenes=4, seed=6, total_cols=400)
    copies = []
    for _ in range(2):
        ds = [scene_to_device(s) for s in scs]
        for d in ds:
            d["dgncur_awet"] = d["dgncur_a"]
        copies.append(ds)
    chans = [Chain(schedule=_abi.SCHED_QUEUE), Chain(schedule=_abi.SCHED_QUEUE)]
    batches = []
    for ch, ds in zip(chans, copies):
        outs = [_outputs(torch, s) for s in scs]
        batches.append(ch.batch(ds, DT, [o[0] for o in outs], dgn_i=[d["dgncur_a"] for d in ds],
                                state_out=[d["state"] for d in ds]))
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        batches[0].run()
    torch.cuda.current_stream().wait_stream(side)
    batches[1].run()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        batches[0].run()
    # a single-domain step larger than the batch on the capturing handle:
    # the handle's own workspace grows (its old buffers are freed)
    big = scene_to_device(scenes.generate(1200, 72, seed=21))
    tb, gb = _outputs(torch, scenes.generate(1200, 72, seed=21))
    for _ in range(2):
        chans[0].compute_tendencies(big, DT, tb, dgn_i=gb)
        torch.cuda.synchronize()
        for step in range(2):
            graph.replay()
            batches[1].run()
            torch.cuda.synchronize()
            for da, db in zip(copies[0], copies[1]):
                assert torch.equal(da["state"], db["state"]), step
                assert torch.equal(da["dgncur_a"], db["dgncur_a"]), step
        big = scene_to_device(scenes.generate(2400, 72, seed=22))
        tb, gb = _outputs(torch, scenes.generate(2400, 72, seed=22))
    for b in batches:
        b.close()


def test_autotuner_picks_a_candidate(torch_cuda, tmp_path):
    from mam4xx_amd import tune
    from mam4xx_amd.chain import scene_to_device
    ds = [scene_to_device(s) for s in scenes.uneven_batch(nscenes=3, seed=5, total_cols=300)]
    before = [d["state"].clone() for d in ds]
    cache = str(tmp_path / "tune.json")
    best, tried = tune.tune(ds, DT, steps=2, cache=cache)
    assert best in [t[0] for t in tried] and len(tried) >= len(tune.CANDIDATES)
    assert all(t[1] > 0 for t in tried)
    for d, s0 in zip(ds, before):  # tuning ran on copies
        assert torch_cuda.equal(d["state"], s0)
    assert tune.tune(ds, DT, cache=cache)[0] == best  # cached


def test_launch_configurations_are_bitwise_neutral(torch_cuda):
    """Every knob the autotuner turns leaves the results bit for bit:
    queue granularity, resident blocks, level tiles (scenes whose column
    counts let 2-, 4- and 8-level tiles cover them, and one that falls back),
    on an uneven batch and on a single domain."""
    from mam4xx_amd import _abi
    from mam4xx_amd import tune
    from mam4xx_amd.chain import Chain, default_config, scene_to_device
    torch = torch_cuda
    scs = [scenes.generate(nc, 72, seed=70 + i) for i, nc in enumerate((512, 64, 200, 1024))]
    configs = [dict(tune.DEFAULTS, schedule=_abi.SCHED_QUEUE, soa_blocks_per_cu=0)]
    configs += [dict(configs[0], **{k: v}) for k, vs in tune.QUEUE_KNOBS for v in vs]
    configs += [dict(configs[0], levels_per_wave=v) for v in tune.TILE_KNOB[1]]
    configs += [dict(configs[0], soa_blocks_per_cu=8)]  # as many as fit
    configs += [dict(configs[0], schedule=_abi.SCHED_FUSED, levels_per_wave=4),
                dict(configs[0], soa_blocks_per_cu=1, soa_chunk=16, soa_short_chunk=64,
                     soa_short_bucket=4, levels_per_wave=2)]
    ref = None
    for cfgd in configs:
        ch = Chain(config=default_config(**cfgd))
        ds = [scene_to_device(s) for s in scs]
        outs = [torch.empty((L.NSTATE, 72, s["temperature"].shape[1]), dtype=torch.float64,
                            device="cuda") for s in scs]
        b = ch.batch(ds, DT, outs)
        b.run()
        single = torch.empty_like(outs[3])
        ch.compute_tendencies(ds[3], DT, single)
        torch.cuda.synchronize()
        got = [o.cpu() for o in outs] + [single.cpu()]
        b.close()
        ch.close()
        if ref is None:
            ref = got
            continue
        for r, g in zip(ref, got):
            assert torch.equal(r, g), cfgd
