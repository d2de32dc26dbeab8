"""Autotuner for the launch configuration of a microphysics step.

The results of a step do not depend on its launch configuration (fused and
work-queue schedules are bitwise identical), only its speed does, and the
best choice depends on the workload: the number and size of the scenes, the
share of cells with an SOA loop and the spread of their sub-step counts.
``tune`` times every candidate on copies of the caller's own scenes and
returns the fastest configuration.  The candidates are:

* schedule: the fused one-lane-per-cell kernel, or the work-queue pipeline;
* soa_blocks_per_cu: resident blocks per CU of the persistent SOA kernel
  (fewer waves raise each wave's issue share on the long serial chains,
  more waves hide latency);

each timed as a batch (one launch sequence for all scenes) over a few
chained steps, eagerly and replayed from a CUDA graph.  Results are cached
per workload signature in a JSON file when ``cache`` is given.
"""
import hashlib
import json
import os
import time

import torch

from . import _abi
from . import layout as L
from .chain import Chain, default_config

CANDIDATES = [
    {"schedule": _abi.SCHED_FUSED, "soa_blocks_per_cu": 0},
    {"schedule": _abi.SCHED_QUEUE, "soa_blocks_per_cu": 0},
    {"schedule": _abi.SCHED_QUEUE, "soa_blocks_per_cu": 1},
    {"schedule": _abi.SCHED_QUEUE, "soa_blocks_per_cu": 2},
    {"schedule": _abi.SCHED_QUEUE, "soa_blocks_per_cu": 3},
]


def signature(scenes, dt, precision):
    shapes = [tuple(s["temperature"].shape) for s in scenes]
    key = json.dumps({"shapes": shapes, "dt": dt, "precision": precision})
    return hashlib.sha1(key.encode()).hexdigest()[:16]


def _copy(scenes):
    out = []
    for s in scenes:
        d = {k: v.clone() for k, v in s.items()}
        if s.get("dgncur_awet") is s.get("dgncur_a"):
            d["dgncur_awet"] = d["dgncur_a"]
        out.append(d)
    return out


def _time(cand, scenes, dt, precision, steps, graph):
    cfg = default_config(schedule=cand["schedule"], soa_blocks_per_cu=cand["soa_blocks_per_cu"])
    ch = Chain(config=cfg, precision=precision)
    ds = _copy(scenes)
    outs = [torch.empty((L.NSTATE,) + tuple(d["temperature"].shape), dtype=torch.float64,
                        device="cuda") for d in ds]
    b = ch.batch(ds, dt, outs, dgn_i=[d["dgncur_a"] for d in ds],
                 state_out=[d["state"] for d in ds])
    b.run()  # first use sizes the workspace
    torch.cuda.synchronize()
    run = b.run
    if graph:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            b.run()
        run = g.replay
        run()
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    b.close()
    ch.close()
    return ms


def tune(scenes, dt, precision=_abi.PREC_FP64, steps=5, graph=True, cache=None,
         candidates=None, verbose=False):
    """Return (best candidate dict, {candidate index: ms per step}).

    scenes: list of scene dicts of CUDA float64 tensors (not modified)."""
    sig = signature(scenes, dt, precision)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            known = json.load(f)
        if sig in known:
            return known[sig]["best"], {int(k): v for k, v in known[sig]["ms"].items()}
    cands = candidates or CANDIDATES
    ms = {}
    for i, c in enumerate(cands):
        ms[i] = _time(c, scenes, dt, precision, steps, graph)
        if verbose:
            print("tune: %s -> %.3f ms/step" % (c, ms[i]), flush=True)
    best = cands[min(ms, key=ms.get)]
    if cache:
        known = {}
        if os.path.exists(cache):
            with open(cache) as f:
                known = json.load(f)
        known[sig] = {"best": best, "ms": ms}
        with open(cache, "w") as f:
            json.dump(known, f, indent=1)
    return best, ms
