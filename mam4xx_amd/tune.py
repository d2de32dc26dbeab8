"""Autotuner for the launch configuration of a microphysics step.

The results of a step do not depend on its launch configuration (fused and
work-queue schedules are bitwise identical), only its speed does, and the
best choice depends on the workload: the number and size of the scenes, the
share of cells with an SOA loop and the spread of their sub-step counts.
``tune`` times candidates on copies of the caller's own scenes and returns
the fastest configuration, searching one knob at a time from the best so far:

1. schedule x soa_blocks_per_cu (CANDIDATES): the fused one-lane-per-cell
   kernel, or the work-queue pipeline with 1, 2 or 3 (the default) resident
   blocks per CU of the persistent SOA kernel, or as many as fit (fewer waves
   per SIMD raise each wave's issue share on the long serial chains, more
   waves hide latency);
2. the queue's granularity (queue schedule only): entries per refill of the
   long part (soa_chunk) and of the short part (soa_short_chunk), and the
   bucket that splits them (soa_short_bucket);
3. levels_per_wave: the per-cell kernels' tiles of 2, 4 or 8 levels per
   wave instead of 64 columns of one level.

Each candidate is timed as a batch (one launch sequence for all scenes) over
a few chained steps, eagerly or replayed from a CUDA graph.  Results are
cached per workload signature in a JSON file when ``cache`` is given.
"""
import hashlib
import json
import os
import time

import torch

from . import _abi
from . import layout as L
from .chain import Chain, default_config

DEFAULTS = {"soa_chunk": 0, "soa_short_chunk": 0, "soa_short_bucket": 0, "levels_per_wave": 1}
# soa_blocks_per_cu: 0 is the library's default (3), 8 means as many as fit
CANDIDATES = [dict(DEFAULTS, schedule=sc, soa_blocks_per_cu=b) for sc, b in (
    (_abi.SCHED_FUSED, 0), (_abi.SCHED_QUEUE, 0), (_abi.SCHED_QUEUE, 1), (_abi.SCHED_QUEUE, 2),
    (_abi.SCHED_QUEUE, 8))]
# stage 2 (queue schedule) and stage 3: one knob at a time
QUEUE_KNOBS = [("soa_chunk", (32, 128)), ("soa_short_chunk", (128, 512)),
               ("soa_short_bucket", (8, 12))]
TILE_KNOB = ("levels_per_wave", (2, 4, 8))


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
    cfg = default_config(**cand)
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
    """Return (best candidate dict, [(candidate, ms per step), ...] in the
    order tried).  candidates: a list to time instead of the staged search.

    scenes: list of scene dicts of CUDA float64 tensors (not modified)."""
    sig = signature(scenes, dt, precision)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            known = json.load(f)
        if sig in known:
            return known[sig]["best"], [tuple(t) for t in known[sig]["tried"]]
    tried = []

    def time_it(c):
        for t in tried:
            if t[0] == c:
                return t[1]
        ms = _time(c, scenes, dt, precision, steps, graph)
        tried.append((dict(c), ms))
        if verbose:
            print("tune: %s -> %.3f ms/step" % (c, ms), flush=True)
        return ms

    if candidates:
        for c in candidates:
            time_it(c)
    else:
        for c in CANDIDATES:
            time_it(c)
        best = min(tried, key=lambda t: t[1])
        knobs = (QUEUE_KNOBS if best[0]["schedule"] == _abi.SCHED_QUEUE else []) + [TILE_KNOB]
        for name, values in knobs:
            base = dict(best[0])
            for v in values:
                time_it(dict(base, **{name: v}))
            best = min(tried, key=lambda t: t[1])
    best = min(tried, key=lambda t: t[1])[0]
    if cache:
        known = {}
        if os.path.exists(cache):
            with open(cache) as f:
                known = json.load(f)
        known[sig] = {"best": best, "tried": tried}
        with open(cache, "w") as f:
            json.dump(known, f, indent=1)
    return best, tried
