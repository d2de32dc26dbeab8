#!/usr/bin/env python3
"""Benchmark: MAM4 microphysics column-steps/s on one MI355X.

Headline workload (the study's precision configuration): 21,600 columns x 72
levels, seeded synthetic inputs (mam4xx_amd.scenes, seed 0), consecutive
1800 s steps 1..K from the seeded state: each step runs the whole chain +
calcsize through the C ABI and writes the tendencies, the calcsize diameters
and the updated state in place (the diameters feed the next step).  W warm-up
steps run first on the seeded state, which is then restored, so the K timed
steps are steps 1..K.  Inputs are resident in HBM before the timed region.

The FP64 build and the mixed (FP32-bulk) build run the same K steps.  The
metric admits a build at <= 1e-5 relative error: the mixed build's step-1
tendencies are compared with FP64's (state-relative measure of the parity
tests) and it becomes the headline only if every cell and field is within
1e-5 and it is faster; otherwise the headline is the FP64 figure.  The line
carries both builds' times and the error tables (step 1, and the state after
K steps), plus FP64 and mixed against the CPU restatement after K steps.

A step is four kernels (work-queue schedule, include/mam4x.h):
  set-up (per cell) -> queue build -> SOA work-queue kernel -> finish (per cell)
The library records HIP events on the launch stream between them while stats
are enabled (mam4x_last_phase_times), so each phase is timed live over the
timed steps.

Roofline (dominant kernel): the SOA work-queue kernel.  Its algorithmic work
is B FP64 flops per SOA sub-step times the sub-steps the GPU reports for the
timed launches; the rest of the step costs A flops per cell.  A and B are the
least-squares fit flops = A + B * substeps over a 216-column sample of this
workload, counted by the op-counting build of the CPU restatement
(oracle/opcount.cpp, driven by oracle/opcount_fit.py: + - * / sqrt = 1,
transcendentals at their gfx950 device-libm flop counts).  Bound: FP64 VALU
-- no step is a dense contraction, so MFMA is not used.

cpu_baseline: the CPU restatement (oracle/, OpenMP over columns) timed on the
host's cores for the same steps 1..K of the same scene, plus a single-core
figure on 216 columns.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NCOL, NLEV, DT = 21600, 72, 1800.0
# FP64 flops per cell (A) and per SOA sub-step (B); see module docstring.
OPS_PER_CELL = 11850.2
OPS_PER_SOA_STEP = 130.87
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# algorithmic HBM bytes per cell of a whole step: reads 52 state + 5 met + 12
# modal doubles, writes 52 tendencies + 52 updated state + 8 diameters
BYTES_PER_CELL = 8 * (52 + 5 + 12 + 52 + 52 + 8)
# algorithmic bytes of the SOA kernel per queued cell: 14 doubles of loop
# inputs and a 4-byte queue entry read, 6 doubles of results written
SOA_BYTES_PER_CELL = 14 * 8 + 4 + 6 * 8


def cpu_baseline(scene, nsteps, nthreads):
    """The CPU restatement (oracle/, OpenMP over columns) on the same workload:
    steps 1..nsteps of the full scene from its seeded state, chained as the
    GPU chains them, timed once after a one-step warm-up; plus a single-core
    figure on 216 columns.  Returns the baseline record and the oracle's
    final state (the reference of the parity figures)."""
    import numpy as np
    import oracle
    orc = oracle.Oracle()
    s = {k: np.array(v) for k, v in scene.items()}
    orc.chain_step(s, dt=DT, nthreads=nthreads)  # warm-up (page-in, threads)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        t, di, _, _ = orc.chain_step(s, dt=DT, nthreads=nthreads)
        s["state"] = s["state"] + DT * t
        s["dgncur_a"] = di
        s["dgncur_awet"] = di
    wall = time.perf_counter() - t0
    ncol1 = 216
    sub = {k: np.ascontiguousarray(v[..., :ncol1] if v.ndim >= 2 else v[:ncol1])
           for k, v in scene.items()}
    t0 = time.perf_counter()
    orc.chain_step(sub, dt=DT, nthreads=1)
    t1 = time.perf_counter() - t0
    rec = {
        "value": NCOL * nsteps / wall,
        "unit": "column-steps/s",
        "cores": nthreads,
        "kind": "port",
        "sample": "steps 1-%d of the full %d x %d scene from its seeded state (the GPU's timed "
                  "steps), chained, timed once after a one-step warm-up; single-core figure "
                  "from step 1 of 216 columns" % (nsteps, NCOL, NLEV),
        "wall_s": wall,
        "single_core_value": ncol1 / t1,
        "cpu": _cpu_model(),
    }
    return rec, s["state"]


def _rel_err(x, ref, scale):
    """|x - ref| / scale per element (numpy)"""
    import numpy as np
    return np.abs(x - ref) / scale


def _err_summary(e):
    """max / median / p99.9 of a [NSTATE, nlev, ncol] error array, the worst
    field and the fraction of cells with any field above 1e-5 and 1e-10"""
    import numpy as np
    from mam4xx_amd import layout as L
    cell = e.max(axis=0)
    per_field = e.reshape(e.shape[0], -1).max(axis=1)
    return {"max": float(e.max()), "median": float(np.median(e)),
            "p99.9": float(np.quantile(cell, 0.999)),
            "worst_field": L.field_names()[int(per_field.argmax())],
            "frac_cells_gt_1e-5": float((cell > 1e-5).mean()),
            "frac_cells_gt_1e-10": float((cell > 1e-10).mean())}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _traffic(name):
    """HBM bytes per launch from a committed counter summary, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        for line in open(path):
            if line.startswith("bytes_per_launch"):
                return float(line.split("=")[1])
    return None


def bench_batch(args):
    """32 uneven scenes (scenes.uneven_batch, seed 0) stepped by one launch
    sequence per step: eager, replayed from a CUDA graph, and autotuned."""
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd import layout as L
    from mam4xx_amd import scenes, tune
    from mam4xx_amd.chain import Chain, scene_to_device
    scs = scenes.uneven_batch(nscenes=32, seed=0)
    ncols = sum(s["temperature"].shape[1] for s in scs)
    ds = [scene_to_device(s) for s in scs]
    for d in ds:
        d["dgncur_awet"] = d["dgncur_a"]
    out = {"metric": "batched uneven scenes: column-steps/sec at 72 lev", "unit": "column-steps/s",
           "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "32 scenes, %d columns x 72 levels, column counts log-uniform "
                                  "in [96, 21600], chained 1800 s steps" % ncols,
                      "scenes": 32, "columns": ncols}}
    for name, sched in (("queue", _abi.SCHED_QUEUE), ("fused", _abi.SCHED_FUSED)):
        ch = Chain(schedule=sched)
        tends = [torch.empty((L.NSTATE,) + tuple(d["temperature"].shape), dtype=torch.float64,
                             device="cuda") for d in ds]
        b = ch.batch(ds, DT, tends, dgn_i=[d["dgncur_a"] for d in ds],
                     state_out=[d["state"] for d in ds])
        for _ in range(args.warmup):
            b.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.run()
        torch.cuda.synchronize()
        eager = (time.perf_counter() - t0) / args.steps
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            b.run()
        g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        graph = (time.perf_counter() - t0) / args.steps
        out[name] = {"eager_ms_per_step": eager * 1e3, "graph_ms_per_step": graph * 1e3,
                     "graph_column_steps_per_s": ncols / graph}
        b.close()
        ch.close()
    best, tried = tune.tune(ds, DT, steps=5)
    out["tuned"] = {"best": best, "tried": [[c, ms] for c, ms in tried]}
    out["value"] = out["queue"]["graph_column_steps_per_s"]
    print(json.dumps(out), flush=True)


def bench_large(args):
    """345,600 columns x 128 levels (top 10 Pa): a seeded 21,600 x 128 scene
    tiled 16 times across columns on the device (cells are independent, so
    the value distributions are kept).  FP64 and mixed forward, the tangent,
    and the device memory in use."""
    import numpy as np
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd import layout as L
    from mam4xx_amd import scenes
    from mam4xx_amd.chain import Chain, scene_to_device
    ncol0, nlev, rep = 21600, 128, 16
    free0, _ = torch.cuda.mem_get_info()
    d0 = scene_to_device(scenes.generate(ncol0, nlev, seed=0, ptop=10.0))
    tile = lambda v: v.repeat(rep) if v.dim() == 1 else v.repeat(*([1] * (v.dim() - 1) + [rep]))
    d = {k: tile(v).contiguous() for k, v in d0.items()}
    del d0
    ncol = ncol0 * rep
    tend = torch.empty((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda")
    dgc = torch.empty((4, nlev, ncol), dtype=torch.float64, device="cuda")
    out = {"metric": "large scene: column-steps/sec at 128 lev", "unit": "column-steps/s",
           "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "345,600 columns x 128 levels (21,600-column seeded scene "
                                  "tiled 16x), chained 1800 s steps",
                      "columns": ncol, "levels": nlev}}
    for name, prec in (("fp64", _abi.PREC_FP64), ("mixed", _abi.PREC_MIXED)):
        ch = Chain(precision=prec)
        st, dg = d["state"].clone(), d["dgncur_a"].clone()
        dd = dict(d, state=st, dgncur_a=dg, dgncur_awet=dg)
        ts = []
        for i in range(1 + args.steps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            ch.compute_tendencies(dd, DT, tend, dgn_i=dg, dgn_c=dgc, state_out=st)
            e1.record()
            torch.cuda.synchronize()
            if i:
                ts.append(e0.elapsed_time(e1))
        free1, _ = torch.cuda.mem_get_info()
        out[name] = {"ms_per_step": float(np.mean(ts)), "column_steps_per_s": ncol / (np.mean(ts) * 1e-3),
                     "device_memory_in_use_gb": (free0 - free1) / 1e9}
        del st, dg, dd
        ch.close()
    ch = Chain()
    dtend = torch.empty_like(tend)
    v = d["state"] * torch.randn_like(d["state"])
    ch.compute_tangent(d, DT, tend, dtend, dstate=v)
    times = {}
    for name, fn in (("tangent", lambda: ch.compute_tangent(d, DT, tend, dtend, dstate=v)),
                     ("forward", lambda: ch.compute_tendencies(d, DT, tend))):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        times[name] = e0.elapsed_time(e1)
    out["tangent"] = {"ms": times["tangent"], "forward_ms_same_input": times["forward"],
                      "ratio": times["tangent"] / times["forward"]}
    out["peak_torch_allocated_gb"] = torch.cuda.max_memory_allocated() / 1e9
    out["value"] = out["fp64"]["column_steps_per_s"]
    print(json.dumps(out), flush=True)


def bench_water(args):
    """Aerosol water uptake (mam4x_compute_water_uptake; --workload water) or
    the wet-size producer, E3SM calcsize then water uptake
    (mam4x_compute_wet_size; --workload wetsize), over the 21,600 x 72 scene:
    kernel time from HIP events on the launch stream, the HBM roofline of its
    compulsory traffic, and the oracle on the host cores."""
    wet = args.workload == "wetsize"
    import numpy as np
    import torch
    import ctypes
    from mam4xx_amd import _abi
    from mam4xx_amd import scenes
    from mam4xx_amd.chain import Chain, scene_to_device
    import oracle
    scene = scenes.generate(NCOL, NLEV, seed=0)
    d = scene_to_device(scene)
    shape = (4, NLEV, NCOL)
    out = [torch.empty(shape, dtype=torch.float64, device="cuda") for _ in range(3)]
    ch = Chain()
    s = torch.cuda.current_stream()
    inp = ch.make_inputs(d)
    # the argument blocks are built once, so the timed loop is one C call per
    # launch and the stream stays busy
    if wet:
        out.append(torch.empty(shape, dtype=torch.float64, device="cuda"))  # dry diameters
        wo = _abi.WetSizeOutputs()
        for m in range(4):
            wo.dgncur_awet[m], wo.qaerwat[m], wo.wetdens[m], wo.dgncur_i[m] = (
                o[m].data_ptr() for o in out)
        fn, h, sp = ch.lib.mam4x_compute_wet_size, ch.h, ctypes.c_void_p(s.cuda_stream)
        ri, ro = ctypes.byref(inp), ctypes.byref(wo)
        run = lambda: fn(h, DT, ri, ro, sp)
    else:
        wo = _abi.WaterOutputs()
        for m in range(4):
            wo.dgncur_awet[m], wo.qaerwat[m], wo.wetdens[m] = (o[m].data_ptr() for o in out)
        fn, h, sp = ch.lib.mam4x_compute_water_uptake, ch.h, ctypes.c_void_p(s.cuda_stream)
        ri, ro = ctypes.byref(inp), ctypes.byref(wo)
        run = lambda: fn(h, ri, ro, sp)
    assert run() == 0
    for _ in range(args.warmup):
        assert run() == 0
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(args.steps):
        run()
    e1.record(s)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    ms = e0.elapsed_time(e1) / args.steps
    ncell = NCOL * NLEV
    # compulsory traffic per cell -- water: 21 interstitial masses, T, p, qv,
    # cloud fraction and 4 dry diameters read, 4 x 3 outputs written;
    # wetsize: 50 numbers and masses, T, p, qv, cloud fraction read, 4 x 4
    # outputs written
    bytes_cell = ((50 + 4) * 8 + 16 * 8) if wet else ((21 + 4 + 4) * 8 + 12 * 8)
    achieved = ncell * bytes_cell / (ms * 1e-3) / 1e9
    what = ("wet-size producer (E3SM calcsize + water uptake)" if wet else
            "aerosol water uptake")
    res = {"metric": "%s: column-steps/sec at 72 lev" % what, "unit": "column-steps/s",
           "value": NCOL / (ms * 1e-3), "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
           "data": "synthetic", "steps": args.steps, "warmup": args.warmup,
           "config": {"workload": "21600 columns x 72 levels, %s per cell" % (
                          "compute_calcsize_and_water_uptake_dr" if wet else
                          "modal_aero_water_uptake_dr with wet density"),
                      "columns": NCOL, "levels": NLEV},
           "kernel_ms": ms, "ms_per_call_host_clock": wall_ms,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                        "frac": achieved / 8000.0,
                        "traffic": _traffic("%s_kernel_%dx%d.txt" % (
                            "wet_size" if wet else "water_uptake", NCOL, NLEV)),
                        "algorithmic_bytes_per_cell": bytes_cell}}
    if not args.no_cpu:
        orc = oracle.Oracle()
        nth = min(16, os.cpu_count() or 1)
        cpu = (lambda: orc.wet_size(scene, DT)) if wet else (lambda: orc.water_uptake(scene))
        cpu()
        best = 1e30
        for _ in range(5):
            t0 = time.perf_counter()
            cpu()
            best = min(best, time.perf_counter() - t0)
        res["cpu_baseline"] = {"value": NCOL / best, "unit": "column-steps/s", "cores": nth,
                               "kind": "port", "sample": "the full 21600 x 72 scene, best of 5 "
                               "after a warm-up (OpenMP over cells)", "cpu": _cpu_model()}
    ch.close()
    print(json.dumps(res), flush=True)


def bench_tracers(args):
    """The host model's wire format (--workload tracers): the E3SM tracer
    arrays state_q / qqcw of the 21,600 x 72 domain, (ncol, nlev, 40) stacked
    by column, to the 52 state planes in the EAMxx (ncol, nlev) layout
    (mam4x_tracers_to_fields) and back (mam4x_fields_to_tracers, state mode):
    the transfer around one step.  Kernel times from HIP events on the launch
    stream; HBM roofline of the values moved; the oracle's numpy walk
    (oracle/tracers.py) on one host core."""
    import ctypes
    import numpy as np
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    import oracle.tracers as T
    g = torch.Generator(device="cuda").manual_seed(0)
    q = torch.rand((NCOL, NLEV, T.PCNST), dtype=torch.float64, device="cuda", generator=g)
    c = torch.rand((NCOL, NLEV, T.PCNST), dtype=torch.float64, device="cuda", generator=g)
    fields = torch.empty((52, NCOL, NLEV), dtype=torch.float64, device="cuda")
    ch = Chain()
    s = torch.cuda.current_stream()
    tr = ch._tracers(q, c)
    fl = ch._fields(fields, NCOL, NLEV, _abi.LAYOUT_COL_LEV)
    h, sp, rt, rf = ch.h, ctypes.c_void_p(s.cuda_stream), ctypes.byref(tr), ctypes.byref(fl)
    unpack = lambda: ch.lib.mam4x_tracers_to_fields(h, rt, rf, sp)
    pack = lambda: ch.lib.mam4x_fields_to_tracers(h, rf, rt, _abi.PACK_STATE, sp)
    assert unpack() == 0 and pack() == 0
    for _ in range(args.warmup):
        unpack()
        pack()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t0 = time.perf_counter()
    ms_u = ms_p = 0.0
    for _ in range(args.steps):
        ev[0].record(s)
        unpack()
        ev[1].record(s)
        pack()
        ev[2].record(s)
        ev[2].synchronize()
        ms_u += ev[0].elapsed_time(ev[1])
        ms_p += ev[1].elapsed_time(ev[2])
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    ms_u /= args.steps
    ms_p /= args.steps
    ncell = NCOL * NLEV
    # values moved per cell and direction: 27 state_q + 25 qqcw tracers and
    # the 52 planes
    bytes_cell = (52 + 52) * 8
    roof = lambda ms, name: {
        "bound": "hbm", "achieved": ncell * bytes_cell / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": ncell * bytes_cell / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "traffic": _traffic("%s_kernel_%dx%d.txt" % (name, NCOL, NLEV)),
        "algorithmic_bytes_per_cell": bytes_cell}
    res = {"metric": "tracer-array transfer (state_q/qqcw <-> state planes): column-steps/sec "
                     "at 72 lev", "unit": "column-steps/s",
           "value": NCOL / ((ms_u + ms_p) * 1e-3), "n_gpus": 1, "higher_is_better": True,
           "dtype": "f64", "data": "synthetic", "steps": args.steps, "warmup": args.warmup,
           "config": {"workload": "21600 columns x 72 levels: state_q and qqcw (ncol, nlev, 40) "
                                  "to 52 (ncol, nlev) planes and back", "columns": NCOL,
                      "levels": NLEV},
           "ms_unpack": ms_u, "ms_pack": ms_p, "ms_per_round_host_clock": wall_ms,
           "roofline": roof(ms_u, "tracers_to_fields"),
           "roofline_pack": roof(ms_p, "fields_to_tracers")}
    if not args.no_cpu:
        # one host core: the reference's index walk over a 2,160-column sample
        nc = 2160
        qh = q[:nc].cpu().numpy().reshape(-1, T.PCNST)
        chh = c[:nc].cpu().numpy().reshape(-1, T.PCNST)
        best = 1e30
        for _ in range(3):
            t0 = time.perf_counter()
            progs = T.allocated_progs(qh.shape[0])
            T.inject_qqcw_to_prognostics(chh, progs)
            T.inject_stateq_to_prognostics(qh, progs)
            st = T.progs_to_state(progs)
            qo, co = qh.copy(), chh.copy()
            p2 = T.state_to_progs(st, [qh[:, 9], qh[:, 10], qh[:, 12], qh[:, 13]])
            T.extract_stateq_from_prognostics(p2, qo)
            T.extract_qqcw_from_prognostics(p2, co)
            best = min(best, time.perf_counter() - t0)
        res["cpu_baseline"] = {"value": nc / best, "unit": "column-steps/s", "cores": 1,
                               "kind": "port", "sample": "2160 of the 21600 columns x 72 levels, "
                               "inject + extract, best of 3 (numpy, one core)",
                               "cpu": _cpu_model()}
    ch.close()
    print(json.dumps(res), flush=True)


def _setsox_domain(n, seed=0):
    """synthetic in-cloud cells for the aqueous sulfur chemistry: met planes
    and the gas_pcnst tracer arrays qcw / qin [31, n]"""
    import numpy as np
    rng = np.random.default_rng(seed)
    T = rng.uniform(255, 300, n)
    p = rng.uniform(5e4, 1.01e5, n)
    met = dict(press=p, tfld=T, lwc=10 ** rng.uniform(-5, -3, n),
               cldfrc=rng.uniform(0.05, 1.0, n), cldnum=10 ** rng.uniform(7, 9, n),
               xhnm=p / (1.38e-23 * T) * 1e-6)
    met["cldfrc"][rng.uniform(size=n) < 0.3] = 0.0  # clear cells
    qcw, qin = np.zeros((31, n)), np.zeros((31, n))
    for i in (6, 14, 21):
        qcw[i] = 10 ** rng.uniform(-11, -9, n)
    for i in (13, 18, 26, 30):
        qcw[i] = 10 ** rng.uniform(9, 12, n)
    qin[0] = 10 ** rng.uniform(-8, -7.3, n)
    qin[1] = 10 ** rng.uniform(-10, -8.5, n)
    qin[2] = 10 ** rng.uniform(-13, -11, n)
    qin[3] = 10 ** rng.uniform(-11, -9, n)
    return met, qcw, qin


def bench_setsox(args):
    """Aqueous sulfur chemistry (mam4x_compute_setsox; --workload setsox)
    over 21,600 x 72 synthetic cells, 30 % cloud-free: kernel time from HIP
    events on the launch stream, HBM roofline of the values moved, the oracle
    on the host cores."""
    import ctypes
    import numpy as np
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    import oracle
    n = NCOL * NLEV
    met, qcw, qin = _setsox_domain(n)
    keys = {"pressure": "press", "temperature": "tfld", "liquid_water": "lwc",
            "cloud_fraction": "cldfrc", "droplet_number": "cldnum", "air_number_density": "xhnm"}
    dm = {k: torch.as_tensor(met[v].reshape(NLEV, NCOL)).cuda() for k, v in keys.items()}
    qc0 = torch.as_tensor(qcw.reshape(31, NLEV, NCOL)).cuda()
    qi0 = torch.as_tensor(qin.reshape(31, NLEV, NCOL)).cuda()
    qc, qi = qc0.clone(), qi0.clone()
    ch = Chain()
    s = torch.cuda.current_stream()
    a = _abi.SetsoxArgs()
    a.nlev, a.ncol, a.ldc, a.layout = NLEV, NCOL, NCOL, _abi.LAYOUT_LEV_COL
    for k in keys:
        setattr(a, k, dm[k].data_ptr())
    for i in range(31):
        a.qcw[i] = qc[i].data_ptr()
        a.qin[i] = qi[i].data_ptr()
    fn, h, sp, ra = ch.lib.mam4x_compute_setsox, ch.h, ctypes.c_void_p(s.cuda_stream), ctypes.byref(a)
    # each launch restarts from the same tracers (a device-to-device copy
    # outside the timed kernel interval), so every launch does the same work
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    for _ in range(args.warmup + 1):
        qc.copy_(qc0)
        qi.copy_(qi0)
        assert fn(h, DT, ra, sp) == 0
    torch.cuda.synchronize()
    ms = 0.0
    for _ in range(args.steps):
        qc.copy_(qc0)
        qi.copy_(qi0)
        ev[0].record(s)
        fn(h, DT, ra, sp)
        ev[1].record(s)
        ev[1].synchronize()
        ms += ev[0].elapsed_time(ev[1])
    ms /= args.steps
    # values per cell: 6 met + 4 gases + 3 cloud-borne sulfate + 3 numbers
    # read, 3 gases + 3 cloud-borne sulfate written
    bytes_cell = (6 + 4 + 6 + 3 + 3) * 8
    achieved = n * bytes_cell / (ms * 1e-3) / 1e9
    res = {"metric": "aqueous sulfur chemistry (setsox): column-steps/sec at 72 lev",
           "unit": "column-steps/s", "value": NCOL / (ms * 1e-3), "n_gpus": 1,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic", "steps": args.steps,
           "warmup": args.warmup,
           "config": {"workload": "21600 columns x 72 levels of synthetic in-cloud cells (30 % "
                                  "cloud-free), setsox_single_level per cell",
                      "columns": NCOL, "levels": NLEV},
           "kernel_ms": ms,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": _traffic("setsox_kernel_%dx%d.txt" % (NCOL, NLEV)),
                        "algorithmic_bytes_per_cell": bytes_cell}}
    if not args.no_cpu:
        orc = oracle.Oracle()
        nth = min(16, os.cpu_count() or 1)
        orc.setsox(DT, met, qcw, qin)
        best = 1e30
        for _ in range(5):
            t0 = time.perf_counter()
            orc.setsox(DT, met, qcw, qin)
            best = min(best, time.perf_counter() - t0)
        res["cpu_baseline"] = {"value": NCOL / best, "unit": "column-steps/s", "cores": nth,
                               "kind": "port", "sample": "the full 21600 x 72 domain, best of 5 "
                               "after a warm-up (OpenMP over cells)", "cpu": _cpu_model()}
    ch.close()
    print(json.dumps(res), flush=True)


def bench_gas_chem(args):
    """Gas-phase chemistry (mam4x_compute_gas_chemistry; --workload gaschem)
    over 21,600 x 72 synthetic cells: kernel time from HIP events on the
    launch stream, HBM roofline of the values moved, the oracle on the host
    cores."""
    import ctypes
    import numpy as np
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    import oracle
    n = NCOL * NLEV
    rng = np.random.default_rng(0)
    temp = rng.uniform(200, 310, n)
    photo = 10 ** rng.uniform(-8, -4, n)
    M = 101325.0 / (1.380649e-23 * 298.0) * 1e-6 * rng.uniform(0.1, 1.0, n)
    inv = np.stack([M, 0.78 * M, 0.21 * M, rng.uniform(0, 0.02, n) * M,
                    10 ** rng.uniform(5, 7, n), 10 ** rng.uniform(5, 8, n),
                    10 ** rng.uniform(7, 9, n), np.zeros(n)])
    extfrc = 10 ** rng.uniform(0, 6, (9, n))
    het = 10 ** rng.uniform(-8, -2, (31, n))
    qq = 10 ** rng.uniform(-14, -7, (31, n))
    sh = lambda a: torch.as_tensor(np.ascontiguousarray(a).reshape(a.shape[:-1] + (NLEV, NCOL)))
    d = {k: sh(v).cuda() for k, v in (("temp", temp), ("photo", photo), ("inv", inv),
                                       ("ext", extfrc), ("het", het))}
    q0 = sh(qq).cuda()
    q = q0.clone()
    ch = Chain()
    s = torch.cuda.current_stream()
    a = _abi.GasChemArgs()
    a.nlev, a.ncol, a.ldc, a.layout = NLEV, NCOL, NCOL, _abi.LAYOUT_LEV_COL
    a.temperature, a.photolysis_rate = d["temp"].data_ptr(), d["photo"].data_ptr()
    for i in range(9):
        a.extfrc[i] = d["ext"][i].data_ptr()
    for i in range(8):
        a.invariants[i] = d["inv"][i].data_ptr()
    for i in range(31):
        a.het_rates[i] = d["het"][i].data_ptr()
        a.qq[i] = q[i].data_ptr()
    fn, h, sp, ra = (ch.lib.mam4x_compute_gas_chemistry, ch.h, ctypes.c_void_p(s.cuda_stream),
                     ctypes.byref(a))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    for _ in range(args.warmup + 1):
        q.copy_(q0)
        assert fn(h, DT, ra, sp) == 0
    torch.cuda.synchronize()
    ms = 0.0
    for _ in range(args.steps):
        q.copy_(q0)  # every launch advances the same mixing ratios
        ev[0].record(s)
        fn(h, DT, ra, sp)
        ev[1].record(s)
        ev[1].synchronize()
        ms += ev[0].elapsed_time(ev[1])
    ms /= args.steps
    # values per cell: T, photolysis, 9 forcings, 5 invariants, 30 washout
    # rates and 30 mixing ratios read, 30 mixing ratios written
    bytes_cell = (2 + 9 + 5 + 30 + 30 + 30) * 8
    achieved = n * bytes_cell / (ms * 1e-3) / 1e9
    res = {"metric": "gas-phase chemistry (implicit solve, 30 species): column-steps/sec at 72 "
                     "lev", "unit": "column-steps/s", "value": NCOL / (ms * 1e-3), "n_gpus": 1,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic", "steps": args.steps,
           "warmup": args.warmup,
           "config": {"workload": "21600 columns x 72 levels of synthetic cells, "
                                  "gas_phase_chemistry per cell, dt 1800 s",
                      "columns": NCOL, "levels": NLEV},
           "kernel_ms": ms,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": _traffic("gas_chem_kernel_%dx%d.txt" % (NCOL, NLEV)),
                        "algorithmic_bytes_per_cell": bytes_cell}}
    if not args.no_cpu:
        orc = oracle.Oracle()
        nth = min(16, os.cpu_count() or 1)
        run = lambda: orc.gas_phase_chemistry(DT, temp, photo, extfrc, inv, het, qq)
        run()
        best = 1e30
        for _ in range(5):
            t0 = time.perf_counter()
            run()
            best = min(best, time.perf_counter() - t0)
        res["cpu_baseline"] = {"value": NCOL / best, "unit": "column-steps/s", "cores": nth,
                               "kind": "port", "sample": "the full 21600 x 72 domain, best of 5 "
                               "after a warm-up (OpenMP over cells)", "cpu": _cpu_model()}
    ch.close()
    print(json.dumps(res), flush=True)


def bench_linoz(args):
    """Stratospheric ozone (mam4x_compute_linoz; --workload linoz) over
    21,600 synthetic columns x 72 levels, with the invariant densities
    (mam4x_compute_invariants) of the same cells timed beside it: kernel
    times from HIP events on the launch stream, HBM roofline of the values
    moved, the oracle on the host cores."""
    import ctypes
    import numpy as np
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    import oracle
    rng = np.random.default_rng(0)
    shp = (NLEV, NCOL)
    ps = rng.uniform(95000, 103000, NCOL)
    edges = 10.0 * (ps / 10.0) ** (np.arange(NLEV + 1)[:, None] / NLEV)
    pmid, pdel = np.sqrt(edges[1:] * edges[:-1]), edges[1:] - edges[:-1]
    temp = np.where(pmid < 2e4, rng.uniform(183, 215, shp), rng.uniform(215, 300, shp))
    o3 = 10 ** rng.uniform(-8, -5, shp)
    col_du = np.cumsum(2.12e20 * pdel * o3, axis=0) / 2.687e16
    clim = np.stack([o3 * rng.uniform(0.7, 1.3, shp), temp + rng.normal(0, 5, shp),
                     col_du * rng.uniform(0.8, 1.2, shp), rng.normal(0, 1e-12, shp),
                     -10 ** rng.uniform(-8, -4.5, shp), rng.normal(0, 1e-13, shp),
                     rng.normal(0, 1e-14, shp), 10 ** rng.uniform(-7, -5, shp)])
    cols = dict(temp=temp, pmid=pmid, pdel=pdel, o3=o3, clim=clim,
                sza=rng.uniform(0, np.pi * 100 / 180, NCOL), rlats=rng.uniform(-1.5, 1.5, NCOL),
                col0=rng.uniform(0, 1e17, NCOL), ltrop=rng.integers(20, 50, NCOL))
    conf = dict(chlorine_loading=3.0, psc_T=193.0, o3_sfc=3e-8, o3_tau=172800.0, o3_lbl=4)
    qv = rng.uniform(0, 0.02, shp)
    offline = 10 ** rng.uniform(-12, -6, (4,) + shp)
    dev = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).cuda()
    d = {k: dev(v) for k, v in (("temp", temp), ("pmid", pmid), ("pdel", pdel), ("clim", clim),
                                ("sza", cols["sza"]), ("lat", cols["rlats"]),
                                ("col0", cols["col0"]), ("qv", qv), ("off", offline))}
    ltrop = dev(cols["ltrop"], torch.int32)
    o3_0 = dev(o3)
    o3_d = o3_0.clone()
    inv = torch.empty((8,) + shp, dtype=torch.float64, device="cuda")
    ch = Chain()
    s = torch.cuda.current_stream()
    a = _abi.LinozArgs()
    a.nlev, a.ncol, a.ldc, a.layout = NLEV, NCOL, NCOL, _abi.LAYOUT_LEV_COL
    a.temperature, a.pressure, a.pdel = (d[k].data_ptr() for k in ("temp", "pmid", "pdel"))
    for i in range(8):
        a.climatology[i] = d["clim"][i].data_ptr()
    a.zenith_angle, a.latitude, a.column_above_top = (d[k].data_ptr()
                                                      for k in ("sza", "lat", "col0"))
    a.ltrop = ltrop.data_ptr()
    for k in ("chlorine_loading", "psc_T", "o3_sfc", "o3_tau"):
        setattr(a, k, conf[k])
    a.o3_lbl = conf["o3_lbl"]
    a.o3 = o3_d.data_ptr()
    b = _abi.SetinvArgs()
    b.nlev, b.ncol, b.ldc, b.layout = NLEV, NCOL, NCOL, _abi.LAYOUT_LEV_COL
    b.temperature, b.pressure, b.vapor_mixing_ratio = (d[k].data_ptr()
                                                       for k in ("temp", "pmid", "qv"))
    for i in range(4):
        b.cnst_offline[i] = d["off"][i].data_ptr()
    for i in range(8):
        b.invariants[i] = inv[i].data_ptr()
    sp = ctypes.c_void_p(s.cuda_stream)
    launches = {"linoz": lambda: ch.lib.mam4x_compute_linoz(ch.h, DT, ctypes.byref(a), sp),
                "setinv": lambda: ch.lib.mam4x_compute_invariants(ch.h, ctypes.byref(b), sp)}
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ms = {}
    for name, fn in launches.items():
        for _ in range(args.warmup + 1):
            o3_d.copy_(o3_0)
            assert fn() == 0
        torch.cuda.synchronize()
        t = 0.0
        for _ in range(args.steps):
            o3_d.copy_(o3_0)  # every launch advances the same ozone
            ev[0].record(s)
            fn()
            ev[1].record(s)
            ev[1].synchronize()
            t += ev[0].elapsed_time(ev[1])
        ms[name] = t / args.steps
    n = NCOL * NLEV
    # linoz, compulsory values: every level reads pdel and the ozone (the
    # overhead column); the levels above the tropopause also read T, p and 8
    # climatology values and write the ozone; the surface-sink levels below
    # the tropopause write the ozone.  The optional diagnostics are not
    # requested.
    lt = np.minimum(cols["ltrop"], NLEV).astype(np.int64)
    lbl = conf["o3_lbl"]
    sink_only = np.maximum(0, lbl - np.maximum(0, lt - (NLEV - lbl)))
    bytes_launch = 8 * (2 * n + 11 * int(lt.sum()) + int(sink_only.sum()))
    bytes_cell = bytes_launch / n
    achieved = bytes_launch / (ms["linoz"] * 1e-3) / 1e9
    # setinv: T, p, qv, 4 offline species read, 8 invariants written
    inv_bytes = (3 + 4 + 8) * 8
    res = {"metric": "stratospheric ozone (linoz) + overhead column: column-steps/sec at 72 lev",
           "unit": "column-steps/s", "value": NCOL / (ms["linoz"] * 1e-3), "n_gpus": 1,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic", "steps": args.steps,
           "warmup": args.warmup,
           "config": {"workload": "21600 columns x 72 levels of synthetic columns, "
                                  "lin_strat_chem_solve above the tropopause (levels 20-49) "
                                  "and the 4-level surface sink, dt 1800 s",
                      "columns": NCOL, "levels": NLEV},
           "kernel_ms": ms["linoz"],
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": _traffic("linoz_kernel_%dx%d.txt" % (NCOL, NLEV)),
                        "kernel": "linoz_block_kernel",
                        "algorithmic_bytes_per_cell": bytes_cell},
           "setinv": {"kernel_ms": ms["setinv"],
                      "achieved_GBs": n * inv_bytes / (ms["setinv"] * 1e-3) / 1e9,
                      "traffic": _traffic("setinv_kernel_%dx%d.txt" % (NCOL, NLEV)),
                      "algorithmic_bytes_per_cell": inv_bytes}}
    if not args.no_cpu:
        orc = oracle.Oracle()
        nth = min(16, os.cpu_count() or 1)
        run = lambda: orc.linoz(DT, cols, conf)
        run()
        best = 1e30
        for _ in range(5):
            t0 = time.perf_counter()
            run()
            best = min(best, time.perf_counter() - t0)
        res["cpu_baseline"] = {"value": NCOL / best, "unit": "column-steps/s", "cores": nth,
                               "kind": "port", "sample": "the full 21600 x 72 domain, best of 5 "
                               "after a warm-up (OpenMP over columns)", "cpu": _cpu_model()}
    ch.close()
    print(json.dumps(res), flush=True)


# table_photo's level pass, per wavelength: two eight-corner sums (8 mul + 7
# add each), the pressure interpolation and the flux (4), the cross-section
# interpolation (3), the product and its sum (2); per cell the two corner
# weight sets (2 x 12) and the five weights' clamps (10)
PHOTO_OPS_PER_WAVELENGTH = 39
PHOTO_OPS_PER_CELL = 34
PHOTO_TABLE = dict(nw=67, nump=51, numsza=10, numcolo3=10, numalb=6, np_xs=15, numj=1)


def bench_photo(args):
    """Photolysis rates (mam4x_compute_photolysis, table_photo; --workload
    photo) over 21,600 synthetic columns x 72 levels with a table of 67
    wavelengths x 51 pressures x 10 zenith angles x 10 ozone ratios x 6
    albedos (synthetic values): one call (column pass + level pass) timed with
    HIP events on the launch stream, the oracle on the host cores."""
    import numpy as np
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd import photo_tables as PT
    from mam4xx_amd.chain import Chain
    import oracle
    rng = np.random.default_rng(0)
    T = PHOTO_TABLE
    ax = PT.wide_axes(nump=T["nump"], numsza=T["numsza"], numcolo3=T["numcolo3"],
                      numalb=T["numalb"], np_xs=T["np_xs"])
    tab = PT.random_table(rng, ax, nw=T["nw"], numj=T["numj"])
    cols = PT.columns(rng, NLEV, NCOL, ax)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    dtab = {k: dev(v) for k, v in tab.items()}
    dtab["lng_index"] = 0
    names = dict(pressure="pmid", pdel="pdel", temperature="temp", o3_column="colo3",
                 liquid_water="lwc", cloud_fraction="clouds")
    planes = {k: dev(cols[v]) for k, v in names.items()}
    zen, alb = dev(cols["zen_angle"]), dev(cols["srf_alb"])
    photo0 = dev(cols["photo"])
    photo = photo0.clone()
    ch = Chain()
    s = torch.cuda.current_stream()
    work = ch.compute_photolysis(dtab, planes, zen, alb, photo, esfact=1.0)  # sizes the work

    def call():
        ch.compute_photolysis(dtab, planes, zen, alb, photo, esfact=1.0, work=work)

    for _ in range(args.warmup + 1):
        photo.copy_(photo0)
        call()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    tot = 0.0
    for _ in range(args.steps):
        photo.copy_(photo0)  # every call updates the same rates
        ev[0].record(s)
        call()
        ev[1].record(s)
        ev[1].synchronize()
        tot += ev[0].elapsed_time(ev[1])
    ms = tot / args.steps
    day = np.rad2deg(cols["zen_angle"]) < 88.85
    ncell_day = int(day.sum()) * NLEV
    flops = ncell_day * (PHOTO_OPS_PER_WAVELENGTH * T["nw"] + PHOTO_OPS_PER_CELL)
    achieved = flops / (ms * 1e-3) / 1e12
    # HBM values per daylight cell: the column pass reads cloud fraction,
    # lwc and pdel and writes the two work planes; the level pass reads
    # pmid, T, the ozone column, the two work planes and the rate and writes
    # the rate (the tables stay in L2 and the infinity cache)
    bytes_cell = 8 * (3 + 2 + 6 + 1)
    res = {"metric": "photolysis rates (table_photo): column-steps/sec at 72 lev",
           "unit": "column-steps/s", "value": NCOL / (ms * 1e-3), "n_gpus": 1,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic", "steps": args.steps,
           "warmup": args.warmup,
           "config": {"workload": "21600 columns x 72 levels (%d in daylight), table %d "
                                  "wavelengths x %d pressures x %d zenith angles x %d ozone "
                                  "ratios x %d albedos, one photolysis reaction"
                                  % (int(day.sum()), T["nw"], T["nump"], T["numsza"],
                                     T["numcolo3"], T["numalb"]),
                      "columns": NCOL, "levels": NLEV},
           "call_ms": ms,
           "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                        "traffic": _traffic("photolysis_%dx%d.txt" % (NCOL, NLEV)),
                        "kernel": "photo_column_block_kernel + photo_level_kernel",
                        "algorithmic_gflop_per_call": flops / 1e9,
                        "hbm_algorithmic_gbs": ncell_day * bytes_cell / (ms * 1e-3) / 1e9,
                        "algorithmic_bytes_per_cell": bytes_cell}}
    if not args.no_cpu:
        orc = oracle.Oracle()
        nth = min(16, os.cpu_count() or 1)
        import oracle.pyoracle as P
        otab = P.photo_table(tab, lng_index=0)
        run = lambda: orc.table_photo(otab, cols, 1.0)
        run()
        best = 1e30
        for _ in range(3):
            t0 = time.perf_counter()
            run()
            best = min(best, time.perf_counter() - t0)
        res["cpu_baseline"] = {"value": NCOL / best, "unit": "column-steps/s", "cores": nth,
                               "kind": "port", "sample": "the full 21600 x 72 domain, best of 3 "
                               "after a warm-up (OpenMP over columns)", "cpu": _cpu_model()}
    ch.close()
    print(json.dumps(res), flush=True)


def bench_washout(args):
    """Washout rates (mam4x_compute_washout, sethet; --workload washout) over
    21,600 synthetic columns x 72 levels: kernel time from HIP events on the
    launch stream, HBM roofline of the values moved, the oracle on the host
    cores."""
    import ctypes
    import numpy as np
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    import oracle
    rng = np.random.default_rng(0)
    shp = (NLEV, NCOL)
    ps = rng.uniform(95000, 103000, NCOL)
    edges = 10.0 * (ps / 10.0) ** (np.arange(NLEV + 1)[:, None] / NLEV)
    press = np.sqrt(edges[1:] * edges[:-1])
    T = np.where(press < 2e4, rng.uniform(200, 220, shp), rng.uniform(220, 300, shp))
    cloudy = rng.uniform(size=shp) < 0.35
    cols = dict(press=press, zmid=7000.0 * np.log(ps / press), tfld=T,
                cmfdqr=np.where(cloudy, 10 ** rng.uniform(-10, -7, shp), 0.0),
                nrain=np.where(cloudy, 10 ** rng.uniform(-10, -7, shp), 0.0),
                nevapr=np.where(rng.uniform(size=shp) < 0.3, 10 ** rng.uniform(-10, -7.5, shp),
                                0.0),
                inv_m=press / (1.380649e-23 * T) * 1e-6, h2o2=10 ** rng.uniform(-10, -9, shp),
                so2=np.zeros(shp), rlat=rng.uniform(-1.5, 1.5, NCOL),
                phis=rng.uniform(0, 3e4, NCOL))
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).cuda()
    d = {k: dev(v) for k, v in cols.items()}
    het = torch.empty((31,) + shp, dtype=torch.float64, device="cuda")
    ch = Chain()
    s = torch.cuda.current_stream()
    a = _abi.SethetArgs()
    a.nlev, a.ncol, a.ldc, a.layout = NLEV, NCOL, NCOL, _abi.LAYOUT_LEV_COL
    for f, k in (("pressure", "press"), ("zmid", "zmid"), ("temperature", "tfld"),
                 ("cmfdqr", "cmfdqr"), ("nrain", "nrain"), ("nevapr", "nevapr"),
                 ("air_density", "inv_m"), ("h2o2", "h2o2"), ("latitude", "rlat"),
                 ("phis", "phis")):
        setattr(a, f, d[k].data_ptr())
    for i in range(31):
        a.het_rates[i] = het[i].data_ptr()
    fn = lambda: ch.lib.mam4x_compute_washout(ch.h, DT, ctypes.byref(a),
                                              ctypes.c_void_p(s.cuda_stream))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    for _ in range(args.warmup + 1):
        assert fn() == 0
    torch.cuda.synchronize()
    ms = 0.0
    for _ in range(args.steps):
        ev[0].record(s)
        fn()
        ev[1].record(s)
        ev[1].synchronize()
        ms += ev[0].elapsed_time(ev[1])
    ms /= args.steps
    n = NCOL * NLEV
    # per cell: pressure, zmid, T, three precipitation rates, M and H2O2 read;
    # 31 rates written
    bytes_cell = (8 + 31) * 8
    achieved = n * bytes_cell / (ms * 1e-3) / 1e9
    res = {"metric": "washout rates (sethet): column-steps/sec at 72 lev",
           "unit": "column-steps/s", "value": NCOL / (ms * 1e-3), "n_gpus": 1,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic", "steps": args.steps,
           "warmup": args.warmup,
           "config": {"workload": "21600 columns x 72 levels of synthetic columns, 35 % of "
                                  "levels precipitating, dt 1800 s",
                      "columns": NCOL, "levels": NLEV},
           "kernel_ms": ms,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": _traffic("sethet_kernel_%dx%d.txt" % (NCOL, NLEV)),
                        "algorithmic_bytes_per_cell": bytes_cell}}
    if not args.no_cpu:
        orc = oracle.Oracle()
        nth = min(16, os.cpu_count() or 1)
        run = lambda: orc.sethet(DT, cols)
        run()
        best = 1e30
        for _ in range(5):
            t0 = time.perf_counter()
            run()
            best = min(best, time.perf_counter() - t0)
        res["cpu_baseline"] = {"value": NCOL / best, "unit": "column-steps/s", "cores": nth,
                               "kind": "port", "sample": "the full 21600 x 72 domain, best of 5 "
                               "after a warm-up (OpenMP over columns)", "cpu": _cpu_model()}
    ch.close()
    print(json.dumps(res), flush=True)


def bench_budgets(args):
    """Column budgets of the 21,600 x 72 step (mam4x_outputs.column_budgets;
    --workload budgets), in both layouts.  The budget kernel runs after the
    step's finish kernel, so its time is the step's event time less the
    library's phase times (set-up, SOA, finish), recorded on the same
    stream; the same steps without budgets are timed beside it.  HBM
    roofline of the compulsory traffic: 52 tendency planes and dp read, 54
    values per column written.  Every step starts from the seeded state."""
    import numpy as np
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd import layout as L
    from mam4xx_amd import scenes
    from mam4xx_amd.chain import Chain, scene_to_device
    scene = scenes.generate(NCOL, NLEV, seed=0)
    nb = L.NSTATE + 2
    legs = {}
    for name, layout in (("lev_col", _abi.LAYOUT_LEV_COL), ("col_lev", _abi.LAYOUT_COL_LEV)):
        if layout == _abi.LAYOUT_COL_LEV:
            sc = {k: (np.ascontiguousarray(np.swapaxes(v, -1, -2)) if v.ndim >= 2 else v)
                  for k, v in scene.items()}
        else:
            sc = scene
        d = scene_to_device(sc)
        tend = torch.empty_like(d["state"])
        bud = torch.empty((nb, NCOL), dtype=torch.float64, device="cuda")
        ch = Chain(schedule=_abi.SCHED_QUEUE)
        s = torch.cuda.current_stream()

        def timed(with_budgets):
            kw = dict(budgets=bud) if with_budgets else {}
            for _ in range(args.warmup):
                ch.compute_tendencies(d, DT, tend, layout=layout, **kw)
            torch.cuda.synchronize()
            ch.enable_stats(True)
            ch.last_phase_times()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(args.steps):
                ch.compute_tendencies(d, DT, tend, layout=layout, **kw)
            e1.record(s)
            torch.cuda.synchronize()
            phases, n = ch.last_phase_times()
            ch.enable_stats(False)
            assert n == args.steps
            step = e0.elapsed_time(e1) / args.steps
            return step, step - sum(phases.values()) / args.steps

        step_b, budget_ms = timed(True)
        step_0, rest_0 = timed(False)
        legs[name] = {"step_ms_with_budgets": step_b, "step_ms_without": step_0,
                      "budget_kernel_ms": budget_ms, "gap_ms_without": rest_0,
                      "finite": bool(torch.isfinite(bud).all().item())}
        if name == "lev_col":
            tend_h = tend.cpu().numpy()
            dp_h = scene["hydrostatic_dp"]
        ch.close()
    bytes_cell = (L.NSTATE + 1) * 8 + nb * 8.0 / NLEV
    ms = legs["lev_col"]["budget_kernel_ms"]
    achieved = NCOL * NLEV * bytes_cell / (ms * 1e-3) / 1e9
    res = {"metric": "column budgets of the step: columns/sec at 72 lev", "unit": "columns/s",
           "value": NCOL / (ms * 1e-3), "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
           "data": "synthetic", "steps": args.steps, "warmup": args.warmup,
           "config": {"workload": "21600 columns x 72 levels: 54 FP64 column integrals of the "
                                  "step's tendencies (52 fields, sulfur and organic closures)",
                      "columns": NCOL, "levels": NLEV, "layouts": legs},
           "kernel_ms": ms,
           "roofline": {"bound": "hbm", "kernel": "column_budget_kernel", "achieved": achieved,
                        "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0,
                        "traffic": _traffic("column_budget_kernel_%dx%d.txt" % (NCOL, NLEV)),
                        "algorithmic_bytes_per_cell": bytes_cell}}
    if not args.no_cpu:
        # numpy over the same tendencies: 52 weighted level sums per column
        w = dp_h * (1.0 / 9.80616)
        best = 1e30
        for _ in range(5):
            t0 = time.perf_counter()
            (tend_h * w[None]).sum(axis=1)
            best = min(best, time.perf_counter() - t0)
        res["cpu_baseline"] = {"value": NCOL / best, "unit": "columns/s", "cores": 1,
                               "kind": "port", "sample": "numpy weighted level sums of the 52 "
                               "tendency planes of the full scene (no closures), best of 5",
                               "cpu": _cpu_model()}
    print(json.dumps(res), flush=True)


def bench_valu(args):
    """Scalar VALU rates of the forms the precision policies compile to
    (tests/cpp/valu_rates.hip, built with the library's flags: packed FP32
    disabled): FP64 / FP32 FMA and the device-libm exp and log."""
    import ctypes
    import numpy as np
    import torch
    lib_path = os.path.join(ROOT, "tests", "_build", "libvalurates.so")
    if not os.path.exists(lib_path):
        raise SystemExit("%s not built: run __graft_entry__.build()" % lib_path)
    torch.cuda.init()
    lib = ctypes.CDLL(lib_path)
    lib.valu_rates.restype = ctypes.c_int
    ops, chk = np.zeros(6), np.zeros(6)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.valu_rates(ops.ctypes.data_as(dp), chk.ctypes.data_as(dp))
    assert rc == 0 and np.all(np.isfinite(chk)), (rc, chk)
    names = ("fma_f64", "fma_f32", "exp_f64", "exp_f32", "log_f64", "log_f32")
    rates = {n: float(v) for n, v in zip(names, ops)}
    out = {
        "metric": "scalar VALU operations/s (FP64 vs FP32 forms of the chain's precision policies)",
        "value": rates["fma_f64"] * 2 / 1e12,
        "unit": "TFLOP/s (FP64 FMA)",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": "8192 blocks x 256 lanes x 8 independent chains; packed FP32 "
                               "forms disabled (the library's flags)"},
        "rates_gops": {n: v / 1e9 for n, v in rates.items()},
        "fma_tflops": {"f64": rates["fma_f64"] * 2 / 1e12, "f32": rates["fma_f32"] * 2 / 1e12},
        "peak_tflops": {"f64_vector": FP64_PEAK_TFLOPS, "f32_vector_packed": 157.3},
        "f32_over_f64": {"fma": rates["fma_f32"] / rates["fma_f64"],
                         "exp": rates["exp_f32"] / rates["exp_f64"],
                         "log": rates["log_f32"] / rates["log_f64"]},
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    """Derivatives of the step (--workload adjoint) over the 21,600 x 72
    scene: the FP64 step (mam4x_compute_tendencies), the tangent along one
    direction (mam4x_compute_tangent) and the reverse sweep with every
    gradient plane (mam4x_compute_adjoint), each timed over K calls from the
    same seeded state.  The reverse sweep synchronises its stream once per
    call (to size its sub-step records), so the host clock is the figure."""
    import torch
    from mam4xx_amd import layout as L
    from mam4xx_amd import scenes
    from mam4xx_amd.chain import Chain, scene_to_device
    scene = scenes.generate(NCOL, NLEV, seed=0)
    d = scene_to_device(scene)
    ch = Chain(ntsubstep=args.ntsubstep)
    shape = (L.NSTATE, NLEV, NCOL)
    g = torch.Generator(device="cuda").manual_seed(1)
    tend = torch.empty(shape, dtype=torch.float64, device="cuda")
    dtend = torch.empty_like(tend)
    v = d["state"] * torch.randn(shape, dtype=torch.float64, device="cuda", generator=g)
    vT = torch.randn((NLEV, NCOL), dtype=torch.float64, device="cuda", generator=g)
    w = torch.randn(shape, dtype=torch.float64, device="cuda", generator=g)
    w = w / d["state"].abs().clamp_min(1e-30)
    gs = torch.empty_like(tend)
    gT, gp, gq = (torch.empty((NLEV, NCOL), dtype=torch.float64, device="cuda")
                  for _ in range(3))
    runs = {
        "forward": lambda: ch.compute_tendencies(d, DT, tend),
        "tangent": lambda: ch.compute_tangent(d, DT, tend, dtend, dstate=v, dtemperature=vT),
        "adjoint": lambda: ch.compute_adjoint(d, DT, w, state_bar=gs, temperature_bar=gT,
                                              pressure_bar=gp, qv_bar=gq),
    }
    stream = torch.cuda.current_stream()
    ms = {}
    for name, run in runs.items():
        for _ in range(args.warmup):
            run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            run()
        e1.record(stream)
        torch.cuda.synchronize()
        ms[name] = {"host_ms": (time.perf_counter() - t0) * 1e3 / args.steps,
                    "event_ms": e0.elapsed_time(e1) / args.steps}
    finite = all(bool(torch.isfinite(x).all().item()) for x in (gs, gT, gp, gq))
    a = ms["adjoint"]["host_ms"]
    out = {
        "metric": "reverse-mode (vector-Jacobian product) column-steps/sec at 72 lev, FP64",
        "value": NCOL / (a * 1e-3),
        "unit": "column-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": a,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator of the reference study's value distributions)",
        "config": {"workload": "21600 columns x 72 levels, one 1800 s step with %d condensation "
                               "sub-step(s) from the seeded state; cotangent of all 52 "
                               "tendencies, gradients w.r.t. the state, T, p and qv"
                               % args.ntsubstep, "columns": NCOL, "levels": NLEV,
                   "ntsubstep": args.ntsubstep},
        "times_ms": ms,
        "cost_vs_forward": {k: ms[k]["host_ms"] / ms["forward"]["host_ms"]
                            for k in ("tangent", "adjoint")},
        "gradients_finite": finite,
    }
    ch.close()
    print(json.dumps(out), flush=True)


def _leg(args, name, prec, sched, d, state0, dgn0):
    """One precision's timed run: W warm-up steps from the seeded state, the
    state restored, an untimed step 1 (its tendencies kept), the state
    restored, then the K timed steps 1..K.  Returns times, the library's
    phase times and SOA sub-step count, the step-1 tendencies and the state
    after K steps."""
    import torch
    from mam4xx_amd import layout as L
    from mam4xx_amd.chain import Chain
    ch = Chain(precision=prec, schedule=sched)
    tend = torch.empty((L.NSTATE, NLEV, NCOL), dtype=torch.float64, device="cuda")
    dgc = torch.empty((4, NLEV, NCOL), dtype=torch.float64, device="cuda")

    def restore():
        d["state"].copy_(state0)
        d["dgncur_a"].copy_(dgn0)

    def step():
        ch.compute_tendencies(d, DT, tend, dgn_i=d["dgncur_a"], dgn_c=dgc, state_out=d["state"])

    for _ in range(args.warmup):
        step()
    restore()
    step()
    tend1 = tend.clone()
    restore()
    torch.cuda.synchronize()
    ch.enable_stats(True)
    ch.last_stats()  # reset the counters
    ch.last_phase_times()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    iters = ch.last_stats()
    phases, ncalls = ch.last_phase_times()
    assert ncalls == args.steps
    res = {"name": name, "wall": wall, "gpu_ms": e0.elapsed_time(e1) / args.steps,
           "phases": {k: v / args.steps for k, v in phases.items()},
           "substeps": iters / args.steps, "tend1": tend1, "stateK": d["state"].clone(),
           "finite": bool(torch.isfinite(d["state"]).all().item())}
    ch.close()
    restore()
    torch.cuda.synchronize()
    return res


def bench_headline(args):
    """Config 2 of the study: 21,600 x 72, consecutive 1800 s steps 1..K from
    the seeded state, FP64 and mixed builds timed on the same steps; the
    mixed build's error against FP64 decides which figure is the headline
    (the metric admits a build only at <= 1e-5 relative error)."""
    import numpy as np
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd import layout as L
    from mam4xx_amd import scenes
    from mam4xx_amd.chain import scene_to_device
    scene = scenes.generate(NCOL, NLEV, seed=0)
    d = scene_to_device(scene)
    # chained steps: calcsize diameters feed the next step's dry and wet
    # diameters (no water uptake in this workload)
    d["dgncur_awet"] = d["dgncur_a"]
    state0 = d["state"].clone()
    dgn0 = d["dgncur_a"].clone()
    sched = {"auto": _abi.SCHED_AUTO, "fused": _abi.SCHED_FUSED,
             "queue": _abi.SCHED_QUEUE}[args.schedule]
    legs = {"fp64": _leg(args, "fp64", _abi.PREC_FP64, sched, d, state0, dgn0)}
    if not args.no_mixed:
        legs["mixed"] = _leg(args, "mixed", _abi.PREC_MIXED, sched, d, state0, dgn0)
    ncell = NCOL * NLEV
    st0 = scene["state"]
    floor = np.where(np.arange(L.NSTATE) < 8, 1e-10, 1e-30)[:, None, None]
    f = legs["fp64"]
    t1_f = f["tend1"].cpu().numpy()
    sK_f = f["stateK"].cpu().numpy()
    precision = {}
    if "mixed" in legs:
        m = legs["mixed"]
        # step-1 tendencies, in the state-relative measure of the parity tests:
        # |d tend| dt / max(|q_in|, |q_post|, floor)
        den1 = np.maximum(np.maximum(np.abs(st0), np.abs(st0 + DT * t1_f)), floor)
        e1 = np.abs(m["tend1"].cpu().numpy() - t1_f) * DT / den1
        # state after K chained steps, relative to the quantity's magnitude
        denK = np.maximum(np.maximum(np.abs(sK_f), np.abs(st0)), floor)
        eK = np.abs(m["stateK"].cpu().numpy() - sK_f) / denK
        gate = float(e1.max()) <= 1e-5
        precision["mixed_vs_fp64"] = {
            "step1_tendencies": _err_summary(e1),
            "state_after_%d_steps" % args.steps: _err_summary(eK),
            "speed_ratio": f["wall"] / m["wall"],
            "gate_1e-5_on_step1_tendencies": gate,
        }
    # the headline build: the faster build that meets the metric's 1e-5 gate
    # (FP64 is the reference precision and always admitted)
    pick = "fp64"
    if precision.get("mixed_vs_fp64", {}).get("gate_1e-5_on_step1_tendencies") and \
            legs["mixed"]["wall"] < f["wall"]:
        pick = "mixed"
    h = legs[pick]
    ph = h["phases"]
    substeps = h["substeps"]
    step_flops = OPS_PER_CELL * ncell + OPS_PER_SOA_STEP * substeps
    queue = ph["soa"] > 0.0
    if queue:
        kern = {"name": "soa_queue_kernel", "ms": ph["soa"], "flops": OPS_PER_SOA_STEP * substeps,
                "traffic": _traffic("traffic_soa_queue_kernel_%dx%d.txt" % (NCOL, NLEV))}
    else:
        kern = {"name": "chain_fused_kernel", "ms": ph["finish"], "flops": step_flops,
                "traffic": _traffic("traffic_chain_fused_kernel_%dx%d.txt" % (NCOL, NLEV))}
    achieved = kern["flops"] / (kern["ms"] * 1e-3) / 1e12
    out = {
        "metric": "microphysics column-steps/sec at 72 lev: mixed vs FP64 GPU at \u22641e-5 rel err",
        "value": NCOL * args.steps / h["wall"],
        "unit": "column-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if pick == "fp64" else "f32+f64",
        "data": "synthetic (seeded generator of the reference study's value distributions)",
        "config": {
            "workload": "%d columns x %d levels, consecutive %g s microphysics steps 1-%d from the "
                        "seeded state (gasaerexch, rename, newnuc, coag, pcarbon aging, calcsize)"
                        % (NCOL, NLEV, DT, args.steps),
            "columns": NCOL, "levels": NLEV, "dt_s": DT, "build": pick,
            "schedule": args.schedule,
            "soa_substeps_per_cell": substeps / ncell,
            "state_finite": h["finite"],
            "fp64": {"ms_per_step": f["wall"] * 1e3 / args.steps,
                     "column_steps_per_s": NCOL * args.steps / f["wall"]},
        },
        "roofline": {
            "bound": "valu",
            "compute_unit": "FP64 VALU (no step is a dense contraction; MFMA unused)",
            "kernel": kern["name"],
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": kern["traffic"],
            "kernel_ms": kern["ms"],
            "algorithmic_gflop_per_launch": kern["flops"] / 1e9,
            "step": {
                "gpu_ms": h["gpu_ms"],
                "phase_ms": ph,
                "algorithmic_gflop": step_flops / 1e9,
                "tflops": step_flops / (h["gpu_ms"] * 1e-3) / 1e12,
                "hbm_algorithmic_gbs": BYTES_PER_CELL * ncell / (h["gpu_ms"] * 1e-3) / 1e9,
            },
        },
        "precision": precision,
    }
    if "mixed" in legs:
        m = legs["mixed"]
        out["config"]["mixed"] = {"ms_per_step": m["wall"] * 1e3 / args.steps,
                                  "column_steps_per_s": NCOL * args.steps / m["wall"]}
    if queue:
        # overcast cells (cloud fraction > 0.999) have no clear subarea and
        # never enter the queue (mam4_amicphys.hpp:241-317)
        nqueued = int((scene["cloud_fraction"] <= 0.999).sum())
        out["roofline"]["queued_cells"] = nqueued
        out["roofline"]["hbm_algorithmic_gbs"] = (SOA_BYTES_PER_CELL * nqueued /
                                                  (kern["ms"] * 1e-3) / 1e9)
    if not args.no_cpu:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        rec, sK_o = cpu_baseline(scene, args.steps, nthreads)
        out["cpu_baseline"] = rec
        # FP64 drop-in against the CPU restatement after the same K steps
        denK = np.maximum(np.maximum(np.abs(sK_o), np.abs(st0)), floor)
        precision["fp64_vs_oracle"] = {
            "state_after_%d_steps" % args.steps: _err_summary(np.abs(sK_f - sK_o) / denK)}
        if "mixed" in legs:
            precision["mixed_vs_oracle"] = {
                "state_after_%d_steps" % args.steps:
                    _err_summary(np.abs(legs["mixed"]["stateK"].cpu().numpy() - sK_o) / denK)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--schedule", choices=["auto", "fused", "queue"], default="auto")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--ntsubstep", type=int, default=1,
                    help="adjoint workload: condensation sub-steps per step")
    ap.add_argument("--no-mixed", action="store_true",
                    help="headline: time the FP64 build only")
    ap.add_argument("--workload", choices=["headline", "batch", "large", "water", "wetsize",
                                           "tracers", "setsox", "gaschem", "linoz", "washout",
                                           "photo", "valu", "adjoint", "budgets"],
                    default="headline",
                    help="headline: the metric's 21,600 x 72 step (default); batch, large, "
                         "water (aerosol water uptake), wetsize (calcsize + water uptake), "
                         "tracers (state_q/qqcw <-> state planes), setsox (aqueous sulfur "
                         "chemistry), gaschem (gas-phase chemistry) and linoz (stratospheric "
                         "ozone, with the invariant densities), washout (sethet), photo "
                         "(table_photo), budgets (column budgets of the step), valu "
                         "(scalar FP64 / FP32 VALU rates) and adjoint (tangent and reverse "
                         "sweep against the step) print their own line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the chain has no cross-GPU step and this build "
              "runs on one MI355X only (--gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd import layout as L
    from mam4xx_amd import scenes
    from mam4xx_amd.chain import Chain, scene_to_device

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "large":
        return bench_large(args)
    if args.workload in ("water", "wetsize"):
        return bench_water(args)
    if args.workload == "tracers":
        return bench_tracers(args)
    if args.workload == "setsox":
        return bench_setsox(args)
    if args.workload == "linoz":
        return bench_linoz(args)
    if args.workload == "washout":
        return bench_washout(args)
    if args.workload == "photo":
        return bench_photo(args)
    if args.workload == "gaschem":
        return bench_gas_chem(args)
    if args.workload == "valu":
        return bench_valu(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "budgets":
        return bench_budgets(args)
    return bench_headline(args)

if __name__ == "__main__":
    main()
