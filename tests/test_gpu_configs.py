"""The library on the study's configurations at their full sizes.

* Ten chained 1800 s steps of 21,600 x 72 (the precision study's scene) in
  FP64: 120 sampled columns end within the parity bar of the oracle run on
  those columns alone, chained the same way, except for the cells whose
  adaptive SOA sub-stepping is ill-conditioned (a last-bit difference changes
  the step sequence and the error then persists over the steps).
* The mixed build on the same scene, step 1: its error against the oracle,
  per output, stays within the profile it is documented with (it does not
  meet 1e-5 everywhere; that is measured, not hidden).
* The 32-scene uneven batch (149,795 columns x 72 levels): sampled columns of
  every scene agree with the oracle, every cell conserves every species
  family, and a graph-captured run replays bit for bit the eager run.
* 345,600 columns x 128 levels generated at that size (model top 10 Pa):
  sampled columns agree with the oracle at 128 levels, every cell conserves
  every species family, two runs agree bit for bit, and the mixed build runs
  finite.
Conservation and bitwise checks run on the device; only sampled columns
come back to the host.
"""
import numpy as np
import pytest

from mam4xx_amd import layout as L
from mam4xx_amd import scenes

from .parity import state_rel_err

pytestmark = pytest.mark.gpu
DT = 1800.0


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


def _columns(sc, cols):
    ncol = sc["temperature"].shape[1]
    return {k: (np.ascontiguousarray(v[..., cols]) if np.shape(v)[-1] == ncol else v)
            for k, v in sc.items()}


def _sample(ncol, n=120):
    return np.unique(np.linspace(3, ncol - 1, n).astype(int))


def _family_residual(torch, state, tend):
    """max over cells of |d(total)/dt| dt / total for every species family
    (SO4 + H2SO4, SOA + SOAG in carbon units, POM, BC, dust, NaCl, MOM), in
    moles, on the device"""
    names = L.field_names()
    worst = {}
    for sp, gas, gw, w in (("so4", L.F_H2SO4, 1.0, 1.0), ("soa", L.F_SOAG, 12.011 / 150.0, 0.08),
                           ("pom", None, 0, 0.08), ("bc", None, 0, 1.0), ("dst", None, 0, 1.0),
                           ("ncl", None, 0, 1.0), ("mom", None, 0, 1.0)):
        fs = [f for f in range(8, 50) if names[f].startswith(sp + "_")]
        total = sum(w * L.to_vmr(f) * state[f].abs() for f in fs)
        change = sum(w * L.to_vmr(f) * tend[f] for f in fs)
        if gas is not None:
            total = total + gw * L.to_vmr(gas) * state[gas].abs()
            change = change + gw * L.to_vmr(gas) * tend[gas]
        rel = change.abs() * DT / total.clamp_min(1e-300)
        worst[sp] = float(rel.max().item())
    return worst


def _sample_bar(orc, sc, cols, tg, dt=DT):
    """the sampled columns against the oracle (bar of test_gpu_parity.py)"""
    sub = _columns(sc, cols)
    to, env = orc.envelope(sub, dt=dt)
    _, err = state_rel_err(tg, to, sub["state"], dt)
    cell_env = env.max(axis=0)
    bound = np.where(cell_env > 1e-11, np.maximum(1e-10, 10 * cell_env), 1e-10)[None]
    assert (err <= bound).all(), (err.max(), np.argwhere(err > bound)[:3])
    assert (cell_env > 1e-10).mean() < 0.01
    return to


# --------------------------------------------------------------- config 2 --

def test_precision_study_fp64_ten_steps_sampled(orc, torch_cuda):
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    ncol, nlev, nsteps = 21600, 72, 10
    sc = scenes.generate(ncol, nlev, seed=0)
    cols = _sample(ncol)
    print("config 2: %d x %d, %d chained steps, FP64" % (ncol, nlev, nsteps), flush=True)
    d = scene_to_device(sc)
    d["dgncur_awet"] = d["dgncur_a"]
    ch = Chain()
    tend = torch.empty((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda")
    dgc = torch.empty((4, nlev, ncol), dtype=torch.float64, device="cuda")
    for _ in range(nsteps):
        ch.compute_tendencies(d, DT, tend, dgn_i=d["dgncur_a"], dgn_c=dgc, state_out=d["state"])
    torch.cuda.synchronize()
    ch.close()
    assert bool(torch.isfinite(d["state"]).all())
    got = d["state"].cpu().numpy()[..., cols]
    # the oracle chained the same way on the sampled columns
    s = _columns(sc, cols)
    s = {k: np.array(v) for k, v in s.items()}
    for _ in range(nsteps):
        t, di, _, _ = orc.chain_step(s, dt=DT)
        s["state"] = s["state"] + DT * t
        s["dgncur_a"] = di
        s["dgncur_awet"] = di
    st0 = _columns(sc, cols)["state"]
    floor = np.where(np.arange(L.NSTATE) < 8, 1e-10, 1e-30)[:, None, None]
    den = np.maximum(np.maximum(np.abs(s["state"]), np.abs(st0)), floor)
    err = (np.abs(got - s["state"]) / den).max(axis=0)
    print("  10-step state error vs oracle: median %.2e, cells > 1e-10: %.4f %%, > 1e-5: %.4f %%, "
          "max %.2e" % (np.median(err), 100 * (err > 1e-10).mean(), 100 * (err > 1e-5).mean(),
                        err.max()), flush=True)
    # most cells agree to round-off; the ill-conditioned SOA cells carry a
    # last-bit difference into another sub-step sequence and keep it
    assert np.median(err) < 1e-12
    assert (err > 1e-10).mean() < 2e-3
    assert (err > 1e-5).mean() < 5e-4


# mixed build, step 1, state-relative error against the oracle per output
# class: median, 99.9th percentile, fraction of cells above 1e-5
MIXED_PROFILE = {"numbers": (1e-7, 2e-3, 0.05), "masses": (1e-7, 2e-3, 0.05),
                 "gases": (1e-7, 1e-3, 0.02)}


def test_precision_study_mixed_step_one_sampled(orc, torch_cuda):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    ncol, nlev = 21600, 72
    sc = scenes.generate(ncol, nlev, seed=0)
    cols = _sample(ncol, 240)
    print("config 2: %d x %d, step 1, mixed" % (ncol, nlev), flush=True)
    d = scene_to_device(sc)
    ch = Chain(precision=_abi.PREC_MIXED)
    tend = torch.empty((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda")
    ch.compute_tendencies(d, DT, tend)
    torch.cuda.synchronize()
    ch.close()
    assert bool(torch.isfinite(tend).all())
    tg = tend.cpu().numpy()[..., cols]
    sub = _columns(sc, cols)
    to = orc.chain_step(sub, dt=DT)[0]
    _, err = state_rel_err(tg, to, sub["state"], DT)
    names = L.field_names()
    for f in range(L.NSTATE):
        ef = err[f]
        print("  mixed %-12s median %.2e  p99.9 %.2e  max %.2e" % (
            names[f], np.median(ef), np.quantile(ef, 0.999), ef.max()), flush=True)
    stats = {}
    for cls, fields in (("numbers", range(0, 8)), ("masses", range(8, 50)),
                        ("gases", range(50, 52))):
        e = err[list(fields)]
        cell = e.max(axis=0)
        stats[cls] = (float(np.median(e)), float(np.quantile(cell, 0.999)),
                      float((cell > 1e-5).mean()))
        print("  mixed %-7s median %.2e  p99.9 %.2e  cells > 1e-5 %.3f %%  max %.2e"
              % (cls, stats[cls][0], stats[cls][1], 100 * stats[cls][2], e.max()), flush=True)
    for cls, (med, p999, frac) in stats.items():
        bmed, bp, bfrac = MIXED_PROFILE[cls]
        assert med < bmed and p999 < bp and frac < bfrac, (cls, stats[cls])


# --------------------------------------------------------------- config 4 --

def test_uneven_batch_full_size(orc, torch_cuda):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    scs = scenes.uneven_batch(32, seed=0)
    ntot = sum(s["temperature"].shape[1] for s in scs)
    print("config 4: 32 scenes, %d columns x 72" % ntot, flush=True)
    copies = []
    for _ in range(2):
        ds = [scene_to_device(s) for s in scs]
        copies.append(ds)
    st0 = [d["state"].clone() for d in copies[0]]
    chans = [Chain(schedule=_abi.SCHED_QUEUE), Chain(schedule=_abi.SCHED_QUEUE)]
    tends = [[torch.empty((L.NSTATE, 72, s["temperature"].shape[1]), dtype=torch.float64,
                          device="cuda") for s in scs] for _ in range(2)]
    batches = [ch.batch(ds, DT, tt) for ch, ds, tt in zip(chans, copies, tends)]
    # eager run of copy 1; copy 0 captured in a graph and replayed twice
    batches[1].run()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        batches[0].run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        batches[0].run()
    for rep in range(2):
        for t in tends[0]:
            t.fill_(float("nan"))
        graph.replay()
        torch.cuda.synchronize()
        _ = torch.ones(1024, device="cuda").sum().item()  # eager work between replays
        for a, b in zip(tends[0], tends[1]):
            assert torch.equal(a, b), rep
    # every cell conserves every family; sampled columns of every scene
    for i, (s, t, s0) in enumerate(zip(scs, tends[1], st0)):
        worst = _family_residual(torch, s0, t)
        assert max(worst.values()) < 1e-12, (i, worst)
        ncol = s["temperature"].shape[1]
        cols = _sample(ncol, 4)
        _sample_bar(orc, s, cols, t.cpu().numpy()[..., cols])
    for b in batches:
        b.close()
    # the mixed build on the same batch runs finite
    chm = Chain(precision=_abi.PREC_MIXED, schedule=_abi.SCHED_QUEUE)
    bm = chm.batch(copies[0], DT, tends[0])
    bm.run()
    torch.cuda.synchronize()
    assert all(bool(torch.isfinite(t).all()) for t in tends[0])
    bm.close()


# --------------------------------------------------------------- config 5 --

def test_large_scene_345600x128(orc, torch_cuda):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    ncol, nlev = 345600, 128
    print("config 5: generating %d x %d" % (ncol, nlev), flush=True)
    sc = scenes.generate(ncol, nlev, seed=5)
    assert 10.0 < sc["pressure"][0].min() and sc["pressure"][0].max() < 10.5  # top at 10 Pa
    d = scene_to_device(sc)
    cols = _sample(ncol)
    sub = _columns(sc, cols)
    del sc
    print("config 5: FP64 step", flush=True)
    ch = Chain(schedule=_abi.SCHED_QUEUE)
    tend = torch.empty((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda")
    ch.compute_tendencies(d, DT, tend)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(tend).all())
    worst = _family_residual(torch, d["state"], tend)
    assert max(worst.values()) < 1e-12, worst
    tg = tend[..., torch.as_tensor(cols, device="cuda")].cpu().numpy()
    _sample_bar(orc, sub, np.arange(len(cols)), tg)
    print("config 5: second FP64 run (bitwise)", flush=True)
    tend2 = torch.empty_like(tend)
    ch.compute_tendencies(d, DT, tend2)
    torch.cuda.synchronize()
    assert torch.equal(tend, tend2)
    ch.close()
    del tend2
    print("config 5: mixed step", flush=True)
    chm = Chain(precision=_abi.PREC_MIXED)
    chm.compute_tendencies(d, DT, tend)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(tend).all())
    chm.close()
