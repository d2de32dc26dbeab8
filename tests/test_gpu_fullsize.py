"""The library at the benchmark's full size, 21,600 columns x 72 levels
(1,555,200 cells), through properties that do not depend on the domain size:

* cells are independent: a sample of columns taken from the full-size run
  agrees with the oracle run on those columns alone, to the bar of
  test_gpu_parity.py (chain) and of the companion tests (wet size,
  chemistry);
* the microphysics conserves every species family in moles -- SO4 with
  H2SO4, SOA with SOAG (in the chain's carbon units), POM, BC, dust, NaCl,
  MOM, over all modes, interstitial and cloud-borne -- in every cell;
* the work-queue schedule's results do not depend on the order in which the
  queue hands out cells: two runs agree bit for bit;
* the column kernels of the chemistry feeders (linoz, washout) at the same
  size agree with the oracle on a sample of columns.
"""
import numpy as np
import pytest

from mam4xx_amd import layout as L
from mam4xx_amd import scenes

from .parity import state_rel_err

pytestmark = pytest.mark.gpu

NCOL, NLEV, DT = 21600, 72, 1800.0
SAMPLE = np.arange(7, NCOL, 180)  # 120 columns spread over the domain


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


@pytest.fixture(scope="module")
def full(torch_cuda):
    """the benchmark scene and two full-size runs of one step"""
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    sc = scenes.generate(NCOL, NLEV, seed=0)
    d = scene_to_device(sc)
    runs = []
    for _ in range(2):
        ch = Chain(schedule=_abi.SCHED_QUEUE)
        tend = torch.empty((L.NSTATE, NLEV, NCOL), dtype=torch.float64, device="cuda")
        dgi = torch.empty((4, NLEV, NCOL), dtype=torch.float64, device="cuda")
        dgc = torch.empty((4, NLEV, NCOL), dtype=torch.float64, device="cuda")
        ch.compute_tendencies(d, DT, tend, dgi, dgc)
        torch.cuda.synchronize()
        runs.append((tend.cpu().numpy(), dgi.cpu().numpy(), dgc.cpu().numpy()))
        ch.close()
    return sc, runs


def _columns(sc, cols):
    return {k: (np.ascontiguousarray(v[..., cols]) if np.ndim(v) >= 1 and np.shape(v)[-1] == NCOL
                else v) for k, v in sc.items()}


def test_full_size_sample_matches_oracle(orc, full):
    sc, runs = full
    tg = runs[0][0][..., SAMPLE]
    sub = _columns(sc, SAMPLE)
    to, env = orc.envelope(sub, dt=DT)
    _, err = state_rel_err(tg, to, sub["state"], DT)
    cell_env = env.max(axis=0)
    bound = np.where(cell_env > 1e-11, np.maximum(1e-10, 10 * cell_env), 1e-10)[None]
    assert np.isfinite(runs[0][0]).all()
    assert (err <= bound).all(), (err.max(), np.argwhere(err > bound)[:3])
    assert (cell_env > 1e-10).mean() < 0.01


def test_full_size_conserves_every_species_family(full):
    sc, runs = full
    st, t = sc["state"], runs[0][0]
    names = L.field_names()
    for sp, gas, gw, w in (("so4", L.F_H2SO4, 1.0, 1.0), ("soa", L.F_SOAG, 12.011 / 150.0, 0.08),
                           ("pom", None, 0, 0.08), ("bc", None, 0, 1.0), ("dst", None, 0, 1.0),
                           ("ncl", None, 0, 1.0), ("mom", None, 0, 1.0)):
        fs = [f for f in range(8, 50) if names[f].startswith(sp + "_")]
        total = sum(w * np.abs(st[f]) * L.to_vmr(f) for f in fs)
        change = sum(w * t[f] * L.to_vmr(f) for f in fs)
        if gas is not None:
            total = total + gw * np.abs(st[gas]) * L.to_vmr(gas)
            change = change + gw * t[gas] * L.to_vmr(gas)
        rel = np.abs(change) * DT / np.maximum(total, 1e-300)
        assert rel.max() < 1e-12, (sp, rel.max(), np.unravel_index(rel.argmax(), rel.shape))


def test_full_size_runs_are_bitwise_reproducible(full):
    _, runs = full
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)


def test_full_size_wet_size_sample_matches_oracle(orc, torch_cuda):
    from mam4xx_amd.chain import Chain, scene_to_device
    torch = torch_cuda
    sc = scenes.generate(NCOL, NLEV, seed=0)
    d = scene_to_device(sc)
    new = lambda *s: torch.empty(s, dtype=torch.float64, device="cuda")
    dgi, dgw, qw, wd = (new(4, NLEV, NCOL) for _ in range(4))
    ch = Chain()
    ch.compute_wet_size(d, DT, dgi, dgw, qw, wetdens=wd)
    torch.cuda.synchronize()
    ch.close()
    got = [x.cpu().numpy()[..., SAMPLE] for x in (dgi, dgw, qw, wd)]
    ref = orc.wet_size(_columns(sc, SAMPLE), DT)
    for name, g, r in zip(("dgncur_i", "dgncur_awet", "qaerwat", "wetdens"), got,
                          (ref[0], ref[2], ref[3], ref[4])):
        den = np.maximum(np.abs(r), 1e-300)
        if name == "qaerwat":  # a difference of wet and dry volumes
            den = den / np.maximum(1.0 - (ref[0] / ref[2]) ** 3, 1e-300)
        err = np.where(g == r, 0.0, np.abs(g - r) / den)
        assert np.median(err) < 1e-15, name
        assert (err > 1e-12).mean() < 1e-3, (name, np.sort(err.ravel())[-3:])
        assert err.max() < 1e-6, (name, err.max())


def test_full_size_chemistry_feeders_sample_matches_oracle(orc, torch_cuda):
    """linoz (LDS block kernel) and washout over 21,600 columns; columns are
    independent, so the sampled columns must match the oracle run on them
    alone (bars of test_gpu_linoz.py and test_gpu_sethet.py)"""
    from mam4xx_amd.chain import Chain
    from .test_linoz import CONF, columns as linoz_columns
    from .test_sethet import columns as sethet_columns
    from .test_gpu_sethet import PLANES
    torch = torch_cuda
    dev = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).cuda()
    sub = lambda cols: {k: np.ascontiguousarray(v[..., SAMPLE]) for k, v in cols.items()}
    ch = Chain()
    # linoz
    cols = linoz_columns(np.random.default_rng(61), NLEV, NCOL)
    o3 = dev(cols["o3"])
    ch.compute_linoz(DT, dev(cols["temp"]), dev(cols["pmid"]), dev(cols["pdel"]),
                     dev(cols["clim"]), dev(cols["sza"]), dev(cols["rlats"]),
                     dev(cols["ltrop"], torch.int32), o3, CONF,
                     column_above_top=dev(cols["col0"]))
    # washout
    wcols = sethet_columns(np.random.default_rng(62), NLEV, NCOL)
    het = torch.empty((31, NLEV, NCOL), dtype=torch.float64, device="cuda")
    ch.compute_washout(DT, {k: dev(wcols[v]) for k, v in PLANES.items()}, dev(wcols["rlat"]),
                       dev(wcols["phis"]), het)
    torch.cuda.synchronize()
    ch.close()
    got = o3.cpu().numpy()[:, SAMPLE]
    ref = orc.linoz(DT, sub(cols), CONF)[0]
    o3_0 = cols["o3"][:, SAMPLE]
    assert (np.abs(got - ref) <= 1e-12 * np.maximum(np.abs(o3_0), np.abs(ref))).all()
    assert np.isfinite(o3.cpu().numpy()).all()
    hg = het.cpu().numpy()[..., SAMPLE]
    hr = orc.sethet(DT, sub(wcols))
    env = np.zeros_like(hr)
    try:
        for s in range(1, 7):
            orc.set_libm_noise(s)
            env = np.maximum(env, np.abs(orc.sethet(DT, sub(wcols)) - hr))
    finally:
        orc.set_libm_noise(0)
    assert (np.abs(hg - hr) <= np.maximum(1e-12 * np.abs(hr), 10 * env)).all()
    assert np.isfinite(het.cpu().numpy()).all() and (hr[1] > 0).sum() > 100
