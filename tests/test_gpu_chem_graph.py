"""The chemistry entry points only launch kernels (no allocation, copy,
memset or synchronisation), so a host can capture its chemistry sequence in
a hipGraph: invariants, forcing, washout, linoz, gas-phase chemistry and
setsox captured once on a torch CUDA graph, replayed several times with
eager work and synchronisations between the replays; every replay equals
the eager sequence bit for bit."""
import numpy as np
import pytest

from .test_gas_chem import _domain as gc_domain
from .test_gpu_sethet import PLANES as HET_PLANES
from .test_linoz import CONF, columns as linoz_columns
from .test_sethet import columns as sethet_columns
from .test_setsox import _domain as sox_domain

pytestmark = pytest.mark.gpu

NLEV, NCOL, DT = 72, 48, 1800.0
SOX_MET = {"pressure": "press", "temperature": "tfld", "liquid_water": "lwc",
           "cloud_fraction": "cldfrc", "droplet_number": "cldnum", "air_number_density": "xhnm"}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


def test_chemistry_sequence_replays_from_a_graph(torch_cuda):
    from mam4xx_amd.chain import Chain
    torch = torch_cuda
    rng = np.random.default_rng(71)
    dev = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).cuda()
    sh = lambda a: np.asarray(a).reshape(np.shape(a)[:-1] + (NLEV, NCOL))
    n = NLEV * NCOL
    # inputs
    T, p, qv = rng.uniform(200, 300, (NLEV, NCOL)), 10 ** rng.uniform(3, 5, (NLEV, NCOL)), \
        rng.uniform(0, 0.02, (NLEV, NCOL))
    inp = {"T": dev(T), "p": dev(p), "qv": dev(qv),
           "off": dev(10 ** rng.uniform(-12, -6, (4, NLEV, NCOL))),
           "sec": [[dev(10 ** rng.uniform(-2, 4, (NLEV, NCOL)))] for _ in range(9)]}
    lz = linoz_columns(rng, NLEV, NCOL)
    wo = sethet_columns(rng, NLEV, NCOL)
    temp, photo, extfrc, inv, het, qq = gc_domain(rng, n)
    met, qcw, qin = sox_domain(rng, n)
    d = {"lz": {k: dev(v) for k, v in lz.items() if k != "ltrop"},
         "ltrop": dev(lz["ltrop"], torch.int32),
         "wo": {k: dev(v) for k, v in wo.items()},
         "gc": [dev(sh(x)) for x in (temp, photo, extfrc, inv, het)],
         "sox_met": {k: dev(sh(met[v])) for k, v in SOX_MET.items()}}
    state0 = {"o3": dev(lz["o3"]), "qq": dev(sh(qq)), "qcw": dev(sh(qcw)), "qin": dev(sh(qin))}
    state = {k: v.clone() for k, v in state0.items()}
    outs = {"inv": torch.empty((8, NLEV, NCOL), dtype=torch.float64, device="cuda"),
            "ext": torch.empty((9, NLEV, NCOL), dtype=torch.float64, device="cuda"),
            "het": torch.empty((31, NLEV, NCOL), dtype=torch.float64, device="cuda")}
    ch = Chain()

    def sequence():
        ch.compute_invariants(inp["T"], inp["p"], inp["qv"], inp["off"], outs["inv"])
        ch.compute_forcing(inp["sec"], outs["ext"])
        w = d["wo"]
        ch.compute_washout(DT, {k: w[v] for k, v in HET_PLANES.items()}, w["rlat"], w["phis"],
                           outs["het"])
        z = d["lz"]
        ch.compute_linoz(DT, z["temp"], z["pmid"], z["pdel"], z["clim"], z["sza"], z["rlats"],
                         d["ltrop"], state["o3"], CONF, column_above_top=z["col0"])
        t, ph, ex, iv, ht = d["gc"]
        ch.compute_gas_chemistry(DT, t, ph, ex, iv, ht, state["qq"])
        ch.compute_setsox(DT, d["sox_met"], state["qcw"], state["qin"])

    def snapshot():
        torch.cuda.synchronize()
        return {k: v.cpu().numpy().copy() for k, v in list(outs.items()) + list(state.items())}

    def reset():
        for k, v in state0.items():
            state[k].copy_(v)
        for v in outs.values():
            v.fill_(float("nan"))

    reset()
    sequence()
    eager = snapshot()
    # capture on a side stream (torch's graph API), then replay
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    reset()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            sequence()
    torch.cuda.synchronize()
    for replay in range(4):
        reset()
        torch.cuda.synchronize()
        g.replay()
        got = snapshot()
        for k in eager:
            assert np.array_equal(got[k], eager[k], equal_nan=True), (replay, k)
        # eager work between replays
        sequence()
        torch.cuda.synchronize()
    assert np.isfinite(eager["het"]).all() and np.isfinite(eager["qq"]).all()
    ch.close()
