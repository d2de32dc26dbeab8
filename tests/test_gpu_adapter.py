"""The C++ host interface (mam4xx_amd/host/mam4x_process.hpp) on the GPU.

tests/cpp/adapter_driver.cpp stores every field in its own device allocation
in the EAMxx layout (ncol x nlev, level fastest), fills the AeroProcess-shaped
containers (aero_config.hpp:74-260, atmosphere.hpp:16-137) species by
species and calls MicrophysicsProcess::validate and ::compute_tendencies, the
way a host model would (also after MicrophysicsProcess::init, and with column
budgets).  Results are held to the bar of test_gpu_parity.py against the
oracle.
"""
import ctypes
import os

import numpy as np
import pytest

from mam4xx_amd import scenes

from .parity import report, state_rel_err

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "_build", "libadapter_driver.so")


@pytest.fixture(scope="module", autouse=True)
def hip_runtime_from_torch():
    """Initialise the GPU through torch before the driver library loads, so
    the whole test process shares torch's HIP runtime (the driver and
    libmam4x.so then bind to the runtime already loaded)."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    torch.cuda.init()


def run_driver(sc, dt, proc=False, budgets=False, do_coag=-1):
    if not os.path.exists(DRIVER):
        pytest.fail("%s not built: run __graft_entry__.build()" % DRIVER)
    lib = ctypes.CDLL(DRIVER)
    dp = ctypes.POINTER(ctypes.c_double)
    p = lambda a: a.ctypes.data_as(dp)
    st = np.ascontiguousarray(sc["state"])
    _, nlev, ncol = st.shape
    arr = [np.ascontiguousarray(sc[k]) for k in ("temperature", "pressure", "height", "qv",
                                                  "cloud_fraction", "pblh", "dgncur_a",
                                                  "dgncur_awet", "wetdens")]
    tend = np.empty_like(st)
    dgi = np.empty((4, nlev, ncol))
    dgc = np.empty_like(dgi)
    pr = np.full((5, 31, nlev, ncol), np.nan) if proc else None
    bud = np.full((54, ncol), np.nan) if budgets else None
    hdp = np.ascontiguousarray(sc["hydrostatic_dp"]) if budgets else None
    err = ctypes.create_string_buffer(512)
    lib.adapter_run.restype = ctypes.c_int
    rc = lib.adapter_run(ncol, nlev, ctypes.c_double(dt), p(st), *[p(a) for a in arr], p(tend),
                         p(dgi), p(dgc), None if pr is None else p(pr), err, 512,
                         None if hdp is None else p(hdp), None if bud is None else p(bud),
                         int(do_coag))
    if budgets:
        return rc, err.value.decode(), tend, bud
    if proc:
        return rc, err.value.decode(), tend, dgi, dgc, pr
    return rc, err.value.decode(), tend, dgi, dgc


def test_host_interface_matches_oracle(orc):
    sc = scenes.generate(160, 72, seed=17)
    dt = 1800.0
    rc, msg, tg, dig, dcg = run_driver(sc, dt)
    assert rc == 0, (rc, msg)
    to, env = orc.envelope(sc, dt)
    _, err = state_rel_err(tg, to, sc["state"], dt)
    cell_env = env.max(axis=0)
    bound = np.where(cell_env > 1e-11, np.maximum(1e-10, 10 * cell_env), 1e-10)[None]
    assert (err <= bound).all(), report(tg, to, sc["state"], dt)
    _, dio, dco, _ = orc.chain_step(sc, dt)
    assert np.all(np.isfinite(dig)) and np.all(np.isfinite(dcg))


def test_host_interface_validate_rejects_negative_state():
    sc = scenes.generate(32, 8, seed=18)
    sc["state"][3, 4, 5] = -1.0
    rc, msg, _, _, _ = run_driver(sc, 1800.0)
    assert rc == 2, (rc, msg)


def test_host_interface_fills_diagnostic_arrays(orc):
    """MicrophysDiagnosticArrays through the host interface (separate
    allocations, EAMxx layout) against the oracle's orc_chain_step_proc, to
    the bar of test_gpu_process.py."""
    from .test_gpu_process import _compare, _scales
    sc = scenes.generate(96, 72, seed=19)
    sc["cloud_fraction"][:, :16] = 0.5
    dt = 1800.0
    rc, msg, tg, _, _, pg = run_driver(sc, dt, proc=True)
    assert rc == 0, (rc, msg)
    to, _, _, _, po = orc.chain_step(sc, dt, proc=True)
    env = np.zeros_like(po)
    try:
        for s in range(1, 7):
            orc.set_libm_noise(s)
            env = np.maximum(env, np.abs(orc.chain_step(sc, dt, proc=True)[4] - po))
    finally:
        orc.set_libm_noise(0)
    _compare(pg, po, env, _scales(sc, to))
    assert np.count_nonzero(pg[:4]) > 1000


def test_host_interface_init_and_budgets(orc):
    """MicrophysicsProcess::init (coagulation switched off on a live process)
    and the column budgets through the host interface: the tendencies hold
    the oracle's bar for that configuration, the budgets are the exact
    column sums of those tendencies and the closures vanish."""
    from . import budget_ref as B
    sc = scenes.generate(96, 72, seed=20)
    dt = 1800.0
    rc, msg, tg, bud = run_driver(sc, dt, budgets=True, do_coag=0)
    assert rc == 0, (rc, msg)
    import oracle.pyoracle as P
    cfg = P.default_config(do_coag=0)
    to, env = orc.envelope(sc, dt, config=cfg)
    _, err = state_rel_err(tg, to, sc["state"], dt)
    cell_env = env.max(axis=0)
    bound = np.where(cell_env > 1e-11, np.maximum(1e-10, 10 * cell_env), 1e-10)[None]
    assert (err <= bound).all(), report(tg, to, sc["state"], dt)
    with_coag = run_driver(sc, dt)[2]
    assert not np.array_equal(tg, with_coag)
    ex, gross = B.exact(tg, sc["hydrostatic_dp"])
    assert (np.abs(bud - ex) <= 1e-14 * gross).all()
    for k in (B.SULFUR, B.ORGANIC):
        assert (np.abs(bud[k]) <= 1e-12 * gross[k]).all()
