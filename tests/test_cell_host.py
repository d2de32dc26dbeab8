"""The HIP kernel's per-cell code, compiled for the CPU, against the oracle.

mam4xx_amd/csrc/chain_cell.hpp restructures the reference's per-cell arrays
(compact per-subarea state, hoisted constants, dropped diagnostics).  Built
with g++ and glibc -- the oracle's own libm -- and without FMA contraction,
it must reproduce the oracle BIT FOR BIT: any difference is a logic error in
the restructuring, not rounding.  (On the GPU only device-libm rounding
differs; see test_gpu_parity.py.)
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from mam4xx_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
SO = os.path.join(BUILD, "libcellhost.so")


@pytest.fixture(scope="module")
def cellhost():
    os.makedirs(BUILD, exist_ok=True)
    # built under a per-process name and renamed into place, so concurrent
    # test processes never overwrite a library another one has loaded
    tmp = SO + ".%d" % os.getpid()
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    "-o", tmp, os.path.join(HERE, "cpp", "cell_host.cpp"),
                    os.path.join(ROOT, "mam4xx_amd", "csrc", "chain_consts.cpp")], check=True)
    os.replace(tmp, SO)
    lib = ctypes.CDLL(SO)
    lib.cell_host_run.restype = ctypes.c_long
    return lib


def run_host(lib, sc, dt, cfg, prec=0):
    dp = ctypes.POINTER(ctypes.c_double)
    p = lambda a: a.ctypes.data_as(dp)
    st = np.ascontiguousarray(sc["state"])
    nf, nlev, ncol = st.shape
    tend = np.empty_like(st)
    dgi = np.empty((4, nlev, ncol))
    dgc = np.empty_like(dgi)
    arr = [np.ascontiguousarray(sc[k]) for k in ("temperature", "pressure", "height", "qv",
                                                  "cloud_fraction", "pblh", "dgncur_a",
                                                  "dgncur_awet", "wetdens")]
    c = (ctypes.c_int * 9)(*cfg)
    it = lib.cell_host_run(c, prec, ncol, nlev, ctypes.c_double(dt), p(st), *[p(a) for a in arr],
                           p(tend), p(dgi), p(dgc))
    return tend, dgi, dgc, it


CASES = [
    ("full", {}, 1800.0, 0),
    ("cond", dict(do_rename=0, do_newnuc=0, do_coag=0), 1800.0, 1),
    ("rename", dict(do_cond=0, do_newnuc=0, do_coag=0), 1800.0, 2),
    ("newnuc", dict(do_cond=0, do_rename=0, do_coag=0), 1800.0, 3),
    ("coag", dict(do_cond=0, do_rename=0, do_newnuc=0), 1800.0, 4),
    ("noxfer", dict(calcsize_do_aitacc_transfer=0), 1800.0, 5),
    ("noadjust", dict(calcsize_do_adjust=0), 1800.0, 6),
    ("nocalcsize", dict(do_calcsize=0), 1800.0, 7),
    ("dt60", {}, 60.0, 8),
    ("dt1", {}, 1.0, 9),
    ("sequential", dict(newnuc_h2so4_conc_optaa=1), 1800.0, 10),
    ("swapped_optaa", dict(gaexch_h2so4_uptake_optaa=1), 1800.0, 11),
]


@pytest.mark.parametrize("name,cfgd,dt,seed", CASES, ids=[c[0] for c in CASES])
def test_kernel_code_bitwise_equals_oracle(orc, cellhost, name, cfgd, dt, seed):
    import oracle.pyoracle as P
    sc = scenes.generate(48, 72, seed=100 + seed)
    # exercise every cloud category and the clipped-negative path
    sc["cloud_fraction"][:, :8] = 0.9995
    sc["cloud_fraction"][:, 8:16] = 0.5
    sc["state"][:, :2, 16:20] *= -1.0
    oc = P.default_config(**cfgd)
    # the kernel receives the swapped optaa switch, as mam4x.hip's fill_args passes it
    cfg = (oc.do_cond, oc.do_rename, oc.do_newnuc, oc.do_coag, oc.newnuc_h2so4_conc_optaa,
           oc.n_so4_monolayers_pcage, oc.do_calcsize, oc.calcsize_do_aitacc_transfer,
           oc.calcsize_do_adjust)
    th, dih, dch, ith = run_host(cellhost, sc, dt, cfg)
    to, dio, dco, ito = orc.chain_step(sc, dt=dt, config=oc)
    assert ith == ito
    assert np.array_equal(th, to), np.argwhere(th != to)[:5]
    assert np.array_equal(dih, dio) and np.array_equal(dch, dco)


PROC_CASES = [c for c in CASES if c[0] in ("full", "cond", "rename", "newnuc", "coag",
                                           "dt60", "sequential", "swapped_optaa")]


@pytest.mark.parametrize("name,cfgd,dt,seed", PROC_CASES, ids=[c[0] for c in PROC_CASES])
@pytest.mark.parametrize("precld", [False, True], ids=["current", "precld"])
def test_process_diagnostics_bitwise_equal_oracle(orc, cellhost, name, cfgd, dt, seed, precld):
    """The diagnostics instantiation (Fp64Diag) of the kernel's cell code:
    the per-process tendency blocks and the tendencies bit for bit against
    the oracle's orc_chain_step_proc, every cloud category, with and without
    a pre-cloud-chemistry state (cloud-borne renaming)."""
    import oracle.pyoracle as P
    sc = scenes.generate(48, 72, seed=200 + seed)
    sc["cloud_fraction"][:, :8] = 0.9995
    sc["cloud_fraction"][:, 8:16] = 0.5
    sc["state"][:, :2, 16:20] *= -1.0
    pre = None
    if precld:
        rng = np.random.default_rng(seed)
        pre = sc["state"] * rng.uniform(0.5, 1.0, sc["state"].shape)
    oc = P.default_config(**cfgd)
    cfg = (oc.do_cond, oc.do_rename, oc.do_newnuc, oc.do_coag, oc.newnuc_h2so4_conc_optaa,
           oc.n_so4_monolayers_pcage, oc.do_calcsize, oc.calcsize_do_aitacc_transfer,
           oc.calcsize_do_adjust)
    dp = ctypes.POINTER(ctypes.c_double)
    p = lambda a: None if a is None else a.ctypes.data_as(dp)
    st = np.ascontiguousarray(sc["state"])
    nf, nlev, ncol = st.shape
    tend = np.empty_like(st)
    proc = np.full((5, 31, nlev, ncol), np.nan)
    arr = [np.ascontiguousarray(sc[k]) for k in ("temperature", "pressure", "height", "qv",
                                                  "cloud_fraction", "pblh", "dgncur_a",
                                                  "dgncur_awet", "wetdens")]
    cellhost.cell_host_proc.restype = ctypes.c_long
    ith = cellhost.cell_host_proc((ctypes.c_int * 9)(*cfg), ncol, nlev, ctypes.c_double(dt),
                                  p(st), *[p(a) for a in arr], p(pre), p(tend), p(proc))
    to, _, _, ito, po = orc.chain_step(sc, dt=dt, config=oc, precldchem=pre, proc=True)
    assert ith == ito
    assert np.array_equal(tend, to)
    assert np.array_equal(proc, po), np.argwhere(proc != po)[:5]
