"""The tangent (forward-mode dual-number) build of the kernel's per-cell code,
compiled for the CPU (tests/cpp/cell_host.cpp, policy Tangent), against
central finite differences of the oracle.

The chain is piecewise smooth (clips at zero, regime switches, the SOA
sub-step count).  An element counts as smooth when the central differences at
steps h and h/2 agree to 1e-4; there the tangent must match them.
"""
import ctypes

import numpy as np
import pytest

from mam4xx_amd import scenes

from .test_cell_host import cellhost  # noqa: F401  (fixture)

DT = 1800.0


def run_tangent(lib, sc, cfg, v_state, v_T, v_qv):
    dp = ctypes.POINTER(ctypes.c_double)
    p = lambda a: a.ctypes.data_as(dp) if a is not None else None
    st = np.ascontiguousarray(sc["state"])
    _, nlev, ncol = st.shape
    arr = [np.ascontiguousarray(sc[k]) for k in ("temperature", "pressure", "height", "qv",
                                                  "cloud_fraction", "pblh", "dgncur_a",
                                                  "dgncur_awet", "wetdens")]
    tend = np.empty_like(st)
    dtend = np.empty_like(st)
    dgi = np.empty((4, nlev, ncol))
    dgc = np.empty_like(dgi)
    lib.cell_host_run2.restype = ctypes.c_long
    lib.cell_host_run2((ctypes.c_int * 9)(*cfg), 2, ncol, nlev, ctypes.c_double(DT), p(st),
                       *[p(a) for a in arr], p(tend), p(dgi), p(dgc), p(v_state), p(v_T), None,
                       p(v_qv), p(dtend))
    return tend, dtend


def central_fd(orc, sc, oc, v_state, v_T, v_qv, h):
    out = []
    for sgn in (1.0, -1.0):
        s = dict(sc)
        s["state"] = sc["state"] + sgn * h * v_state
        s["temperature"] = sc["temperature"] + sgn * h * v_T
        s["qv"] = sc["qv"] + sgn * h * v_qv
        out.append(orc.chain_step(s, DT, config=oc)[0])
    return (out[0] - out[1]) / (2 * h)


@pytest.mark.parametrize("seed", [3, 4])
def test_tangent_matches_central_differences(orc, cellhost, seed):
    import oracle.pyoracle as P
    sc = scenes.generate(24, 72, seed=seed)
    oc = P.default_config()
    cfg = (oc.do_cond, oc.do_rename, oc.do_newnuc, oc.do_coag, oc.newnuc_h2so4_conc_optaa,
           oc.n_so4_monolayers_pcage, oc.do_calcsize, oc.calcsize_do_aitacc_transfer,
           oc.calcsize_do_adjust)
    rng = np.random.default_rng(seed)
    st = sc["state"]
    v_state = st * rng.standard_normal(st.shape)  # relative perturbation of every prognostic
    v_T = rng.standard_normal(sc["temperature"].shape)
    v_qv = sc["qv"] * rng.standard_normal(sc["qv"].shape)
    tend, dtend = run_tangent(cellhost, sc, cfg, v_state, v_T, v_qv)
    # the value part of the dual numbers is the FP64 chain itself
    assert np.array_equal(tend, orc.chain_step(sc, DT, config=oc)[0])
    f1 = central_fd(orc, sc, oc, v_state, v_T, v_qv, 1e-6)
    f2 = central_fd(orc, sc, oc, v_state, v_T, v_qv, 5e-7)
    scale = np.abs(st) / DT + 1e-300
    smooth = np.abs(f1 - f2) <= 1e-4 * (np.abs(f1) + scale)
    err = np.abs(dtend - f1) / (np.abs(f1) + scale)
    assert smooth.mean() > 0.999
    assert np.median(err[smooth]) < 1e-8
    assert (err[smooth] > 1e-4).mean() < 5e-4, np.sort(err[smooth])[-5:]


def test_seeded_contraction_matches_tangent(cellhost):
    """The vector-Jacobian product path of the per-cell code (xbar = sum_f
    tbar_f d tend_f, the building block of mam4x_compute_adjoint) against the
    tangent's derivative planes contracted on the host, for unit seeds of
    single inputs; zero cotangent planes contribute nothing."""
    import oracle.pyoracle as P
    sc = scenes.generate(12, 72, seed=31)
    oc = P.default_config()
    cfg = (oc.do_cond, oc.do_rename, oc.do_newnuc, oc.do_coag, oc.newnuc_h2so4_conc_optaa,
           oc.n_so4_monolayers_pcage, oc.do_calcsize, oc.calcsize_do_aitacc_transfer,
           oc.calcsize_do_adjust)
    dp = ctypes.POINTER(ctypes.c_double)
    p = lambda a: a.ctypes.data_as(dp) if a is not None else None
    st = np.ascontiguousarray(sc["state"])
    _, nlev, ncol = st.shape
    arr = [np.ascontiguousarray(sc[k]) for k in ("temperature", "pressure", "height", "qv",
                                                  "cloud_fraction", "pblh", "dgncur_a",
                                                  "dgncur_awet", "wetdens")]
    rng = np.random.default_rng(32)
    tbar = rng.standard_normal(st.shape)
    tbar[3] = 0.0
    ones = np.ones((nlev, ncol))
    lib = cellhost
    lib.cell_host_run3.restype = ctypes.c_long
    for j in (1, 12, 44, "T", "qv"):
        unit = np.zeros_like(st)
        v_T = v_qv = None
        if j == "T":
            v_T = ones
        elif j == "qv":
            v_qv = ones
        else:
            unit[j] = 1.0
        tend, dtend = run_tangent(lib, sc, cfg, unit, v_T, v_qv)
        xbar = np.full((nlev, ncol), np.nan)
        lib.cell_host_run3((ctypes.c_int * 9)(*cfg), 2, ncol, nlev, ctypes.c_double(DT), p(st),
                           *[p(a) for a in arr], None, None, None, p(unit), p(v_T), None,
                           p(v_qv), None, p(tbar), p(xbar))
        # the same products summed in the same field order
        ref = np.zeros((nlev, ncol))
        for f in range(st.shape[0]):
            ref = ref + tbar[f] * dtend[f]
        assert np.array_equal(xbar, ref), j
