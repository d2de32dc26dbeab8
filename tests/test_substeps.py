"""Condensation sub-steps (ntsubstep > 1) on the CPU.

The reference fixes ntsubstep = 1 (mam4_amicphys.hpp:1288), so nothing of
its own pins more.  These tests pin this build's semantics (the subarea
chain repeated over int(dt / n) sub-steps, the first sub-step's H2SO4 uptake
rates held for the rest, 1/n of the cloud-chemistry growth per cloudy
renaming) by:

* the kernel's per-cell code, built for the CPU, against the oracle: the
  tendencies bit for bit, the per-process diagnostics to the last bits
  (summed in another order), every cloud category, with and without a
  pre-cloud-chemistry state;
* a closed form: with nucleation off nothing but the analytic update
  q(t) = (q0 - P/k) e^{-kt} + P/k moves H2SO4 gas, and with the rates held
  that update composes exactly, so n sub-steps of dt / n leave the same gas
  as one of dt (recomputing the rates per sub-step, or updating over dt
  instead of dt / n, breaks it);
* convergence: the step's result settles as n doubles, as an operator
  splitting does (a process run over dt instead of dt / n makes it drift
  instead);
* species conservation, as for one sub-step;
* the tangent against central differences of the oracle, and the reverse
  sweep against the tangent, at n = 3, including the two-level storage the
  GPU uses for the SOA loop's states.
"""
import ctypes

import numpy as np
import pytest

from mam4xx_amd import layout as L
from mam4xx_amd import scenes

from .test_adjoint_host import ALL, adjoint, row_scaled_error, tangent
from .test_cell_host import cellhost  # noqa: F401  (fixture)
from .test_oracle_invariants import moles
from .test_tangent_host import central_fd

DT = 1800.0
KEYS = ("temperature", "pressure", "height", "qv", "cloud_fraction", "pblh", "dgncur_a",
        "dgncur_awet", "wetdens")


@pytest.fixture
def host3(cellhost):
    cellhost.cell_host_set_ntsubstep(3)
    yield cellhost
    cellhost.cell_host_set_ntsubstep(1)
    cellhost.cell_host_set_tape(0)


def _scene(seed, ncol=24):
    sc = scenes.generate(ncol, 72, seed=seed)
    sc["cloud_fraction"][:, :4] = 0.9995
    sc["cloud_fraction"][:, 4:8] = 0.5
    return sc


@pytest.mark.parametrize("ns", [2, 3])
@pytest.mark.parametrize("precld", [False, True], ids=["current", "precld"])
def test_kernel_code_bitwise_equals_oracle(orc, cellhost, ns, precld):
    import oracle.pyoracle as P
    sc = _scene(300 + ns)
    pre = None
    if precld:
        rng = np.random.default_rng(ns)
        pre = np.ascontiguousarray(sc["state"] * rng.uniform(0.5, 1.0, sc["state"].shape))
    oc = P.default_config(ntsubstep=ns)
    cfg = (1, 1, 1, 1, oc.newnuc_h2so4_conc_optaa, 8, 1, 1, 1)
    dp = ctypes.POINTER(ctypes.c_double)
    p = lambda a: None if a is None else a.ctypes.data_as(dp)
    st = np.ascontiguousarray(sc["state"])
    nf, nlev, ncol = st.shape
    tend = np.empty_like(st)
    proc = np.full((5, 31, nlev, ncol), np.nan)
    arr = [np.ascontiguousarray(sc[k]) for k in KEYS]
    cellhost.cell_host_proc.restype = ctypes.c_long
    cellhost.cell_host_set_ntsubstep(ns)
    try:
        ith = cellhost.cell_host_proc((ctypes.c_int * 9)(*cfg), ncol, nlev, ctypes.c_double(DT),
                                      p(st), *[p(a) for a in arr], p(pre), p(tend), p(proc))
    finally:
        cellhost.cell_host_set_ntsubstep(1)
    to, _, _, ito, po = orc.chain_step(sc, dt=DT, config=oc, precldchem=pre, proc=True)
    assert ith == ito
    assert np.array_equal(tend, to), np.argwhere(tend != to)[:5]
    # the diagnostics: the kernel adds each sub-step's increment to the
    # output in its final units, the oracle converts the summed increments
    # once, so they agree to the last bits of each plane's scale
    scale = np.abs(po).max(axis=(2, 3), keepdims=True) + 1e-300
    assert (np.abs(proc - po) <= 1e-14 * scale).all(), np.abs(proc - po).max()
    # and it is a different step from one sub-step
    assert not np.array_equal(to, orc.chain_step(sc, dt=DT, config=P.default_config())[0])


def test_h2so4_gas_composes_exactly(orc):
    import oracle.pyoracle as P
    sc = _scene(311, ncol=32)
    out = {}
    for ns in (1, 3, 6):
        oc = P.default_config(do_newnuc=0, ntsubstep=ns)
        out[ns] = orc.chain_step(sc, dt=DT, config=oc)[0][L.F_H2SO4]
    h0 = sc["state"][L.F_H2SO4]
    for ns in (3, 6):
        # the gas after the step, relative to its size before and after
        den = np.maximum(np.abs(h0 + DT * out[1]), np.abs(h0)) + 1e-300
        rel = np.abs(out[ns] - out[1]) * DT / den
        assert rel.max() < 1e-11, (ns, rel.max())
    # with the rates recomputed from the number after coagulation the gas
    # would differ by the number change; the scene has coagulation working
    oc = P.default_config(do_newnuc=0, ntsubstep=3)
    t3 = orc.chain_step(sc, dt=DT, config=oc)[0]
    assert np.abs(t3[L.f_ni(1)]).max() > 0


def test_step_converges_as_substeps_double(orc):
    import oracle.pyoracle as P
    dt = 1728.0  # 2^6 x 27: every n below divides it, so the sub-steps sum to dt
    sc = _scene(312, ncol=32)
    st = sc["state"]
    res = {ns: orc.chain_step(sc, dt=dt, config=P.default_config(ntsubstep=ns))[0]
           for ns in (2, 4, 8, 16, 32)}
    floor = np.where(np.arange(L.NSTATE) < 8, 1e-10, 1e-30)[:, None, None]
    den = np.maximum(np.abs(st), floor)
    # the elements the sub-stepping changes at all
    act = np.abs(res[32] - res[2]) * dt / den > 1e-12
    assert act.mean() > 0.2
    d = [np.median((np.abs(res[2 * n] - res[n]) * dt / den)[act]) for n in (2, 4, 8, 16)]
    # each doubling shrinks the change (by 0.64-0.78 on this scene: slower
    # than first order, the SOA loop restarting every sub-step)
    assert all(d[k + 1] < 0.9 * d[k] for k in range(3)) and d[3] < 0.5 * d[0], d


@pytest.mark.parametrize("ns", [3, 8])
def test_species_conserved(orc, ns):
    import oracle.pyoracle as P
    sc = _scene(313 + ns)
    tend = orc.chain_step(sc, dt=DT, config=P.default_config(ntsubstep=ns, do_calcsize=0))[0]
    st = sc["state"]
    for spec in ("so4", "soa", "pom", "bc", "dst", "ncl", "mom"):
        before = moles(st, spec)
        after = moles(st + DT * tend, spec)
        scale = np.abs(before) + 1e-300
        assert (np.abs(after - before) / scale).max() < 1e-10, spec


def test_tangent_matches_central_differences(orc, host3):
    import oracle.pyoracle as P
    from .test_tangent_host import run_tangent
    sc = _scene(320)
    oc = P.default_config(ntsubstep=3)
    cfg = (1, 1, 1, 1, 2, 8, 1, 1, 1)
    rng = np.random.default_rng(320)
    st = sc["state"]
    v_state = st * rng.standard_normal(st.shape)
    v_T = rng.standard_normal(sc["temperature"].shape)
    v_qv = sc["qv"] * rng.standard_normal(sc["qv"].shape)
    tend, dtend = run_tangent(host3, sc, cfg, v_state, v_T, v_qv)
    assert np.array_equal(tend, orc.chain_step(sc, DT, config=oc)[0])
    f1 = central_fd(orc, sc, oc, v_state, v_T, v_qv, 1e-6)
    f2 = central_fd(orc, sc, oc, v_state, v_T, v_qv, 5e-7)
    scale = np.abs(st) / DT + 1e-300
    smooth = np.abs(f1 - f2) <= 1e-4 * (np.abs(f1) + scale)
    err = np.abs(dtend - f1) / (np.abs(f1) + scale)
    assert smooth.mean() > 0.999
    assert np.median(err[smooth]) < 1e-8
    assert (err[smooth] > 1e-4).mean() < 5e-4, np.sort(err[smooth])[-5:]


@pytest.mark.parametrize("tape", [0, 1], ids=["all-states", "two-level"])
def test_reverse_sweep_jacobian_equals_tangent_jacobian(host3, tape):
    host3.cell_host_set_tape(tape)
    sc = _scene(330, ncol=6)
    rng = np.random.default_rng(330)
    pre = np.ascontiguousarray(sc["state"] * rng.uniform(0.5, 1.0, sc["state"].shape))
    st = sc["state"]
    nf, nlev, ncol = st.shape
    z = np.zeros((nlev, ncol))
    one = np.ones((nlev, ncol))
    Jt = np.empty((nf, nf + 3, nlev, ncol))
    Ja = np.empty_like(Jt)
    for j in range(nf + 3):
        v = np.zeros_like(st)
        if j < nf:
            v[j] = 1.0
        Jt[:, j] = tangent(host3, sc, ALL, v, one if j == nf else z, one if j == nf + 1 else z,
                           one if j == nf + 2 else z, pre=pre)
    for f in range(nf):
        tb = np.zeros_like(st)
        tb[f] = 1.0
        sb, Tb, pb, qb = adjoint(host3, sc, ALL, tb, pre=pre)
        Ja[f, :nf] = sb
        Ja[f, nf], Ja[f, nf + 1], Ja[f, nf + 2] = Tb, pb, qb
    assert np.isfinite(Ja).all()
    err = row_scaled_error(sc, Jt, Ja)
    assert np.quantile(err, 0.9999) < 1e-11
    assert err.max() < 1e-8


def test_dot_product_per_cell(host3):
    """<w, J v> = <J^T w, v> per cell at n = 3, on a nucleation-heavy scene
    whose first sub-steps' SOA loops exceed the two-level storage's 32
    states (so its segment recomputation runs)."""
    host3.cell_host_set_tape(1)
    sc = scenes.generate(24, 72, seed=6, h2so4_range=(1e-11, 1e-9))
    rng = np.random.default_rng(12)
    st = sc["state"]
    nlev, ncol = st.shape[1:]
    v = rng.standard_normal(st.shape) * np.abs(st) * 1e-3
    vT = rng.standard_normal((nlev, ncol)) * 0.1
    vp = rng.standard_normal((nlev, ncol)) * 10.0
    vq = rng.standard_normal((nlev, ncol)) * np.abs(sc["qv"]) * 1e-3
    dtend = tangent(host3, sc, ALL, v, vT, vp, vq)
    tb = rng.standard_normal(st.shape) / np.maximum(np.abs(st), 1e-30) * 1e-3
    sb, Tb, pb, qb = adjoint(host3, sc, ALL, tb)
    lhs = (tb * dtend).sum(axis=0)
    rhs = (sb * v).sum(axis=0) + Tb * vT + pb * vp + qb * vq
    scale = (np.abs(tb * dtend).sum(axis=0) + np.abs(sb * v).sum(axis=0) + np.abs(Tb * vT)
             + np.abs(pb * vp) + np.abs(qb * vq) + (np.abs(tb) * np.abs(v)).sum(axis=0) / DT
             + 1e-300)
    rel = np.abs(lhs - rhs) / scale
    assert np.median(rel) < 1e-14
    assert (rel > 1e-8).mean() <= 2e-3
    assert rel.max() < 1e-5

