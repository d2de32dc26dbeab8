"""The reverse sweep of the kernel's per-cell code (chain_adjoint.hpp),
compiled for the CPU (tests/cpp/cell_host.cpp), against the tangent build of
the same code.

The tangent is pinned against central finite differences of the oracle
(test_tangent_host.py), so agreement with it pins the reverse sweep.  Two
comparisons:

* whole Jacobians: the tangent with a unit seed per input (55 passes) gives
  the columns of every cell's 52 x 55 Jacobian block, the reverse sweep with
  a unit cotangent per output (52 passes) its rows.  Element (f, j) is
  compared in units of the largest contribution |J_fj' x_j'| to row f (or
  |x_f| / dt when that is larger), so an element that is negligible in its
  row is not held to a relative bound it cannot carry;
* dot products along random directions, <w, J v> = <J^T w, v> per cell,
  over process subsets, cloud categories, a pre-cloud-chemistry state and
  time steps from 1 s to 3600 s.

The two modes accumulate the chain rule in opposite orders.  In the SOA
sub-step loop (up to some tens of stiff sub-steps) contributions of opposite
sign cancel, so a few cells agree to 1e-9 .. 1e-6 of their scale rather than
to round-off; the bounds below say how many.
"""
import ctypes

import numpy as np
import pytest

from mam4xx_amd import scenes

from .test_cell_host import cellhost  # noqa: F401  (fixture)

DT = 1800.0
ALL = (1, 1, 1, 1, 2, 8, 1, 1, 1)
KEYS = ("temperature", "pressure", "height", "qv", "cloud_fraction", "pblh", "dgncur_a",
        "dgncur_awet", "wetdens")
_dp = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def tangent(lib, sc, cfg, v, vT, vp, vq, dt=DT, pre=None):
    st = np.ascontiguousarray(sc["state"])
    _, nlev, ncol = st.shape
    arr = [np.ascontiguousarray(sc[k]) for k in KEYS]
    tend = np.empty_like(st)
    dtend = np.empty_like(st)
    lib.cell_host_tangent((ctypes.c_int * 9)(*cfg), ncol, nlev, ctypes.c_double(dt), _p(st),
                          *[_p(a) for a in arr], _p(pre), _p(np.ascontiguousarray(v)),
                          _p(np.ascontiguousarray(vT)), _p(np.ascontiguousarray(vp)),
                          _p(np.ascontiguousarray(vq)), _p(tend), _p(dtend))
    return dtend


def adjoint(lib, sc, cfg, tb, dt=DT, pre=None):
    st = np.ascontiguousarray(sc["state"])
    _, nlev, ncol = st.shape
    arr = [np.ascontiguousarray(sc[k]) for k in KEYS]
    sb = np.full_like(st, np.nan)
    Tb, pb, qb = (np.full((nlev, ncol), np.nan) for _ in range(3))
    lib.cell_host_adjoint((ctypes.c_int * 9)(*cfg), ncol, nlev, ctypes.c_double(dt), _p(st),
                          *[_p(a) for a in arr], _p(pre), _p(np.ascontiguousarray(tb)), _p(sb),
                          _p(Tb), _p(pb), _p(qb))
    return sb, Tb, pb, qb


def jacobians(lib, sc, cfg=ALL):
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
        Jt[:, j] = tangent(lib, sc, cfg, v, one if j == nf else z, one if j == nf + 1 else z,
                           one if j == nf + 2 else z)
    for f in range(nf):
        tb = np.zeros_like(st)
        tb[f] = 1.0
        sb, Tb, pb, qb = adjoint(lib, sc, cfg, tb)
        Ja[f, :nf] = sb
        Ja[f, nf], Ja[f, nf + 1], Ja[f, nf + 2] = Tb, pb, qb
    return Jt, Ja


def row_scaled_error(sc, Jt, Ja, dt=DT):
    st = sc["state"]
    xs = np.concatenate([np.abs(st), np.abs(sc["temperature"])[None],
                         np.abs(sc["pressure"])[None], np.abs(sc["qv"])[None]])
    contrib = np.abs(Jt) * xs[None]
    rowscale = np.maximum(contrib.max(axis=1, keepdims=True), np.abs(st)[:, None] / dt) + 1e-300
    return np.abs(Jt - Ja) * xs[None] / rowscale


@pytest.mark.parametrize("name", ["default", "nucleation"])
def test_reverse_sweep_jacobian_equals_tangent_jacobian(cellhost, name):
    if name == "default":
        sc = scenes.generate(6, 72, seed=0)
    else:
        sc = scenes.generate(6, 72, seed=6, h2so4_range=(1e-11, 1e-9))
    Jt, Ja = jacobians(cellhost, sc)
    assert np.isfinite(Ja).all()
    err = row_scaled_error(sc, Jt, Ja)
    print("%s: row-scaled |J_rev - J_tan| max %.2e, 99.99 %% %.2e, median %.2e"
          % (name, err.max(), np.quantile(err, 0.9999), np.median(err)))
    assert np.quantile(err, 0.9999) < 1e-11
    assert err.max() < 1e-8


def _cases():
    out = []
    sc = scenes.generate(24, 72, seed=5)
    rng = np.random.default_rng(5)
    pre = np.ascontiguousarray(sc["state"] * rng.uniform(0.5, 1.0, sc["state"].shape))
    out.append(("precld", sc, ALL, pre, DT))
    out.append(("precld rename only", sc, (0, 1, 0, 0, 2, 8, 0, 1, 1), pre, DT))
    out.append(("nucleation", scenes.generate(24, 72, seed=6, h2so4_range=(1e-11, 1e-9)),
                ALL, None, DT))
    sc3 = scenes.generate(24, 72, seed=7)
    rng = np.random.default_rng(7)
    sc3["state"][:8] *= rng.choice([0.01, 0.1, 1, 10, 100], size=sc3["state"][:8].shape)
    out.append(("numbers off bounds", sc3, ALL, None, DT))
    out.append(("calcsize only", sc3, (0, 0, 0, 0, 2, 8, 1, 1, 1), None, DT))
    out.append(("calcsize without transfer", sc3, (0, 0, 0, 0, 2, 8, 1, 0, 1), None, DT))
    out.append(("condensation only", sc3, (1, 0, 0, 0, 2, 8, 0, 1, 1), None, DT))
    out.append(("coagulation only, optaa 1", sc3, (0, 0, 0, 1, 1, 8, 0, 1, 1), None, DT))
    sc4 = scenes.generate(24, 72, seed=8)
    sc4["cloud_fraction"][:, :8] = 1.0
    sc4["cloud_fraction"][:, 8:16] = rng.uniform(0.05, 0.95, (72, 8))
    pre4 = np.ascontiguousarray(sc4["state"] * rng.uniform(0.5, 1.0, sc4["state"].shape))
    out.append(("cloud categories + precld", sc4, ALL, pre4, DT))
    for dt in (1.0, 60.0, 3600.0):
        out.append(("dt %g" % dt, scenes.generate(16, 72, seed=9), ALL, None, dt))
    return out


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_dot_product_per_cell(cellhost, case):
    name, sc, cfg, pre, dt = case
    rng = np.random.default_rng(11)
    st = sc["state"]
    nlev, ncol = st.shape[1:]
    v = rng.standard_normal(st.shape) * np.abs(st) * 1e-3
    vT = rng.standard_normal((nlev, ncol)) * 0.1
    vp = rng.standard_normal((nlev, ncol)) * 10.0
    vq = rng.standard_normal((nlev, ncol)) * np.abs(sc["qv"]) * 1e-3
    dtend = tangent(cellhost, sc, cfg, v, vT, vp, vq, dt=dt, pre=pre)
    tb = rng.standard_normal(st.shape) / np.maximum(np.abs(st), 1e-30) * 1e-3
    sb, Tb, pb, qb = adjoint(cellhost, sc, cfg, tb, dt=dt, pre=pre)
    for g in (sb, Tb, pb, qb):
        assert np.isfinite(g).all(), name
    lhs = (tb * dtend).sum(axis=0)
    rhs = (sb * v).sum(axis=0) + Tb * vT + pb * vp + qb * vq
    scale = (np.abs(tb * dtend).sum(axis=0) + np.abs(sb * v).sum(axis=0) + np.abs(Tb * vT)
             + np.abs(pb * vp) + np.abs(qb * vq) + (np.abs(tb) * np.abs(v)).sum(axis=0) / dt
             + 1e-300)
    rel = np.abs(lhs - rhs) / scale
    print("%s: median %.2e, 99.9 %% %.2e, max %.2e, cells > 1e-8: %d of %d"
          % (name, np.median(rel), np.quantile(rel, 0.999), rel.max(), (rel > 1e-8).sum(),
             rel.size))
    assert np.median(rel) < 1e-14
    assert (rel > 1e-8).mean() <= 2e-3
    assert rel.max() < 1e-5


def test_cotangent_is_linear_and_processes_off_leave_round_off(cellhost):
    """The gradient is linear in the cotangent (2 w gives twice the bits of
    w).  With every process and calcsize off the step is the identity up to
    the unit conversions in and out, so the state gradient is round-off of
    |w| / dt and T, p and qv have no influence at all."""
    sc = scenes.generate(8, 72, seed=12)
    rng = np.random.default_rng(12)
    tb = rng.standard_normal(sc["state"].shape)
    g1 = adjoint(cellhost, sc, ALL, tb)
    g2 = adjoint(cellhost, sc, ALL, 2.0 * tb)
    for a, b in zip(g1, g2):
        assert np.array_equal(2.0 * a, b)
    off = (0, 0, 0, 0, 2, 8, 0, 1, 1)
    sb, Tb, pb, qb = adjoint(cellhost, sc, off, tb)
    assert np.abs(sb).max() <= 1e-14 * np.abs(tb).max() / DT
    for g in (Tb, pb, qb):
        assert not np.any(g)
