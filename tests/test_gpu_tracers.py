"""mam4x_tracers_to_fields / mam4x_fields_to_tracers on the GPU against the
restatement in oracle/tracers.py (pinned in test_tracers.py).  Integer-
exact work: every comparison is bit for bit, including the tracers the
transfer must leave alone."""
import numpy as np
import pytest

import oracle.tracers as T
from mam4xx_amd import layout as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


def _arrays(rng, ncol, nlev):
    q = rng.standard_normal((ncol, nlev, T.PCNST))
    c = rng.standard_normal((ncol, nlev, T.PCNST))
    return q, c


def _ref_inject(q, c):
    """oracle: state_q/qqcw [ncell, 40] -> the 52 fields [52, ncell]"""
    progs = T.allocated_progs(q.shape[0])
    T.inject_qqcw_to_prognostics(c, progs)
    T.inject_stateq_to_prognostics(q, progs)
    return T.progs_to_state(progs)


def _ref_extract(st, q0, c0, tendency):
    """oracle: fields [52, ncell] into copies of q0/c0 [ncell, 40]"""
    q, c = q0.copy(), c0.copy()
    if tendency:
        others = [np.zeros(q.shape[0])] * 4
    else:  # the gases the chain does not carry keep the host's values
        others = [q0[:, 9], q0[:, 10], q0[:, 12], q0[:, 13]]
    progs = T.state_to_progs(st, others)
    if tendency:
        T.extract_ptend_from_tendencies(progs, q)
    else:
        T.extract_stateq_from_prognostics(progs, q, atm=None)
    T.extract_qqcw_from_prognostics(progs, c)
    return q, c


@pytest.mark.parametrize("layout", ["lev_col", "col_lev"])
@pytest.mark.parametrize("ncol,nlev", [(100, 72), (1, 1), (33, 5)])
def test_inject_matches_oracle(torch_cuda, layout, ncol, nlev):
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    rng = np.random.default_rng(ncol * 7 + nlev)
    q, c = _arrays(rng, ncol, nlev)
    ref = _ref_inject(q.reshape(-1, T.PCNST), c.reshape(-1, T.PCNST))  # cells (icol, k)
    ref = ref.reshape(L.NSTATE, ncol, nlev)
    lay = _abi.LAYOUT_LEV_COL if layout == "lev_col" else _abi.LAYOUT_COL_LEV
    shape = (L.NSTATE, nlev, ncol + 3) if layout == "lev_col" else (L.NSTATE, ncol, nlev + 3)
    out = torch.full(shape, np.nan, dtype=torch.float64, device="cuda")
    fields = out[:, :, :-3]  # padded leading dimension
    ch = Chain()
    ch.tracers_to_fields(torch.as_tensor(q).cuda(), torch.as_tensor(c).cuda(), fields, layout=lay)
    torch.cuda.synchronize()
    got = fields.cpu().numpy()
    if layout == "lev_col":
        got = np.swapaxes(got, 1, 2)
    assert np.array_equal(got, ref)
    assert np.isnan(out[:, :, -3:].cpu().numpy()).all()  # padding untouched
    ch.close()


@pytest.mark.parametrize("mode", ["state", "tendency"])
@pytest.mark.parametrize("stacking", ["by_column", "by_level", "padded"])
def test_extract_matches_oracle(torch_cuda, mode, stacking):
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain
    ncol, nlev = 70, 72
    rng = np.random.default_rng(5)
    q0, c0 = _arrays(rng, ncol, nlev)
    st = rng.standard_normal((L.NSTATE, nlev, ncol))
    qr, cr = _ref_extract(np.swapaxes(st, 1, 2).reshape(L.NSTATE, -1), q0.reshape(-1, T.PCNST),
                          c0.reshape(-1, T.PCNST), mode == "tendency")
    if stacking == "by_column":
        dq, dc = torch.as_tensor(q0).cuda(), torch.as_tensor(c0).cuda()
        view = lambda t: t
    elif stacking == "by_level":  # (nlev, ncol, 40) in memory
        dq = torch.as_tensor(np.ascontiguousarray(np.swapaxes(q0, 0, 1))).cuda()
        dc = torch.as_tensor(np.ascontiguousarray(np.swapaxes(c0, 0, 1))).cuda()
        view = lambda t: t.permute(1, 0, 2)
    else:  # rows of 48 with 8 spare entries
        dq = torch.full((ncol, nlev, 48), 7.0, dtype=torch.float64, device="cuda")
        dc = torch.full((ncol, nlev, 48), 7.0, dtype=torch.float64, device="cuda")
        dq[:, :, :40] = torch.as_tensor(q0).cuda()
        dc[:, :, :40] = torch.as_tensor(c0).cuda()
        view = lambda t: t[:, :, :40]
    ch = Chain()
    m = _abi.PACK_TENDENCY if mode == "tendency" else _abi.PACK_STATE
    ch.fields_to_tracers(torch.as_tensor(st).cuda(), view(dq), view(dc), mode=m)
    torch.cuda.synchronize()
    gq, gc = view(dq).cpu().numpy(), view(dc).cpu().numpy()
    assert np.array_equal(gq.reshape(-1, T.PCNST), qr)
    assert np.array_equal(gc.reshape(-1, T.PCNST), cr)
    if stacking == "padded":
        assert (dq[:, :, 40:].cpu().numpy() == 7.0).all() and (dc[:, :, 40:].cpu().numpy() == 7.0).all()
    ch.close()


def test_absent_planes_and_qqcw_are_left_alone(torch_cuda):
    import torch
    from mam4xx_amd.chain import Chain
    ncol, nlev = 64, 10
    rng = np.random.default_rng(6)
    q0, c0 = _arrays(rng, ncol, nlev)
    st = torch.as_tensor(rng.standard_normal((L.NSTATE, nlev, ncol))).cuda()
    planes = [st[f] if f % 3 else None for f in range(L.NSTATE)]
    dq, dc = torch.as_tensor(q0).cuda(), torch.as_tensor(c0).cuda()
    ch = Chain()
    ch.fields_to_tracers(planes, dq, None)  # no qqcw: cloud-borne fields skipped
    torch.cuda.synchronize()
    gq = dq.cpu().numpy()
    assert np.array_equal(dc.cpu().numpy(), c0)
    for f in range(L.NSTATE):
        if L.is_cloud_borne(f):
            continue
        t = L.pcnst_slot(f) + 9
        want = np.swapaxes(st[f].cpu().numpy(), 0, 1) if f % 3 else q0[:, :, t]
        assert np.array_equal(gq[:, :, t], want), f
    untouched = [t for t in range(T.PCNST)
                 if t not in {L.pcnst_slot(f) + 9 for f in range(L.NSTATE) if f % 3}]
    assert np.array_equal(gq[:, :, untouched], q0[:, :, untouched])
    # and the other direction: absent planes are not written
    out = torch.full((L.NSTATE, nlev, ncol), 3.0, dtype=torch.float64, device="cuda")
    planes = [out[f] if f % 2 else None for f in range(L.NSTATE)]
    ch.tracers_to_fields(torch.as_tensor(q0).cuda(), torch.as_tensor(c0).cuda(), planes)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert (o[0::2] == 3.0).all() and not (o[1::2] == 3.0).any()
    ch.close()


def test_round_trip_full_domain(torch_cuda):
    """21,600 x 72 cells: fields -> tracers -> fields is the identity, and
    the tracers the transfer does not own keep their values."""
    import torch
    from mam4xx_amd.chain import Chain
    ncol, nlev = 21600, 72
    g = torch.Generator(device="cuda").manual_seed(7)
    st = torch.randn((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda", generator=g)
    q = torch.randn((ncol, nlev, T.PCNST), dtype=torch.float64, device="cuda", generator=g)
    c = torch.randn((ncol, nlev, T.PCNST), dtype=torch.float64, device="cuda", generator=g)
    q0, c0 = q.clone(), c.clone()
    ch = Chain()
    ch.fields_to_tracers(st, q, c)
    back = torch.empty_like(st)
    ch.tracers_to_fields(q, c, back)
    torch.cuda.synchronize()
    assert torch.equal(back, st)
    own = sorted({L.pcnst_slot(f) + 9 for f in range(L.NSTATE) if not L.is_cloud_borne(f)})
    other = [t for t in range(T.PCNST) if t not in own]
    assert torch.equal(q[:, :, other], q0[:, :, other])
    assert torch.equal(c[:, :, :15], c0[:, :, :15])
    ch.close()


def test_overlapping_rows_are_rejected(torch_cuda):
    import torch
    from mam4xx_amd.chain import Chain, ChainError
    ch = Chain()
    q = torch.zeros((4, 3, 40), dtype=torch.float64, device="cuda")
    st = torch.zeros((L.NSTATE, 3, 4), dtype=torch.float64, device="cuda")
    bad = q.as_strided((4, 3, 40), (40, 20, 1))  # rows 20 apart overlap
    with pytest.raises(ChainError, match="overlap"):
        ch.tracers_to_fields(bad, None, st)
    with pytest.raises(ChainError, match="mode"):
        ch.fields_to_tracers(st, q, None, mode=5)
    ch.close()
