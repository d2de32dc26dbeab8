"""Per-process tendency diagnostics (mam4x_outputs.process_tendencies: the
gas_aero_exchange_* arrays of modal_aero_amicphys_intr) on the GPU against
the oracle's orc_chain_step_proc.  The oracle's blocks are pinned by budgets
and conservation laws in test_process_diagnostics.py, and the kernel's cell
code reproduces them bit for bit when built for the CPU (test_cell_host.py).

Bar, as for the tendencies (test_gpu_parity.py): a diagnostic is measured on
the magnitude of its tracer in the cell, |diag_gpu - diag_ref| * dt /
max(|q_in|, |q_post|, floor) with q in mol/mol (#/kmol), and must be
<= 1e-10, or within 10x the oracle's own spread in cells where single-ulp
libm perturbations move the oracle's diagnostics by more than 1e-11.
Tracer slots outside MAM4 stay exactly 0.  Forming the diagnostics must not
change the step: the tendencies are bitwise those of the plain step.
"""
import numpy as np
import pytest

from mam4xx_amd import layout as L
from mam4xx_amd import scenes

pytestmark = pytest.mark.gpu

DT = 1800.0
TOL = 1e-10
SENSITIVE = 1e-11
ENV_FACTOR = 10.0


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch


def _scene(ncol, seed):
    sc = scenes.generate(ncol, 72, seed=seed)
    rng = np.random.default_rng(seed)
    # all cloud categories of setup_subareas (mam4_amicphys.hpp:241-317)
    sc["cloud_fraction"][:, : ncol // 6] = rng.uniform(0.9991, 1.0, (72, ncol // 6))
    sc["cloud_fraction"][:, ncol // 6: ncol // 3] = rng.uniform(0.05, 0.95, (72, ncol // 6))
    sc["cloud_fraction"][:, ncol // 3: ncol // 2] = 0.0
    # cloud-chemistry growth: renaming in the cloudy subarea
    pre = sc["state"] * rng.uniform(0.5, 1.0, sc["state"].shape)
    return sc, pre


def _gpu(torch, sc, pre, schedule, process=True, layout=None, blocks=None, pad=0, **cfg):
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    ch = Chain(schedule=schedule, **cfg)
    nlev, ncol = sc["temperature"].shape
    d = scene_to_device(sc)
    tend = torch.empty((L.NSTATE, nlev, ncol), dtype=torch.float64, device="cuda")
    proc = None
    if process:
        full = torch.full((5, 31, nlev + pad, ncol), np.nan, dtype=torch.float64, device="cuda")
        proc = full[:, :, :nlev]
        if blocks is not None:
            proc = [proc[k] if k in blocks else None for k in range(5)]
    ch.compute_tendencies(d, DT, tend, precldchem=torch.as_tensor(pre).cuda(), process=proc)
    torch.cuda.synchronize()
    ch.close()
    t = tend.cpu().numpy()
    if not process:
        return t, None
    return t, full[:, :, :nlev].cpu().numpy()


def _oracle(orc, sc, pre, seeds=(1, 2, 3, 4, 5, 6), **cfg):
    """oracle diagnostics and their spread under libm noise (per element)"""
    import oracle.pyoracle as P
    oc = P.default_config(**cfg)
    t, _, _, _, base = orc.chain_step(sc, DT, config=oc, precldchem=pre, proc=True)
    env = np.zeros_like(base)
    try:
        for s in seeds:
            orc.set_libm_noise(s)
            pr = orc.chain_step(sc, DT, config=oc, precldchem=pre, proc=True)[4]
            env = np.maximum(env, np.abs(pr - base))
    finally:
        orc.set_libm_noise(0)
    return t, base, env


def _scales(sc, t):
    """[5][31][nlev][ncol] magnitudes of each block's tracers in mol/mol;
    NaN where a block has no such tracer"""
    st = sc["state"]
    fl = np.where(np.arange(L.NSTATE) < 8, 1e-10, 1e-30)[:, None, None]
    mag = np.maximum(np.maximum(np.abs(st), np.abs(st + DT * t)), fl)
    out = np.full((5, 31) + st.shape[1:], np.nan)
    for f in range(L.NSTATE):
        v = mag[f] * L.to_vmr(f)
        if L.is_cloud_borne(f):
            out[4, L.pcnst_slot(f)] = v
        else:
            out[:4, L.pcnst_slot(f)] = v
    return out


def _compare(g, ref, env, scale):
    none = np.isnan(scale)
    assert np.all(g[none] == 0.0), "a slot without a MAM4 tracer is not zero"
    s = np.where(none, 1.0, scale)
    err = np.where(none, 0.0, np.abs(g - ref) * DT / s)
    cell_env = np.where(none, 0.0, env * DT / s).max(axis=(0, 1))
    bound = np.where(cell_env > SENSITIVE, np.maximum(TOL, ENV_FACTOR * cell_env), TOL)
    bad = np.argwhere(err > bound[None, None])
    assert len(bad) == 0, ("%d out of bound, first [block, slot, lev, col] %s err %.3e "
                           "bound %.3e" % (len(bad), bad[0], err[tuple(bad[0])],
                                           bound[tuple(bad[0][2:])]))
    return err


@pytest.mark.parametrize("sched", ["queue", "fused"])
def test_process_tendencies_match_oracle(orc, torch_cuda, sched):
    from mam4xx_amd import _abi
    schedule = _abi.SCHED_QUEUE if sched == "queue" else _abi.SCHED_FUSED
    sc, pre = _scene(96, 21)
    tg, pg = _gpu(torch_cuda, sc, pre, schedule)
    to, po, env = _oracle(orc, sc, pre)
    assert np.isfinite(pg).all()
    err = _compare(pg, po, env, _scales(sc, to))
    print("worst error per block:", err.reshape(5, -1).max(axis=1))
    for k in range(5):  # every block is exercised
        assert np.count_nonzero(pg[k]) > 100, L.PROCESS_NAMES[k]


@pytest.mark.parametrize("proc", ["do_cond", "do_rename", "do_newnuc", "do_coag"])
def test_single_process_blocks(orc, torch_cuda, proc):
    from mam4xx_amd import _abi
    cfg = dict(do_cond=0, do_rename=0, do_newnuc=0, do_coag=0)
    cfg[proc] = 1
    sc, pre = _scene(48, 22)
    _, pg = _gpu(torch_cuda, sc, pre, _abi.SCHED_QUEUE, **cfg)
    to, po, env = _oracle(orc, sc, pre, **cfg)
    _compare(pg, po, env, _scales(sc, to))


def test_diagnostics_leave_the_step_unchanged(torch_cuda):
    from mam4xx_amd import _abi
    sc, pre = _scene(320, 23)
    for schedule in (_abi.SCHED_QUEUE, _abi.SCHED_FUSED):
        a, _ = _gpu(torch_cuda, sc, pre, schedule, process=False)
        b, _ = _gpu(torch_cuda, sc, pre, schedule)
        assert np.array_equal(a, b)


def test_partial_blocks_stride_and_layout(torch_cuda):
    """A subset of blocks, a tracer-plane stride larger than the plane, and
    the [ncol][nlev] layout give bitwise the values of the full request."""
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, scene_to_device
    sc, pre = _scene(60, 24)
    _, full = _gpu(torch_cuda, sc, pre, _abi.SCHED_QUEUE)
    _, part = _gpu(torch_cuda, sc, pre, _abi.SCHED_QUEUE, blocks=(1, 3), pad=5)
    for k in range(5):
        if k in (1, 3):
            assert np.array_equal(part[k], full[k])
        else:
            assert np.isnan(part[k]).all()  # not requested: untouched
    tr = lambda v: np.ascontiguousarray(np.swapaxes(v, -1, -2))
    sct = {k: (tr(v) if v.ndim >= 2 else v) for k, v in sc.items()}
    ch = Chain(schedule=_abi.SCHED_QUEUE)
    tend = torch.empty((L.NSTATE, 60, 72), dtype=torch.float64, device="cuda")
    proc = torch.full((5, 31, 60, 72), np.nan, dtype=torch.float64, device="cuda")
    ch.compute_tendencies(scene_to_device(sct), DT, tend, precldchem=torch.as_tensor(tr(pre)).cuda(),
                          layout=_abi.LAYOUT_COL_LEV, process=proc)
    torch.cuda.synchronize()
    ch.close()
    assert np.array_equal(np.swapaxes(proc.cpu().numpy(), -1, -2), full)


def test_process_tendencies_need_fp64(torch_cuda):
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd.chain import Chain, ChainError, scene_to_device
    sc = scenes.generate(16, 8, seed=25)
    d = scene_to_device(sc)
    tend = torch.empty((L.NSTATE, 8, 16), dtype=torch.float64, device="cuda")
    proc = torch.empty((5, 31, 8, 16), dtype=torch.float64, device="cuda")
    ch = Chain(precision=_abi.PREC_MIXED)
    with pytest.raises(ChainError, match="FP64"):
        ch.compute_tendencies(d, DT, tend, process=proc)
    ch.close()
