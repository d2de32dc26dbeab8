"""The per-cell flag for the SOA loop's 1000-sub-step cap
(mam4x_outputs.soa_capped; the reference warns and goes on,
gasaerexch_soaexch.hpp:93, 170-173), on the CPU build of the kernel's code:
the flag is 1 exactly where the cell's loop stopped at the cap.
"""
import ctypes

import numpy as np

from mam4xx_amd import scenes

from .test_adjoint_host import ALL
from .test_cell_host import cellhost  # noqa: F401  (fixture)

DT = 1800.0


def _single_cell(sc, k, i):
    out = {}
    for key, v in sc.items():
        if key == "pblh":
            out[key] = np.ascontiguousarray(v[i:i + 1])
        elif v.ndim == 3:
            out[key] = np.ascontiguousarray(v[:, k:k + 1, i:i + 1])
        else:
            out[key] = np.ascontiguousarray(v[k:k + 1, i:i + 1])
    return out


def capped_scene(ncol=8, seed=0):
    """A scene with 100x the interstitial numbers: uptake fast enough that
    some SOA loops need more than their 1000 sub-steps."""
    sc = scenes.generate(ncol, 72, seed=seed)
    sc["state"][:4] *= 100.0
    return sc


def host_capped(lib, sc, ns=1):
    """The SOA-cap flag plane of the kernel's CPU build."""
    from .test_cell_host import run_host
    flags = np.full(sc["temperature"].shape, -1, dtype=np.int32)
    lib.cell_host_set_soacap(flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    lib.cell_host_set_ntsubstep(ns)
    try:
        run_host(lib, sc, DT, ALL)
    finally:
        lib.cell_host_set_soacap(None)
        lib.cell_host_set_ntsubstep(1)
    return flags


def test_soa_cap_flag_marks_the_capped_loops(cellhost):
    """The flag is 1 exactly where the cell's loop ran its 1000 sub-steps
    (its sub-step count, from the cell run alone, is the cap's 1001)."""
    from .test_cell_host import run_host
    sc = capped_scene()
    flags = host_capped(cellhost, sc)
    assert set(np.unique(flags)) <= {0, 1}
    nlev, ncol = flags.shape
    for k in range(0, nlev, 6):
        for i in range(ncol):
            it = run_host(cellhost, _single_cell(sc, k, i), DT, ALL)[3]
            assert flags[k, i] == (1 if it > 1000 else 0), (k, i, it)
    assert 0 < flags.mean() < 1
    # a scene of ordinary numbers never reaches the cap
    assert not host_capped(cellhost, scenes.generate(8, 72, seed=0)).any()
