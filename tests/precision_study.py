"""Precision study: 21,600 x 72 cells, 10 chained 1800 s steps.

Steps the bench scene with the FP64 and the mixed GPU builds and with the
CPU oracle, chaining each the same way (state += dt * tendencies; calcsize's
dry diameter feeds the next step's dry and wet diameters), and prints, for
every mode number, species mass and gas, the maximum and median relative
error of the final state against the oracle's, relative to the quantity's
magnitude over the run:
  |x_gpu - x_oracle| / max(|x_oracle|, |x_initial|, floor),
floor 1e-10 #/kg for numbers and 1e-30 kg/kg for masses (a species the run
depletes to round-off is measured against what it started with).
Run on a GPU box from the repository root:  python tests/precision_study.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

NCOL, NLEV, DT, NSTEP = 21600, 72, 1800.0, 10


def gpu_run(sc, prec):
    import torch
    from mam4xx_amd import _abi
    from mam4xx_amd import layout as L
    from mam4xx_amd.chain import Chain, scene_to_device
    d = scene_to_device(sc)
    d["dgncur_awet"] = d["dgncur_a"]
    ch = Chain(precision=prec)
    tend = torch.empty((L.NSTATE, NLEV, NCOL), dtype=torch.float64, device="cuda")
    dgc = torch.empty((4, NLEV, NCOL), dtype=torch.float64, device="cuda")
    for _ in range(NSTEP):
        ch.compute_tendencies(d, DT, tend, dgn_i=d["dgncur_a"], dgn_c=dgc, state_out=d["state"])
    torch.cuda.synchronize()
    return d["state"].cpu().numpy()


def oracle_run(sc):
    import oracle
    orc = oracle.Oracle()
    s = {k: np.array(v) for k, v in sc.items()}
    for _ in range(NSTEP):
        t, di, _, _ = orc.chain_step(s, dt=DT)
        s["state"] = s["state"] + DT * t
        s["dgncur_a"] = di.copy()
        s["dgncur_awet"] = di.copy()
    return s["state"]


def main():
    from mam4xx_amd import _abi
    from mam4xx_amd import layout as L
    from mam4xx_amd import scenes
    sc = scenes.generate(NCOL, NLEV, seed=0)
    ref = oracle_run(sc)
    res = {"fp64": gpu_run(sc, _abi.PREC_FP64), "mixed": gpu_run(sc, _abi.PREC_MIXED)}
    floor = np.where(np.arange(L.NSTATE) < 8, 1e-10, 1e-30)[:, None, None]
    den = np.maximum(np.maximum(np.abs(ref), np.abs(sc["state"])), floor)
    names = L.field_names()
    print("%-14s %12s %12s %12s %12s" % ("field", "fp64 max", "fp64 median", "mixed max",
                                          "mixed median"))
    for f in range(L.NSTATE):
        row = []
        for k in ("fp64", "mixed"):
            e = np.abs(res[k][f] - ref[f]) / den[f]
            row += [e.max(), np.median(e)]
        print("%-14s %12.2e %12.2e %12.2e %12.2e" % (names[f], *row))
    for k in ("fp64", "mixed"):
        e = np.abs(res[k] - ref) / den
        cell = e.max(axis=0)
        print("%s: all fields max %.2e, median %.2e; cells with any field above 1e-5: %.3f %%, "
              "above 1e-3: %.3f %%" % (k, e.max(), np.median(e), 100 * (cell > 1e-5).mean(),
                                        100 * (cell > 1e-3).mean()))


if __name__ == "__main__":
    main()
