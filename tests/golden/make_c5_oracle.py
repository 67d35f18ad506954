"""Writes c5_oracle.json: the CPU oracle's T_LR / T_RL for a few (omega, k)
points of the large mode-resolved workload (Si film 8 x 4 cells, b = 768,
L = 100, seeded 3-layer rough Si/Ge interface), which the oracle needs
minutes per point for.  The GPU test regenerates the same seeded blocks and
compares.  Run from the repository root:  python tests/golden/make_c5_oracle.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O                                     # noqa: E402
from phinfortran_amd import lattice as lt              # noqa: E402

CASES = [dict(kfrac=[0.0625, 0.1875], omega=2.0), dict(kfrac=[0.0625, 0.1875], omega=4.5)]


def main():
    op = O.phonon_params()
    out = dict(workload="rough_interface_film(nx=8, ny=4, L=100, rough_layers=3, seed=0)",
               params=dict(eta=op.eta, eta0=op.eta0, convergence=op.convergence,
                           device_eta_factor=op.device_eta_factor, max_iter=op.max_iter), points=[])
    for c in CASES:
        blk = lt.rough_interface_film(nx=8, ny=4, L=100, rough_layers=3, seed=0, kfrac=tuple(c["kfrac"]))
        t = time.time()
        tl, tr, st = O.solve_tdb(blk.as_dict(), c["omega"] ** 2, op)
        print(c, tl, tr, st, f"{time.time() - t:.0f}s", flush=True)
        out["points"].append(dict(kfrac=c["kfrac"], E=c["omega"] ** 2, T_LR=tl, T_RL=tr, iters=st))
    with open(os.path.join(ROOT, "tests", "golden", "c5_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
