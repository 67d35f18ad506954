"""Devices with the same leads and points in one multi-device call: the
engine computes their lead Green's functions once and runs each device's
sweep on them.  Results equal the per-device solves (which the oracle pins)
and the oracle on sampled points."""
import numpy as np
import pytest

import oracle as O
from phinfortran_amd import lattice as lt

pytestmark = pytest.mark.gpu


def test_shared_leads_equal_single_solves(engine):
    from phinfortran_amd import transport as T
    p = T.params("phonon")
    op = O.OrParams(p.eta, p.eta0, p.convergence, p.device_eta_factor, p.eta_mode, p.max_iter)
    ks = lt.mp_grid(2, 2)
    f = 16.0 * np.arange(1, 33) / 32
    E = np.tile(f ** 2, len(ks))
    kidx = np.repeat(np.arange(len(ks), dtype=np.int32), len(f))
    # three n = 1 superlattice films (same pristine Si leads) and one n = 2
    specs = [dict(n=1, period=2, L=9, seed=1), dict(n=1, period=3, L=14, seed=2), dict(n=2, period=2, L=6, seed=3),
             dict(n=1, period=1, L=5, seed=4)]
    devices, blocks = [], []
    for sp in specs:
        blk = [lt.si_ge_superlattice_film(sp["n"], sp["period"], sp["L"], kf, seed=sp["seed"]) for kf in ks]
        blocks.append(blk)
        devices.append(dict(blocks=engine.upload_blocks(blk), E=E, kidx=kidx))
    print("multi-device, shared leads", flush=True)
    outs = engine.solve_tdb_multi(devices, p)
    for i, (d, o) in enumerate(zip(devices, outs)):
        r = engine.solve_tdb(d["blocks"], E, p, kidx=kidx)
        assert (o["status"].cpu().numpy() == 0).all()
        for key in ("T_LR", "T_RL"):
            a, b = o[key].cpu().numpy(), r[key].cpu().numpy()
            assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max()), (i, key)
        for j in (3, 40, 101):
            tl, tr, _ = O.solve_tdb(blocks[i][kidx[j]].as_dict(), E[j], op)
            assert abs(o["T_LR"][j].item() - tl) < 1e-9 * max(1, tl)
            assert abs(o["T_RL"][j].item() - tr) < 1e-9 * max(1, tr)
