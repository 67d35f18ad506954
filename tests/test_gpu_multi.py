"""Several devices of uneven block size and length in one call
(nt_solve_tdb_multi_dev) give exactly the per-device batched results, which
the oracle checks on a few points."""
import numpy as np
import pytest

import oracle as O
from phinfortran_amd import lattice as lt

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from phinfortran_amd import transport
    return transport


def test_multi_device_equals_single(engine, T):
    p = T.params("phonon", eta=1e-5, eta0=1e-6)
    op = O.OrParams(p.eta, p.eta0, p.convergence, p.device_eta_factor, p.eta_mode, p.max_iter)
    ks = lt.mp_grid(2, 2)
    specs = [dict(n=1, period=2, L=7), dict(n=2, period=3, L=4), dict(n=1, period=1, L=12)]
    devices, blocks = [], []
    rng = np.random.default_rng(0)
    for i, sp in enumerate(specs):
        blk = [lt.si_ge_superlattice_film(sp["n"], sp["period"], sp["L"], kf, seed=i) for kf in ks]
        blocks.append(blk)
        npts = 9
        E = (rng.uniform(0.5, 14.0, npts)) ** 2
        kidx = rng.integers(0, len(ks), npts).astype(np.int32)
        devices.append(dict(blocks=engine.upload_blocks(blk), E=E, kidx=kidx))
    print("multi-device", flush=True)
    outs = engine.solve_tdb_multi(devices, p)
    for i, (d, o) in enumerate(zip(devices, outs)):
        r = engine.solve_tdb(d["blocks"], d["E"], p, kidx=d["kidx"])
        assert (o["status"].cpu().numpy() == 0).all()
        assert np.abs(o["T_LR"].cpu().numpy() - r["T_LR"].cpu().numpy()).max() < 1e-12
        assert np.abs(o["T_RL"].cpu().numpy() - r["T_RL"].cpu().numpy()).max() < 1e-12
        for j in (0, 4):
            tl, tr, _ = O.solve_tdb(blocks[i][d["kidx"][j]].as_dict(), d["E"][j], op)
            assert abs(o["T_LR"][j].item() - tl) < 1e-9 * max(1, tl)
