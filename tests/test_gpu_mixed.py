"""Mixed precision (FP64 decimation, FP32-MFMA sweep) against the FP64 path on
the same inputs.  The tolerance is the measured accuracy class of an FP32
sweep (relative T error below 1e-3 where T > 1e-2); the exact figure per
workload is reported by bench.py."""
import numpy as np
import pytest

from phinfortran_amd import lattice as lt

pytestmark = pytest.mark.gpu


def test_mixed_vs_fp64_film(engine):
    from phinfortran_amd import transport as T
    blk = lt.si_ge_interface_film(nx=2, ny=1, L=6, kfrac=(0.25, 0.0))
    d = engine.upload_blocks(blk)
    E = (np.linspace(0.5, 15.5, 48)) ** 2
    r64 = engine.solve_tdb(d, E, T.params("phonon"))
    r32 = engine.solve_tdb(d, E, T.params("phonon", precision="mixed"))
    t64 = r64["T_LR"].cpu().numpy()
    t32 = r32["T_LR"].cpu().numpy()
    assert (r32["status"].cpu().numpy() == 0).all()
    m = t64 > 1e-2
    err = np.abs(t32[m] - t64[m]) / t64[m]
    print("mixed vs fp64: max rel err", err.max(), "median", np.median(err))
    assert err.max() < 1e-3


def test_mixed_extras_are_fp64_layout(engine):
    from phinfortran_amd import transport as T
    blk = lt.random_blocks(16, 16, 3, seed=2)
    d = engine.upload_blocks(blk)
    E = np.array([0.2, 0.5])
    p64 = T.params("phonon", eta=1e-3, eta0=1e-4, convergence=1e-13)
    p32 = T.params("phonon", eta=1e-3, eta0=1e-4, convergence=1e-13, precision="mixed")
    a = engine.solve_tdb(d, E, p64, extras=("G_D_LR", "G_L"))
    b = engine.solve_tdb(d, E, p32, extras=("G_D_LR", "G_L"))
    for k in ("G_D_LR", "G_L"):
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        assert np.abs(x - y).max() / np.abs(x).max() < 1e-3
