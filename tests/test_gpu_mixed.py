"""Mixed precision (FP64 decimation; FP32-MFMA factorisation of each sweep
direction with FP64 iterative refinement of its block column; FP64 recompute
of points that fail the error estimate) against the FP64 path on the same
inputs.

Bar: relative T error <= 1e-6 wherever T > 1e-3 max T (the bench's
criterion), and |dT| <= 1e-6 max T everywhere."""
import numpy as np
import pytest

from phinfortran_amd import lattice as lt

pytestmark = pytest.mark.gpu

REL = 1e-6


def compare(name, engine, T, d, E, kidx=None, extras=()):
    r64 = engine.solve_tdb(d, E, T.params("phonon"), kidx=kidx, extras=extras)
    engine.profile_reset()
    r32 = engine.solve_tdb(d, E, T.params("phonon", precision="mixed"), kidx=kidx, extras=extras)
    counts = engine.mixed_counts()
    assert (r64["status"].cpu().numpy() == 0).all()
    assert (r32["status"].cpu().numpy() == 0).all()
    worst = 0.0
    for key in ("T_LR", "T_RL"):
        t64, t32 = r64[key].cpu().numpy(), r32[key].cpu().numpy()
        big = t64 > 1e-3 * t64.max()
        err = np.abs(t32[big] - t64[big]) / t64[big]
        worst = max(worst, err.max())
        assert np.abs(t32 - t64).max() <= REL * max(1.0, t64.max()), (name, key)
    print(f"{name}: max rel dT {worst:.2e}; points / second steps / FP64 recomputes {counts}", flush=True)
    assert worst <= REL, (name, worst)
    assert counts[0] == len(E)
    return r64, r32


def test_mixed_vs_fp64_film(engine):
    from phinfortran_amd import transport as T
    blk = lt.si_ge_interface_film(nx=2, ny=1, L=6, kfrac=(0.25, 0.0))
    E = (np.linspace(0.5, 15.5, 48)) ** 2
    compare("film b=48 L=6", engine, T, engine.upload_blocks(blk), E)


def test_mixed_vs_fp64_headline_film(engine):
    """b = 192, L = 40 interface film, one k-point, every 16th of the 1024
    benchmark frequencies."""
    from phinfortran_amd import transport as T
    ks = lt.mp_grid(8, 8)
    blk = lt.si_ge_interface_film(nx=4, ny=2, L=40, kfrac=ks[27])
    E = (16.0 * np.arange(1, 1025) / 1024)[7::16] ** 2
    compare("film b=192 L=40", engine, T, engine.upload_blocks(blk), E)


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_mixed_vs_fp64_uneven(engine, n):
    """one superlattice device per block size of the uneven-batching workload"""
    from phinfortran_amd import transport as T
    spec = next(d for d in lt.superlattice_set(256, seed=0) if d["n"] == n)
    ks = lt.mp_grid(4, 4)
    blocks = [lt.si_ge_superlattice_film(spec["n"], spec["period"], spec["L"], ks[k], seed=spec["seed"])
              for k in (0, 5)]
    npts = 32 if n < 4 else 16
    E = (16.0 * np.arange(1, 257) / 256)[np.linspace(2, 253, npts).astype(int)] ** 2
    kidx = (np.arange(npts) % 2).astype(np.int32)
    compare(f"uneven b={spec['b']} L={spec['L']}", engine, T, engine.upload_blocks(blocks), E, kidx=kidx)


def test_mixed_extras_are_fp64_layout(engine):
    from phinfortran_amd import transport as T
    blk = lt.random_blocks(16, 16, 3, seed=2)
    d = engine.upload_blocks(blk)
    E = np.array([0.2, 0.5])
    p64 = T.params("phonon", eta=1e-3, eta0=1e-4, convergence=1e-13)
    p32 = T.params("phonon", eta=1e-3, eta0=1e-4, convergence=1e-13, precision="mixed")
    keys = ("G_D_LR", "G_D_RL", "G_L", "G_R")
    a = engine.solve_tdb(d, E, p64, extras=keys)
    b = engine.solve_tdb(d, E, p32, extras=keys)
    for k in keys:
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        assert np.abs(x - y).max() / np.abs(x).max() < 1e-7, k
