"""Adjoint gradient of the thermal conductance with respect to the interface
masses at the large mode-resolved workload's shape: Si film of 8x4 cells
(768 dof per layer), 100 device layers, seeded 3-layer rough Si/Ge interface
(768 trainable masses).

The gradient (adjoint solve over the interface layers, contracted with
dD/dm = -D/(2m) per row and column of each atom) is checked against central
finite differences of the GPU FP64 forward solve (pinned to the oracle at
this shape by the configuration parity tests) on 4 seeded masses,
h = 1e-3 amu, to 1e-5 relative; with checkpointed recomputation of the layer
Green's functions (about every sqrt(L - lo)-th layer kept) the gradient and
T are the ones of the full storage, at a third of the device memory."""
import numpy as np
import pytest

from phinfortran_amd import lattice as lt

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mfilm(engine):
    from phinfortran_amd import transport as T
    from phinfortran_amd.adjoint import MassFilm
    print("mass film 8x4 L=100: building", flush=True)
    f = np.array([0.05, 1.3, 2.9])          # the first frequency is the reference's ie = 1 point (T = 0)
    return MassFilm(engine, 8, 4, 100, [(0.3125, 0.4375)], f, rough_layers=3, seed=0, params=T.params("phonon"))


def test_mass_gradient_vs_fd(engine, mfilm):
    import torch
    assert mfilm.b == 768 and mfilm.nparam == 768
    m = mfilm.initial()
    engine.set_adjoint_checkpoint(0)
    full_bytes = engine.adjoint_bytes_per_point(mfilm.dblocks, mfilm.w_lo, mfilm.w_hi)
    G, g, tl = mfilm.conductance(m)
    G0, _, _ = mfilm.conductance(m, with_grad=False)
    assert abs(float(G) - float(G0)) <= 1e-12 * abs(float(G0))
    g = g.cpu().numpy()
    rng = np.random.default_rng(3)
    idx = rng.choice(mfilm.nparam, 4, replace=False)
    h = 1e-3
    scale = np.abs(g).max()
    worst = 0.0
    for n in idx:
        e = torch.zeros_like(m)
        e[n] = h
        Gp, _, _ = mfilm.conductance(m + e, with_grad=False)
        Gm, _, _ = mfilm.conductance(m - e, with_grad=False)
        fd = (float(Gp) - float(Gm)) / (2 * h)
        err = abs(fd - g[n]) / max(abs(fd), 1e-3 * scale)
        worst = max(worst, err)
        assert err < 1e-5, (int(n), fd, g[n])
    # checkpointed recomputation: same T and gradient, less memory
    engine.set_adjoint_checkpoint(-1)
    ck_bytes = engine.adjoint_bytes_per_point(mfilm.dblocks, mfilm.w_lo, mfilm.w_hi)
    Gc, gc, tlc = mfilm.conductance(m)
    engine.set_adjoint_checkpoint(0)
    gc = gc.cpu().numpy()
    assert torch.equal(tlc, tl)
    assert np.abs(gc - g).max() <= 1e-12 * scale
    print(f"mass film: G = {float(G):.6e} W/m^2/K, max |dG/dm| {scale:.3e}, worst FD mismatch {worst:.2e}; "
          f"adjoint storage per point {full_bytes / 2**20:.0f} MiB, checkpointed {ck_bytes / 2**20:.0f} MiB",
          flush=True)
    assert ck_bytes < 0.5 * full_bytes
