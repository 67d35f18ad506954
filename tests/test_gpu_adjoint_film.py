"""Adjoint gradient of the thermal conductance at the differentiation
workload's cross-section: Si|Ge interface film of 4x2 conventional cells
(192 dof per layer), 40 device layers, 2048 trainable per-bond angular
springs (the bonds owned by the 16 layers around the interface),
initialised to 13.8 + N(0, 0.5) N/m from the seed.

dG/dbeta from the adjoint solve against central finite differences
(h = 1e-4 N/m) of the GPU FP64 forward solve, which the configuration parity
tests pin to the oracle at this shape, on 32 seeded bonds, agreement to 1e-5
relative; then Adam steps maximising G.  The (omega, k) sample is two
k-points of the 8x8 grid times 24 of the 1024 frequencies."""
import numpy as np
import pytest

from phinfortran_amd import lattice as lt

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def film(engine):
    import torch
    from phinfortran_amd import transport as T
    from phinfortran_amd.adjoint import InterfaceFilm
    ks = lt.mp_grid(8, 8)
    f = (16.0 * np.arange(1, 1025) / 1024)[20::43]
    print(f"film 4x2 L=40: {len(f)} frequencies, building the 2048-bond window", flush=True)
    fl = InterfaceFilm(engine, 4, 2, 40, [ks[9], ks[36]], f, n_iface=16, params=T.params("phonon"))
    assert fl.b == 192 and fl.nparam == 2048
    rng = np.random.default_rng(0)
    fl.beta_init = torch.as_tensor(13.8 + 0.5 * rng.standard_normal(fl.nparam), device=fl.dev)
    return fl


def test_film_2048_bonds_gradient_vs_fd(film):
    import torch
    beta = film.beta_init.clone()
    G, g, tl = film.conductance(beta)
    g = g.cpu().numpy()
    G0, _, tl0 = film.conductance(beta, with_grad=False)
    # the adjoint solve's T_LR is the forward's
    assert abs(float(G) - float(G0)) <= 1e-12 * abs(float(G0))
    rng = np.random.default_rng(1)
    bonds = rng.choice(film.nparam, 32, replace=False)
    h = 1e-4
    scale = np.abs(g).max()
    worst = 0.0
    for n in bonds:
        e = torch.zeros_like(beta)
        e[n] = h
        Gp, _, _ = film.conductance(beta + e, with_grad=False)
        Gm, _, _ = film.conductance(beta - e, with_grad=False)
        fd = (float(Gp) - float(Gm)) / (2 * h)
        err = abs(fd - g[n]) / max(abs(fd), 1e-3 * scale)
        worst = max(worst, err)
        assert err < 1e-5, (int(n), fd, g[n])
    print(f"2048-bond film: G = {float(G):.6e} W/m^2/K, max |dG/dbeta| {scale:.3e}, "
          f"worst relative FD mismatch over 32 bonds {worst:.2e}", flush=True)


def test_film_2048_bonds_adam(film):
    from phinfortran_amd.adjoint import maximize_conductance
    beta, hist = maximize_conductance(film, film.beta_init.clone(), steps=50, lr=0.2)
    print(f"Adam, 50 steps: G {hist[0]:.6e} -> {hist[-1]:.6e}", flush=True)
    assert len(hist) == 50 and hist[-1] > hist[0]
