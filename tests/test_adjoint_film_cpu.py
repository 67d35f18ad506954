"""Spring parameterisation of the interface film (CPU): the window blocks
assembled from the derivative entries J_b must equal the dynamical-matrix
blocks that lattice.phonon_blocks builds at the same springs."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("nx,ny,L,n_iface", [(1, 1, 6, 2), (2, 1, 8, 3)])
def test_window_blocks_linear_in_springs(nx, ny, L, n_iface):
    from phinfortran_amd.adjoint import InterfaceFilm
    kpts = [(0.0, 0.0), (0.25, -0.125)]
    film = InterfaceFilm(None, nx, ny, L, kpts, np.linspace(0.5, 12, 5), n_iface=n_iface)
    assert film.nparam == n_iface * film.layer.nbond // 2
    rng = np.random.default_rng(0)
    beta = 13.8 + rng.standard_normal(film.nparam)
    hd, hv = film.window_blocks(torch.as_tensor(beta))
    for k in range(len(kpts)):
        blk = film.blocks_numpy(beta, k)
        ref_d = blk.Hamd[film.w_lo:film.w_hi + 1]
        ref_v = blk.Hamdv[film.w_lo:film.w_lo + film.nc]
        assert np.abs(hd[k].numpy() - ref_d).max() < 1e-9 * np.abs(ref_d).max()
        assert np.abs(hv[k].numpy() - ref_v).max() < 1e-9 * np.abs(ref_v).max()
        # blocks outside the window do not depend on the trainable springs
        b0 = film.blocks_numpy(np.full(film.nparam, 13.8), k)
        others = [i for i in range(L) if not film.w_lo <= i <= film.w_hi]
        assert np.abs(blk.Hamd[others] - b0.Hamd[others]).max() == 0.0
        oth_v = [i for i in range(L - 1) if not film.w_lo <= i < film.w_lo + film.nc]
        assert np.abs(blk.Hamdv[oth_v] - b0.Hamdv[oth_v]).max() == 0.0


def test_conductance_weights_trapezoid():
    """sum_i w_i T_i reproduces condt.py's trapz of hbar w dn/dT T / (2 pi A)"""
    from phinfortran_amd.adjoint import conductance_weights, HBAR, KB
    f = np.linspace(0.1, 16, 300)
    T = 1.0 + 0.3 * np.sin(f)
    A = 3e-18
    w = f * 2e12 * np.pi
    x = HBAR * w / KB / 300.0
    ex = np.exp(x)
    itd = 1 / A / 2 / np.pi * (HBAR * w * ex / (ex - 1) ** 2 * HBAR * w / KB / 300.0 ** 2 * T)
    ref = np.trapezoid(itd, x=w) if hasattr(np, "trapezoid") else np.trapz(itd, x=w)
    assert abs(np.sum(conductance_weights(f, 300.0, A) * T) - ref) < 1e-12 * abs(ref)
