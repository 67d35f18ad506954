"""Batched complex eigen-decomposition and matrix square root (the LAPACK
calls of the mode-resolved transmission) against numpy/scipy."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    return torch, torch.device("cuda", 0)


def to_dev(torch, dev, mats):
    a = np.ascontiguousarray(np.swapaxes(np.stack(mats), -1, -2))
    return torch.from_numpy(a).to(dev)


def from_dev(t):
    return np.swapaxes(t.cpu().numpy(), -1, -2)


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def lead_F(n, seed):
    """F = g V of a random lead: eigenvalues on and inside the unit circle,
    V of low rank (many zero eigenvalues), as in the mode-resolved path."""
    import oracle as O
    rng = np.random.default_rng(seed)
    H = crand(rng, n, n) * 0.3
    H = H + H.conj().T
    V = np.zeros((n, n), complex)
    k = max(1, n // 3)
    V[:, -k:] = crand(rng, n, k) * 0.3
    op = O.OrParams(1e-6, 1e-7, 1e-13, 10.0, 0, 2000)
    sg = O.surface_green(H, V, H, V.conj().T, 0.4, op, want_plus_minus=False)
    return sg["gl_adv"] @ V


def geev(engine, torch, dev, mats):
    from phinfortran_amd import _native as N
    n = mats[0].shape[0]
    b = len(mats)
    dA = to_dev(torch, dev, mats)
    dV = torch.zeros_like(dA)
    dw = torch.zeros((b, n), dtype=torch.complex128, device=dev)
    st = torch.zeros(b, dtype=torch.int32, device=dev)
    rc = N.lib().nt_zgeev_batched_dev(engine.ctx, n, C.c_void_p(dA.data_ptr()), n * n, C.c_void_p(dw.data_ptr()),
                                      C.c_void_p(dV.data_ptr()), n * n, C.c_void_p(st.data_ptr()), b)
    assert rc == 0
    engine.sync()
    return dw.cpu().numpy(), from_dev(dV), st.cpu().numpy(), from_dev(dA)


@pytest.mark.parametrize("n", [1, 2, 5, 24, 48, 96, 192, 300])
def test_zgeev_random(engine, torch_dev, n):
    torch, dev = torch_dev
    rng = np.random.default_rng(n)
    mats = [crand(rng, n, n) for _ in range(3)]
    print(f"zgeev n={n}", flush=True)
    w, V, st, A_after = geev(engine, torch, dev, mats)
    assert (st == 0).all()
    for i, A in enumerate(mats):
        assert np.abs(A_after[i] - A).max() == 0.0          # input preserved
        scale = np.abs(A).max()
        assert np.abs(A @ V[i] - V[i] * w[i]).max() < 1e-12 * scale * n
        ref = np.linalg.eigvals(A)
        # same spectrum (each computed eigenvalue matches a reference one)
        d = np.abs(w[i][:, None] - ref[None, :]).min(axis=1)
        assert d.max() < 1e-11 * scale * n
        # ZGEEV normalisation
        assert np.abs(np.linalg.norm(V[i], axis=0) - 1).max() < 1e-13
        k = np.argmax(np.abs(V[i]) ** 2, axis=0)
        top = V[i][k, np.arange(n)]
        assert np.all(top.imag == 0.0) and np.all(top.real > 0)


@pytest.mark.parametrize("n", [12, 48, 96])
def test_zgeev_lead_transfer_matrix(engine, torch_dev, n):
    torch, dev = torch_dev
    F = lead_F(n, n)
    w, V, st, _ = geev(engine, torch, dev, [F])
    assert st[0] == 0
    ref = np.linalg.eigvals(F)
    assert np.abs(F @ V[0] - V[0] * w[0]).max() < 1e-11 * max(1.0, np.abs(F).max())
    # propagating count (|lambda| = 1 within lambda_eta) as the reference selects
    sel = lambda lam: int(np.sum(np.abs(np.abs(lam) - 1.0) < 5e-3))
    assert sel(w[0]) == sel(ref)
    prop = np.sort(np.angle(w[0][np.abs(np.abs(w[0]) - 1) < 5e-3]))
    prop_ref = np.sort(np.angle(ref[np.abs(np.abs(ref) - 1) < 5e-3]))
    assert np.abs(prop - prop_ref).max() < 1e-9


@pytest.mark.parametrize("n", [1, 4, 33, 96, 192])
def test_zsqrtm(engine, torch_dev, n):
    import scipy.linalg as sl
    from phinfortran_amd import _native as N
    torch, dev = torch_dev
    rng = np.random.default_rng(7 + n)
    # skew-Hermitian velocity-like matrices and general ones
    X = crand(rng, n, n)
    mats = [1j * (X + X.conj().T) / 2, crand(rng, n, n) + 2 * np.eye(n)]
    dA = to_dev(torch, dev, mats)
    dS = torch.zeros_like(dA)
    st = torch.zeros(2, dtype=torch.int32, device=dev)
    print(f"zsqrtm n={n}", flush=True)
    rc = N.lib().nt_zsqrtm_batched_dev(engine.ctx, n, C.c_void_p(dA.data_ptr()), n * n, C.c_void_p(dS.data_ptr()),
                                       n * n, C.c_void_p(st.data_ptr()), 2)
    assert rc == 0
    engine.sync()
    assert (st.cpu().numpy() == 0).all()
    S = from_dev(dS)
    for i, A in enumerate(mats):
        ref = sl.sqrtm(A)
        assert np.abs(S[i] @ S[i] - A).max() < 1e-11 * np.abs(A).max() * n
        assert np.abs(S[i] - ref).max() < 1e-9 * np.abs(ref).max()
