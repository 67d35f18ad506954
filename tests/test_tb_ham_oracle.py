"""The tight-binding H(k) oracle (oracle/tb_ham.py) against what the
mathematics fixes: covariance of the Slater-Koster table under rotations of
the bond, its transposition symmetry, the power law of the two-centre
integrals, Hermiticity and Bloch periodicity of H(k), and the band edges of
bulk Si with the example's sp3d5s* parameters (examples/tb/param, copied as
data)."""
import os

import numpy as np
import pytest

from oracle import tb_ham as TB

HERE = os.path.dirname(os.path.abspath(__file__))
PARAM = os.path.join(HERE, "golden", "tb", "param")
ORBS = ["s", "px", "py", "pz", "s*", "dxy", "dyz", "dxz", "dx2-y2", "dz2"]


@pytest.fixture(scope="module")
def P():
    return TB.read_param(PARAM)


def test_param_reader(P):
    assert P["orbitals"] == [ORBS, ORBS] and P["orb_num"] == [10, 10]
    V = P["V"]
    assert V[0, 0, 0, 0, 0] == -3.40253 and V[0, 0, 3, 3, 0] == 23.88167 and V[1, 1, 2, 2, 0] == 14.01053
    # hopping block "1 0" = species (2, 1); its line "0 1 0 2.82890 2.37280": s-p sigma
    assert V[1, 0, 0, 1, 1] == 2.82890 and P["Vn"][1, 0, 0, 1, 1] == 2.37280
    # "5 5 0" is s*-s* sigma, "1 1 1" p-p pi, "2 2 2" d-d delta
    assert V[0, 0, 3, 3, 1] == -4.86118 and V[0, 0, 1, 1, 2] == -1.58305 and V[0, 0, 2, 2, 3] == -1.19386
    assert np.allclose(P["r_d"], [[2.44999e-10, 2.39792e-10], [2.39792e-10, 2.35169e-10]])


def block(P, i, j, r):
    return np.array([[TB.v_sk(P, i, j, a, b, r) for b in ORBS] for a in ORBS])


def rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def d_functions(u):
    """the SK angular functions of the d orbitals (xy, yz, xz, x2-y2, z2)"""
    x, y, z = u
    s3 = np.sqrt(3.0)
    return np.array([s3 * x * y, s3 * y * z, s3 * x * z, s3 / 2 * (x * x - y * y), z * z - (x * x + y * y) / 2])


def orbital_rotation(R):
    """D(R) in the orbital order s px py pz s* d(5): (D f)(u) = f(R^T u)"""
    rng = np.random.default_rng(1)
    U = rng.normal(size=(40, 3))
    U /= np.linalg.norm(U, axis=1)[:, None]
    A = np.array([d_functions(u) for u in U])
    B = np.array([d_functions(R.T @ u) for u in U])
    Dd = np.linalg.lstsq(A, B, rcond=None)[0]
    D = np.zeros((10, 10))
    D[0, 0] = D[4, 4] = 1.0
    D[1:4, 1:4] = R
    D[5:, 5:] = Dd
    return D


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_table_is_rotation_covariant(P, seed):
    """B(R r) = D(R) B(r) D(R)^T for every pair block (both species orders):
    fails on a wrong sign, a wrong direction cosine or a swapped integral in
    any of the table's cases"""
    rng = np.random.default_rng(10 + seed)
    r = rng.normal(size=3)
    r *= 2.37 / np.linalg.norm(r)
    R = rotation(seed)
    D = orbital_rotation(R)
    assert np.abs(D @ D.T - np.eye(10)).max() < 1e-12
    for i, j in ((1, 2), (2, 1), (2, 2)):
        B0, B1 = block(P, i, j, r), block(P, i, j, R @ r)
        assert np.abs(B1 - D @ B0 @ D.T).max() < 1e-6 * np.abs(B0).max(), (i, j)


def test_table_transposition_and_power_law(P):
    r = np.array([0.9, -1.4, 1.6])
    for i, j in ((1, 2), (2, 2)):
        assert np.abs(block(P, i, j, r) - block(P, j, i, -r).T).max() < 1e-12
    # V(d) = V0 (d0 / d)^n: halving nothing but the length scales every sigma term
    v1 = TB.v_term(P, 2, 1, 1, 2, 2, 2.30)
    v2 = TB.v_term(P, 2, 1, 1, 2, 2, 2.53)
    assert abs(v1 / v2 - (2.53 / 2.30) ** P["Vn"][1, 0, 0, 1, 1]) < 1e-12


def si_bulk():
    a = 5.431
    lat = a / 2 * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    pos = np.array([[0.0, 0.0, 0.0], [a / 4, a / 4, a / 4]])
    return lat, pos, np.array([2, 2])


def test_hamiltonian_hermitian_and_periodic(P):
    lat, pos, sp = si_bulk()
    k = np.array([0.13, -0.21, 0.34])
    H = TB.gen_ham(P, pos, sp, lat, k, 2.6)
    assert np.abs(H - H.conj().T).max() == 0.0
    # a reciprocal lattice vector leaves the spectrum unchanged
    e1 = np.linalg.eigvalsh(H)
    e2 = np.linalg.eigvalsh(TB.gen_ham(P, pos, sp, lat, k + np.array([1.0, 0.0, -1.0]), 2.6))
    assert np.abs(e1 - e2).max() < 1e-10


def test_bulk_silicon_band_edges(P):
    """sp3d5s* Si: the valence top at Gamma is threefold degenerate and the
    conduction minimum lies on Gamma-X near 0.85 X with an indirect gap of
    about 1.1 eV"""
    lat, pos, sp = si_bulk()
    rec = 2 * np.pi * np.linalg.inv(lat).T
    eg = np.linalg.eigvalsh(TB.gen_ham(P, pos, sp, lat, np.zeros(3), 2.6))
    vbm = eg[3]                      # two atoms x four electrons fill bands 0..3
    assert np.abs(eg[1:4] - vbm).max() < 1e-9
    assert np.abs(eg[4:7] - eg[4]).max() < 1e-9     # Gamma15 conduction triplet
    X = np.linalg.solve(rec.T, np.array([0.0, 0.0, 2 * np.pi / 5.431]))   # crystal coordinates of X
    cb = []
    for t in np.linspace(0.6, 1.0, 21):
        cb.append(np.linalg.eigvalsh(TB.gen_ham(P, pos, sp, lat, t * X, 2.6))[4])
    cb = np.array(cb)
    gap = cb.min() - vbm
    tmin = np.linspace(0.6, 1.0, 21)[cb.argmin()]
    assert 1.0 < gap < 1.25, gap
    assert 0.78 <= tmin <= 0.92, tmin
