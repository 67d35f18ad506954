"""Pins of the tight-binding primitive-cell unfolding oracle (oracle/kpc_tb.py)
and of the host table it shares with the GPU path (tb.pc_table), on the CPU.

* A one-orbital chain (atoms d apart) in a two-atom primitive cell (bands
  eps +- 2 t cos(k d), velocities -+ 2 t d sin(k d)) seen from a four-atom
  lead cell: every mode unfolds to the primitive k with the band's energy
  and its group velocity, on the lead cell's zone boundary to its own k.
* Bulk Si with the example's parameters: the two-atom primitive cell's bands
  at k are bands of the eight-atom cubic cell's H(k) (oracle.tb_ham.gen_ham),
  and its velocity operator gives the derivative of the bands (central
  differences of the eigenvalues).
* tb.pc_table summed on the host equals the oracle's H(k) and HF(k).
* read_pc_fdf: the rotation of theta = 0 is the identity up to the
  single-precision rounding of sqrt(2.0) / 2, the atoms inside the cell are
  kept."""
import os

import numpy as np
import pytest

from oracle import kpc_tb as K
from oracle import tb_ham as TB
from phinfortran_amd import tb

HERE = os.path.dirname(os.path.abspath(__file__))
PARAM = os.path.join(HERE, "golden", "tb", "param")


def chain_param(eps, t):
    V = np.zeros((1, 1, 4, 4, 4))
    V[0, 0, 0, 0, 0] = eps          # onsite s
    V[0, 0, 0, 0, 1] = t            # s s sigma
    return dict(V=V, Vn=np.zeros_like(V), r_d=np.array([[1.0e-10]]), orbitals=[["s"]], orb_num=[1])


def chain_pc(d, eps, t, box=10.0):
    """a chain of equal atoms d apart described by a two-atom primitive cell
    (length 2 d): with the bond vectors' phases H(k) = [[eps, 2 t cos(k d)],
    [2 t cos(k d), eps]], bands eps +- 2 t cos(k d), velocities
    -+ 2 t d sin(k d).  (A one-atom cell would not do: gen_ham_pc_l gives
    every image of the atom itself the onsite block.)"""
    P = chain_param(eps, t)
    lat = np.diag([box, box, 2 * d])
    pos = np.array([[0.0, 0.0, 0.0, 1.0], [0.0, 0.0, d, 1.0]])
    return P, K.TbPrimitiveCell(P, pos, lat, 2.6), lat


def lattices(d, lat):
    reci_pc = 2 * np.pi * np.linalg.inv(lat).T
    uc = lat.copy()
    uc[2, 2] = 4 * d                                        # the lead's cell: two primitive cells
    return reci_pc, 2 * np.pi * np.linalg.inv(uc).T


@pytest.mark.parametrize("k0d", [0.2, 0.6, 1.0, 1.4])
def test_chain_unfolds_to_the_band(k0d):
    d, eps, t = 2.0, 0.1, -0.6
    P, pc, lat = chain_pc(d, eps, t)
    reci_pc, reci_uc = lattices(d, lat)
    E = eps + 2 * t * np.cos(k0d)
    assert abs(np.linalg.eigvalsh(pc.ham(np.array([0, 0, k0d / d]))) - E).min() < 1e-12
    for sgn in (1.0, -1.0):
        kz = sgn * k0d / d
        # the lead cell's k: folded into (-pi / 4d, pi / 4d]
        kf = (kz + np.pi / (4 * d)) % (np.pi / (2 * d)) - np.pi / (4 * d)
        kp, v = K.mode_pc(E, np.array([0.0, 0.0, kf]), reci_uc, reci_pc, 1, 1, pc)
        assert np.allclose(kp, [0.0, 0.0, kz], atol=1e-12), (k0d, sgn, kp, kz)
        assert v[2] == pytest.approx(-2 * t * d * np.sin(kz * d), rel=1e-10)
        assert abs(v[0]) < 1e-12 and abs(v[1]) < 1e-12


def test_chain_zone_boundary_is_direct():
    d, eps, t = 2.0, 0.1, -0.6
    P, pc, lat = chain_pc(d, eps, t)
    reci_pc, reci_uc = lattices(d, lat)
    k = np.array([0.0, 0.0, np.pi / (4 * d)])
    E = eps + 2 * t * np.cos(np.pi / 4)
    kpcs, vs = K.find_k_pc(E, k, reci_uc, reci_pc, 1, 1, pc)
    assert len(kpcs) == 1 and np.allclose(kpcs[0], k)
    assert vs[0][2] == pytest.approx(-2 * t * d * np.sin(np.pi / 4), rel=1e-10)
    # off the band by more than 1e-3: the candidate stays, its velocity is zero
    kpcs, vs = K.find_k_pc(E + 2e-3, k, reci_uc, reci_pc, 1, 1, pc)
    assert len(kpcs) == 1 and not vs[0].any()


def si_pc(a=5.43):
    P = TB.read_param(PARAM)
    fcc = np.array([[0.0, a / 2, a / 2], [a / 2, 0.0, a / 2], [a / 2, a / 2, 0.0]])
    pos = np.array([[0.0, 0.0, 0.0, 1.0], [a / 4, a / 4, a / 4, 1.0]])
    return P, K.TbPrimitiveCell(P, pos, fcc, 2.6), fcc


def test_si_primitive_bands_fold_into_the_cubic_cell():
    a = 5.43
    P, pc, fcc = si_pc(a)
    cub = np.diag([a, a, a])
    base = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]]) * a
    pos = np.concatenate([base, base + a / 4])
    sp = np.ones(8, int)
    rng = np.random.default_rng(3)
    for _ in range(3):
        k = rng.uniform(-0.5, 0.5, 3) * 2 * np.pi / a
        e_pc = np.linalg.eigvalsh(pc.ham(k))
        e_c = np.linalg.eigvalsh(TB.gen_ham(P, pos, sp, cub, k @ cub.T / (2 * np.pi), 2.6))
        for e in e_pc:
            assert np.abs(e_c - e).min() < 1e-9, e


def test_si_velocity_operator_is_the_band_derivative():
    P, pc, fcc = si_pc()
    k = np.array([0.13, -0.21, 0.37])
    eigs, _, vels = pc.group_velocity(k)
    h = 1e-6
    for c3 in range(3):
        dk = np.zeros(3)
        dk[c3] = h
        d = (np.linalg.eigvalsh(pc.ham(k + dk)) - np.linalg.eigvalsh(pc.ham(k - dk))) / (2 * h)
        nondeg = np.ones(len(eigs), bool)
        nondeg[1:] &= np.diff(eigs) > 1e-4
        nondeg[:-1] &= np.diff(eigs) > 1e-4
        assert np.abs(vels[nondeg, c3] - d[nondeg]).max() < 1e-6


def test_host_table_equals_oracle():
    P = tb.read_param(PARAM)
    Po, pc, fcc = si_pc()
    el_off, coef, dr, n = tb.pc_table(P, np.array([[0.0, 0.0, 0.0], [5.43 / 4] * 3]), [1, 1], fcc, 2.6)
    assert n == pc.n and el_off[-1] == len(coef) == len(dr)
    k = np.array([0.3, 0.1, -0.2])
    H = np.zeros((n, n), complex)
    F = np.zeros((3, n, n), complex)
    for e in range(n * n):
        r, c = e % n, e // n
        for j in range(el_off[e], el_off[e + 1]):
            ph = np.exp(1j * (dr[j] @ k))
            H[r, c] += coef[j] * ph
            F[:, r, c] += coef[j] * dr[j] * ph
    assert np.abs(0.5 * (H + H.conj().T) - pc.ham(k)).max() < 1e-12
    for c3 in range(3):
        assert np.abs(1j * 0.5 * (F[c3] - F[c3].conj().T) - pc.hf(k, c3)).max() < 1e-12


def test_read_pc_fdf(tmp_path):
    a = 5.5
    f = tmp_path / "pc.fdf"
    f.write_text("NumberOfAtoms 3\nNumberOfSpecies 1\nLatticeConstant 1.0 Ang\n%block LatticeVectors\n"
                 f"0.0 {a / 2} {a / 2}\n{a / 2} 0.0 {a / 2}\n{a / 2} {a / 2} 0.0\n%endblock LatticeVectors\n"
                 "%block AtomicScaledCoordinatesAndAtomicSpecies\n0.0 0.0 0.0 2\n0.25 0.25 0.25 2\n"
                 "1.0 0.25 0.25 2\n%endblock AtomicScaledCoordinatesAndAtomicSpecies\n")
    pc = tb.read_pc_fdf(str(f))
    o = K.read_pc_fdf(str(f))
    c = float(np.float32(np.sqrt(np.float32(2.0))) / np.float32(2.0))
    assert np.allclose(np.diag(tb.misorientation_rotation(0.0)), [2 * c * c, 2 * c * c, 1.0], rtol=0, atol=1e-16)
    assert abs(2 * c * c - 1.0) > 1e-9                       # single-precision sqrt(2.0) / 2
    assert np.allclose(pc["latvec_pc"], o["latvec_pc"], rtol=0, atol=1e-15)
    assert np.allclose(pc["latvec_pc"], [[0, a / 2, a / 2], [a / 2, 0, a / 2], [a / 2, a / 2, 0]], atol=1e-6)
    assert pc["natoms_uc"] == 3 and len(pc["positions_pc"]) == 2 and list(pc["species_pc"]) == [2, 2]
    assert np.allclose(pc["positions_pc"], o["positions_pc"][:, :3], atol=1e-15)
    assert np.allclose(pc["recivec_pc"] @ pc["latvec_pc"].T, 2 * np.pi * np.eye(3), atol=1e-12)
