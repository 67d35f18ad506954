"""Host side of the tight-binding H(k) builder (phinfortran_amd/tb.py)
against the oracle (oracle/tb_ham.py): the parameter and structure readers,
every Slater-Koster block for both species orders, and the phase-sum table,
contracted here on the CPU, against gen_ham (tb/hamiltonian.f90:444-519)."""
import os

import numpy as np
import pytest

from oracle import tb_ham as TB
from phinfortran_amd import tb
from tests import tb_structures as S

HERE = os.path.dirname(os.path.abspath(__file__))
PARAM = os.path.join(HERE, "golden", "tb", "param")


@pytest.fixture(scope="module")
def P():
    return tb.read_param(PARAM)


def test_read_param_matches_oracle(P):
    Q = TB.read_param(PARAM)
    for k in ("V", "Vn", "r_d"):
        assert np.array_equal(P[k], Q[k]), k
    assert P["orbitals"] == Q["orbitals"] and P["orb_num"] == Q["orb_num"]


def test_read_fdf_matches_oracle(tmp_path):
    lat, pos, sp = S.si_ge_superlattice(5.5, 1, 1, 2)
    f = str(tmp_path / "input.fdf")
    S.write_fdf(f, lat, pos, sp, scale=1.0)
    a, b = tb.read_fdf(f), TB.read_fdf(f)
    assert np.array_equal(a["latvec"], b["latvec"]) and np.array_equal(a["positions"], b["positions"])
    assert np.array_equal(a["species"], b["species"])


@pytest.mark.parametrize("si,sj", [(1, 1), (1, 2), (2, 1), (2, 2)])
def test_sk_blocks_match_oracle(P, si, sj):
    rng = np.random.default_rng(10 * si + sj)
    r = rng.normal(size=(12, 3))
    r *= (rng.uniform(2.2, 2.7, size=12) / np.linalg.norm(r, axis=1))[:, None]
    r[:3] = [[1.4, 1.4, 1.4], [0.0, 0.0, 2.4], [-1.7, 0.0, 1.7]]       # axis and high-symmetry bonds
    got = tb.sk_blocks(P, si, sj, r)
    names = P["orbitals"]
    for q in range(len(r)):
        want = np.array([[TB.v_sk(P, si, sj, a, b, r[q]) for b in names[sj - 1]] for a in names[si - 1]])
        assert np.abs(got[q] - want).max() <= 1e-14 * max(1.0, np.abs(want).max()), (q, r[q])


def contract(A, k):
    """the table's phase sum and (H + H^+)/2, on the CPU"""
    kc = A.kcart(k)[0]
    H = np.zeros((A.norb, A.norb), complex)
    g = A.seg_geom.reshape(-1, 4)
    for s in range(A.nseg):
        r0, c0, ni, nj = g[s]
        for q in range(A.seg_off[s], A.seg_off[s + 1]):
            H[r0:r0 + ni, c0:c0 + nj] += A.coef[q, :ni * nj].reshape(ni, nj) * np.exp(1j * (A.dr[q] @ kc))
    return 0.5 * (H + H.conj().T)


@pytest.mark.parametrize("name", ["bulk", "superlattice", "chain"])
def test_table_matches_gen_ham(P, name):
    lat, pos, sp = {"bulk": S.si_bulk, "superlattice": lambda: S.si_ge_superlattice(5.5, 1, 1, 4),
                    "chain": S.short_chain}[name]()
    A = tb.TbAssembler(P, lat, pos, sp, 2.6)
    rng = np.random.default_rng(3)
    for k in (np.zeros(3), rng.uniform(-0.5, 0.5, 3), np.array([0.5, 0.0, 0.25])):
        want = TB.gen_ham(P, pos, sp, lat, k, 2.6)
        got = contract(A, k[None])
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max(), (name, k)
    if name == "chain":      # the atom's own images carry the onsite energies, as in the reference
        assert A.nent == 3 and A.nseg == 1
