"""CPU restatement of the tight-binding run's primitive-cell unfolding of
lead modes, find_k_pc_l / find_k_pc_r (test infrastructure only: tests/
import it, the product path never does).

For every propagating mode the tight-binding vel_mat (tb/velocity_mat.f90:
593-605 and its three mirrors) looks for the primitive-cell wave vectors the
mode unfolds to and the group velocity there:

    rot_all              the misorientation rotation of tb/input.f90:173-182
                         (rot_z with the default-real sqrt(2.0) / 2)
    read_pc_fdf          the primitive-cell part of load_input
                         (tb/input.f90:144-234): lattice vectors (rotated),
                         scaled coordinates of the listed atoms, the atoms
                         inside the cell, reciprocal vectors
    TbPrimitiveCell      gen_ham_pc_l (tb/hamiltonian.f90:918-989): the
                         primitive cell's H(k) summed over the images
                         [-2, 2]^3 within r_cutoff, hermitian; gen_hf_pc_l
                         (:1238-1313): the same sum weighted by the bond
                         vector's Cartesian component, i times its
                         antihermitian part; group_velocity_pc_l
                         (:1316-1378): ZHEEV eigenpairs and Re(z^H HF z)
    compare_evs          tb/hamiltonian.f90:1596-1686: a band within 1e-4 of
                         E (the velocity of the last such band)
    find_k_pc            tb/hamiltonian.f90:666-822: the unit cell's own k
                         when it lies on its zone boundary (in_ws == 2; the
                         velocity of a band within 1e-3 of E), else the
                         matching candidates k + n reci_uc in the primitive
                         zone
    mode_pc              the kp / v columns of one mode (find_min_dist of
                         the candidates, the first candidate's velocity)

in_ws, find_min_dist and move_in_ws2 of tb/util.f90 are those of
negf/util.f90 (oracle.kpc.in_ws).

Pinned (tests/test_kpc_tb_oracle.py) by closed forms: a one-orbital chain
in a two-atom primitive cell (bands eps +- 2 t cos(k d), velocities
-+ 2 t d sin(k d)) unfolded from a four-atom lead cell, the bulk Si bands of
the example's parameters against the cubic cell's Hamiltonian of
oracle.tb_ham, and the velocity operator against central differences of
the bands.
"""
import numpy as np

from oracle import tb_ham as TB
from oracle.kpc import in_ws

PI = 3.141592653589793238


def rot_all(theta_deg=0.0):
    """tb/input.f90:173-182: rot_z^T rot_misori rot_z, rot_z holding the
    default-real value sqrt(2.0) / 2.0."""
    c = float(np.float32(np.sqrt(np.float32(2.0))) / np.float32(2.0))
    rz = np.array([[c, -c, 0.0], [c, c, 0.0], [0.0, 0.0, 1.0]])
    th = theta_deg * PI / 180.0
    rm = np.array([[1.0, 0.0, 0.0], [0.0, np.cos(th), -np.sin(th)], [0.0, np.sin(th), np.cos(th)]])
    return rz.T @ rm @ rz


def read_pc_fdf(path, theta_deg=0.0):
    """tb/input.f90:144-234 for one lead: returns latvec_pc (rows, rotated),
    positions_uc (n x 4: Cartesian, species), natoms_uc, positions_pc (the
    atoms with crystal coordinates in (-0.001, 0.999)), recivec_pc."""
    lines = open(path).read().splitlines()
    nat = lscale = None
    latvec, rows = [], []
    for t in (ln.split() for ln in lines):
        if t and t[0] == "NumberOfAtoms":
            nat = int(t[1])
        elif t and t[0] == "LatticeConstant":
            lscale = float(t[1])
    i = 0
    R = rot_all(theta_deg)
    while i < len(lines):
        if "%block LatticeVectors" in lines[i]:
            for k in (1, 2, 3):
                v = np.array([float(x) for x in lines[i + k].split()[:3]]) * lscale
                latvec.append(R @ v)
            i += 3
        elif "%block AtomicScaledCoordinatesAndAtomicSpecies" in lines[i]:
            for k in range(nat):
                t = lines[i + 1 + k].split()
                rows.append([float(x) for x in t[:4]])
            i += nat
        i += 1
    latvec = np.array(latvec)
    pos = np.array(rows)
    pos[:, :3] = pos[:, :3] @ latvec
    inv = np.linalg.inv(latvec)
    keep = [p for p in pos if all(-0.001 < (p[:3] @ inv)[k] < 0.999 for k in range(3))]
    return dict(latvec_pc=latvec, positions_uc=pos, natoms_uc=nat, positions_pc=np.array(keep),
                recivec_pc=2 * PI * inv.T)


class TbPrimitiveCell:
    """The primitive cell's H(k) and velocity operators (k Cartesian, 1/A)."""

    def __init__(self, P, positions_pc, latvec_pc, r_cutoff):
        self.P = P
        self.pos = np.asarray(positions_pc, float)[:, :3]
        self.sp = [int(round(x)) for x in np.asarray(positions_pc)[:, 3]]
        self.lat = np.asarray(latvec_pc, float)
        self.rc = r_cutoff
        self.orbs = [(a, o) for a in range(len(self.pos)) for o in P["orbitals"][self.sp[a] - 1]]
        self.n = len(self.orbs)
        # (row, col, V, r) of every element the sums run over
        self.terms = []
        start = np.concatenate([[0], np.cumsum([P["orb_num"][s - 1] for s in self.sp])]).astype(int)
        for ia in range(len(self.pos)):
            for ja in range(len(self.pos)):
                ti, tj = self.sp[ia], self.sp[ja]
                for i1 in range(-2, 3):
                    for i2 in range(-2, 3):
                        for i3 in range(-2, 3):
                            r = self.pos[ja] - self.pos[ia] + np.array([i1, i2, i3], float) @ self.lat
                            if not np.sqrt(r @ r) < r_cutoff:
                                continue
                            for io, o1 in enumerate(P["orbitals"][ti - 1]):
                                for jo, o2 in enumerate(P["orbitals"][tj - 1]):
                                    v = TB.onsite(P, ti, tj, o1, o2) if ia == ja else TB.v_sk(P, ti, tj, o1, o2, r)
                                    self.terms.append((start[ia] + io, start[ja] + jo, v, r.copy()))

    def ham(self, k):
        H = np.zeros((self.n, self.n), complex)
        for r, c, v, d in self.terms:
            H[r, c] += np.exp(1j * (d @ k)) * v
        return 0.5 * (H + H.conj().T)

    def hf(self, k, c3):
        H = np.zeros((self.n, self.n), complex)
        for r, c, v, d in self.terms:
            H[r, c] += np.exp(1j * (d @ k)) * v * d[c3]
        return 1j * 0.5 * (H - H.conj().T)

    def group_velocity(self, k):
        """eigs (ascending), eigenvectors, vels (n x 3)."""
        eigs, evs = np.linalg.eigh(self.ham(k))
        vels = np.zeros((self.n, 3))
        for c3 in range(3):
            V = evs.conj().T @ self.hf(k, c3) @ evs
            vels[:, c3] = np.real(np.diag(V))
        return eigs, evs, vels


def compare_evs(E, pc, kpc, tol=1.0e-4):
    """(found, vxyz): a band within tol of E, the velocity of the last such."""
    eigs, _, vels = pc.group_velocity(kpc)
    found, v = 0, np.zeros(3)
    for i in range(len(eigs)):
        if abs(E - eigs[i]) < tol:
            v = vels[i].copy()
            found = 1
    return found, v


def find_k_pc(E, k_uc, reci_uc, recivec_pc, nbx, nby, pc):
    """tb/hamiltonian.f90:666-822: lists of candidate primitive-cell k and
    velocities."""
    l = max(nbx, nby)
    k_uc = np.asarray(k_uc, float)
    if in_ws(reci_uc, k_uc) == 2:
        _, v = compare_evs(E, pc, k_uc, tol=1.0e-3)
        return [k_uc.copy()], [v]
    kpcs, vs = [], []
    for i1 in range(-l, l + 1):
        for i2 in range(-l, l + 1):
            for i3 in range(-l, l + 1):
                kpc = k_uc + np.array([i1, i2, i3], float) @ reci_uc
                if in_ws(recivec_pc, kpc) != 0:
                    found, v = compare_evs(E, pc, kpc)
                    if found > 0:
                        kpcs.append(kpc)
                        vs.append(v)
    return kpcs, vs


def mode_pc(E, k_xyz, reci_uc, recivec_pc, nbx, nby, pc):
    """(kp, v) of one propagating mode: the candidate of smallest |k| (first
    on ties) and the first candidate's velocity; zeros without a candidate."""
    kpcs, vs = find_k_pc(E, k_xyz, reci_uc, recivec_pc, nbx, nby, pc)
    if not kpcs:
        return np.zeros(3), np.zeros(3)
    norms = [np.sqrt(k @ k) for k in kpcs]
    return kpcs[int(np.argmin(norms))], vs[0]
