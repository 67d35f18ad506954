"""CPU restatement of the primitive-cell unfolding of lead modes, find_k_pc
(test infrastructure only: tests/ import it, the product path never does).

For every propagating mode, vel_mat_tdb (negf/velocity_mat.f90:2073-2090 and
mirrors) looks for the primitive-cell wave vectors the mode unfolds to and
the group velocity there:

    read_pc_fdf          read_pc_qe (negf/input.f90:520-747): the primitive
                         cell file (lattice rotated by pi/4 about z, atoms
                         inside the cell, reciprocal vectors)
    PcDynamics           gen_dyn_pc_qe_k (negf/dyn.f90:2537-2644): the
                         dynamical matrix of the primitive cell's atoms from
                         the QE force constants, Wigner-Seitz weighted,
                         columns folded by atom index mod natoms_pc, then
                         hermitian (negf/util.f90:940-949);
                         group_velocity_pc_qe (dyn.f90:599-659): ZHEEV
                         eigenpairs and central differences (1e-5) of the
                         matrix in each Cartesian direction
    in_ws                negf/util.f90:589-638
    compare_evs          dyn.f90:662-768: a band within 1e-2 of sqrt(E)
                         (the velocity of the last such band)
    find_k_pc            dyn.f90:770-889
    mode_pc              the kp / v columns of one mode (velocity_mat.f90:
                         2088-2089: find_min_dist of the candidates, the
                         first candidate's velocity)

Pinned by the primitive-cell k and velocity columns of the si-si example's
committed per-mode tables (tests/test_kpc_oracle.py).
"""
import numpy as np

PI = 3.141592653589793238
F32 = lambda x: float(np.float32(x))


def rotate_z(v, ang):
    """negf/util.f90:901-919 (rows are vectors)."""
    out = np.array(v, float, copy=True)
    out[:, 0] = v[:, 0] * np.cos(ang) - v[:, 1] * np.sin(ang)
    out[:, 1] = v[:, 0] * np.sin(ang) + v[:, 1] * np.cos(ang)
    return out


def read_pc_fdf(path, species_mass):
    """read_pc_qe (negf/input.f90:580-747) for one lead: returns latvec_pc
    (rotated), positions_uc (Cartesian), mass_uc, natoms_pc, recivec_pc."""
    lines = open(path).read().splitlines()
    nat = lscale = None
    latvec, coords, species = [], [], []
    i = 0
    while i < len(lines):
        t = lines[i].split()
        if t and t[0] == "NumberOfAtoms":
            nat = int(t[1])
        elif t and t[0] == "LatticeConstant":
            lscale = float(t[1])
        elif "%block LatticeVectors" in lines[i]:
            latvec = [[float(x) for x in lines[i + k].split()[:3]] for k in (1, 2, 3)]
            i += 3
        elif "%block AtomicScaledCoordinatesAndAtomicSpecies" in lines[i]:
            for k in range(nat):
                t = lines[i + 1 + k].split()
                coords.append([float(x) for x in t[:3]])
                species.append(int(t[3]))
            i += nat
        i += 1
    latvec = rotate_z(np.array(latvec) * lscale, PI / 4.0)
    pos = np.array(coords) @ latvec
    mass = np.array([species_mass[s] for s in species])
    inv = np.linalg.inv(latvec)
    npc = 0
    for p in pos:
        c = p @ inv
        if all(-0.001 < c[k] < 0.999 for k in range(3)):
            npc += 1
    recivec_pc = 2 * PI * inv.T
    return dict(latvec_pc=latvec, positions_uc=pos, mass_uc=mass, natoms_pc=npc, recivec_pc=recivec_pc)


def idnint(x):
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


class PcDynamics:
    """gen_dyn_pc_qe_k's k-independent part, tabulated once: for every
    (primitive atom i, QE atom j, image) inside the Wigner-Seitz cell of the
    force-constant supercell, w Phi_ij / sqrt(m_i m_j), dr and the target
    block (i, j mod natoms_pc)."""

    def __init__(self, qe, npc, mass_uc):
        pos, cell, fc = qe["pos"], qe["cell"], qe["fc"]
        nx, ny, nz = qe["nx"], qe["ny"], qe["nz"]
        rws, rd = qe["rws"], qe["rd"]
        ivcell = np.linalg.inv(cell)
        self.n = 3 * npc
        rows, cols, coef, drs = [], [], [], []
        abc = np.array([(a, b, c) for a in range(-2 * nx, 2 * nx + 1) for b in range(-2 * ny, 2 * ny + 1)
                        for c in range(-2 * nz, 2 * nz + 1)], float)
        for i in range(npc):
            for j in range(len(pos)):
                dr = -abc @ cell + pos[j] - pos[i]                       # (images, 3)
                df = dr @ rws.T - rd[None, :]                            # (images, 124)
                ok = ~(df > 1.0e-5).any(axis=1)
                nreq = 1 + (np.abs(df) <= 1.0e-5).sum(axis=1)
                ixyz = (dr - (pos[j] - pos[i])) @ ivcell
                aa = -idnint(ixyz)
                for q in np.nonzero(ok)[0]:
                    t1, t2, t3 = (int(aa[q, 0]) % nx, int(aa[q, 1]) % ny, int(aa[q, 2]) % nz)
                    w = 1.0 / nreq[q]
                    rows.append(i)
                    cols.append(j % npc)
                    coef.append(w * fc[:, :, i, j, t1, t2, t3] / np.sqrt(mass_uc[i] * mass_uc[j]))
                    drs.append(dr[q])
        self.rows, self.cols = np.array(rows), np.array(cols)
        self.coef, self.dr = np.array(coef), np.array(drs)

    def dyn(self, k):
        D = np.zeros((self.n, self.n), complex)
        ph = np.exp(1j * (self.dr @ np.asarray(k, float)))
        for r, c, f, e in zip(self.rows, self.cols, self.coef, ph):
            D[3 * r:3 * r + 3, 3 * c:3 * c + 3] += f * e
        return 0.5 * (D + D.conj().T)

    def group_velocity(self, k, devi=1.0e-5):
        """eigs (ascending), eigenvectors, vels (n x 3)."""
        eigs, evs = np.linalg.eigh(self.dyn(k))
        vels = np.zeros((self.n, 3))
        for i in range(3):
            km = np.array(k, float)
            kp = np.array(k, float)
            km[i] -= devi
            kp[i] += devi
            dH = (self.dyn(kp) - self.dyn(km)) / (2.0 * devi)
            V = evs.conj().T @ dH @ evs
            vels[:, i] = np.real(np.diag(V)) / (2 * np.sqrt(np.abs(eigs)))
        return eigs, evs, vels


def in_ws(a, k):
    """negf/util.f90:589-638: 2 when the two nearest lattice points of k are
    equally near (within 1e-5), 1 when the origin is nearest, else 0."""
    n = np.array([(i1, i2, i3) for i1 in range(-2, 3) for i2 in range(-2, 3) for i3 in range(-2, 3)], float)
    d = k[None, :] - n @ a
    dist = (d * d).sum(axis=1)
    dmin = dist.min()
    imin = int(np.argmin(dist))
    d2 = np.delete(dist, imin).min()
    if abs(d2 - dmin) < 1.0e-5:
        return 2
    if abs(dist[62] - dmin) < 1.0e-5:
        return 1
    return 0


def compare_evs(E, dyn, kpc):
    """(found, vxyz) of dyn.f90:662-768 (the overlap weight it also computes is
    not used by its caller)."""
    eigs, _, vels = dyn.group_velocity(kpc)
    found, v = 0, np.zeros(3)
    for i in range(len(eigs)):
        if abs(np.sqrt(E) - np.sqrt(abs(eigs[i]))) < 1.0e-2:
            v = vels[i].copy()
            found = 1
    return found, v


def find_k_pc(E, k_uc, reci_uc, recivec_pc, nbx, nby, dyn):
    """dyn.f90:770-889: lists of candidate primitive-cell k and velocities."""
    l = max(nbx, nby)
    kpcs, vs = [], []
    if in_ws(reci_uc, k_uc) == 2:
        kpcs.append(np.array(k_uc, float))
        found, v = compare_evs(E, dyn, k_uc)
        vs.append(v)
        return kpcs, vs
    for i1 in range(-l, l + 1):
        for i2 in range(-l, l + 1):
            for i3 in range(-l, l + 1):
                kpc = k_uc + np.array([i1, i2, i3], float) @ reci_uc
                if in_ws(recivec_pc, kpc) != 0:
                    found, v = compare_evs(E, dyn, kpc)
                    if found > 0:
                        kpcs.append(kpc)
                        vs.append(v)
    return kpcs, vs


def mode_pc(E, k_xyz, reci_uc, recivec_pc, nbx, nby, dyn):
    """(kp, v) of one propagating mode (velocity_mat.f90:2088-2089): the
    candidate of smallest |k| (first on ties) and the first candidate's
    velocity; zeros when there is no candidate (the reference reads element
    0 of an empty list there; its committed tables hold zeros)."""
    kpcs, vs = find_k_pc(E, k_xyz, reci_uc, recivec_pc, nbx, nby, dyn)
    if not kpcs:
        return np.zeros(3), np.zeros(3)
    norms = [np.sqrt(k @ k) for k in kpcs]
    return kpcs[int(np.argmin(norms))], vs[0]
