"""Structures for the tight-binding H(k) tests: bulk Si (two-atom cell), the
Si/Ge superlattice of the reference's example (examples/tb/gen_struct.py: a
tetragonal four-atom diamond cell repeated nx x ny x nz, the first half of
the atoms species 2, the rest species 1), and a one-atom chain short enough
that the atom's own images lie inside the cutoff."""
import numpy as np


def si_bulk(a=5.431):
    lat = a / 2 * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    pos = np.array([[0.0, 0.0, 0.0], [a / 4, a / 4, a / 4]])
    return lat, pos, np.array([2, 2])


def si_ge_superlattice(a=5.5, nx=2, ny=2, nz=8):
    cell = np.array([[a / 2, -a / 2, 0.0], [a / 2, a / 2, 0.0], [0.0, 0.0, a]])     # four-atom cell
    basis = np.array([[0.0, 0.0, 0.0], [a / 4, a / 4, a / 4], [a / 2, 0.0, a / 2], [a / 4, -a / 4, 3 * a / 4]])
    pos, sp = [], []
    nat = 4 * nx * ny * nz
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                for b in basis:
                    pos.append(np.array([i, j, k], float) @ cell + b)
                    sp.append(2 if len(sp) < nat // 2 else 1)
    lat = np.diag([nx, ny, nz]).astype(float) @ cell
    return lat, np.array(pos), np.array(sp)


def short_chain(c=2.3):
    lat = np.array([[c, 0.0, 0.0], [0.0, 9.0, 0.0], [0.0, 0.0, 9.0]])
    return lat, np.zeros((1, 3)), np.array([1])


def write_fdf(path, lat, pos, sp, scale=1.0):
    with open(path, "w") as f:
        f.write(f"NumberOfAtoms {len(pos)}\nNumberOfSpecies {int(sp.max())}\nLatticeConstant {scale} Ang\n")
        f.write("%block LatticeVectors\n")
        for r in lat / scale:
            f.write(" ".join(f"{x:.13f}" for x in r) + "\n")
        f.write("%endblock LatticeVectors\nAtomicCoordinatesFormat  Ang\n")
        f.write("%block AtomicCoordinatesAndAtomicSpecies\n")
        for p, s in zip(pos / scale, sp):
            f.write(" ".join(f"{x:.13f}" for x in p) + f" {int(s)}\n")
        f.write("%endblock AtomicCoordinatesAndAtomicSpecies\n")
