"""The reference's si-si example as a golden fixture (test infrastructure).

The only golden data phinfortran ships for this path is the committed output
of examples/negf/si-si (T(omega, k) tables transl_tot.dat / transr_tot.dat,
61 frequencies x 49 transverse k).  Reproducing it needs the H(k) blocks that
the reference builds from the Quantum ESPRESSO force constants in its tdb
branch; this module restates that assembly:

    read_qe_fc          negf/fc.f90:103-325   (IFCs -> THz^2 * amu, cutoff, ASR,
                                              Wigner-Seitz vectors, supercell)
    gen_dyn_qe_tdb_k    negf/dyn.f90:1482-2080 (WS-weighted Fourier sums into
                                              Haml, Hamlv, Hamr, Hamrv, Hamd,
                                              Hamdv, Hamld, Hamrd)
    load_configure      negf/config.f90:156-211 (k grid, Egrid = omega^2)

The inputs are the example's own files (Si444.fc, mass_profile.dat, config);
`make_fixture()` copies them (data, not code) into tests/golden/si_si/ so the
fixture travels without the reference.  Constants written as single-precision
literals in the reference (masses, Ry/bohr^2 conversion, bohr2ang) are rounded
to float32 exactly as the Fortran compiler does.
"""
from __future__ import annotations

import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "si_si")
REF_EXAMPLE = "/root/reference/examples/negf/si-si"

# negf/util.f90:5-8
PI = 3.141592653589793238
ELEVOLT = 1.609e-19
MASS_PROTON = 1.6726219e-27
F32 = lambda x: float(np.float32(x))        # a REAL(4) literal promoted to REAL(8)


# ---------------------------------------------------------------- inputs
def make_fixture():
    """Copy the example's input and output data files into tests/golden/si_si."""
    os.makedirs(FIXTURE, exist_ok=True)
    for f in ("Si444.fc", "mass_profile.dat", "config"):
        shutil.copy(os.path.join(REF_EXAMPLE, f), os.path.join(FIXTURE, f))
    for f in ("transl_tot.dat", "transr_tot.dat"):
        shutil.copy(os.path.join(REF_EXAMPLE, "output", f), os.path.join(FIXTURE, f))
    # per-mode transmission of left-incident modes at a few k-points
    for ik in MODE_KPOINTS:
        f = "trl_specular_param_%06d.dat" % (ik + 1)
        shutil.copy(os.path.join(REF_EXAMPLE, "output", f), os.path.join(FIXTURE, f))


# k-points (0-based) whose per-mode tables are kept as fixtures
MODE_KPOINTS = (0, 10, 24, 40)


def read_config(path=None):
    """The namelist values the path uses (negf/config.f90)."""
    path = path or os.path.join(FIXTURE, "config")
    vals = {}
    for line in open(path):
        line = line.split("!")[0].strip()
        if "=" in line:
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip().strip('"').strip("'")
    mesh = [int(x) for x in vals["k_trans_mesh(:)"].split()]
    g = lambda k: float(vals[k].replace("d", "e").replace("D", "e"))
    return dict(nk1=mesh[0], nk2=mesh[1], ne=int(vals["ne"]), emin=g("emin"), emax=g("emax"),
                eta=g("eta"), eta0=g("eta0"), convergence=g("convergence"), r_cutoff=g("r_cutoff"),
                nz=int(vals["nz"]), n_bloch_x=int(vals["n_bloch_x"]), n_bloch_y=int(vals["n_bloch_y"]),
                period=int(vals["period_left"]), buffer=int(vals["buffer_left"]))


def read_qe_fc(cfg, path=None):
    """negf/fc.f90:103-325 for the tdb branch."""
    path = path or os.path.join(FIXTURE, "Si444.fc")
    lines = open(path).read().splitlines()
    hdr = lines[0].split()
    ntype, nat = int(hdr[0]), int(hdr[1])
    celldm1 = float(hdr[3])
    bohr2ang = F32(0.529177)
    cell = np.array([[float(x) for x in lines[1 + i].split()] for i in range(3)])
    pos = []
    li = 4 + ntype
    for a in range(nat):
        t = lines[li + a].split()
        pos.append([float(x) for x in t[2:5]])
    pos = np.array(pos) * celldm1 * bohr2ang
    cell = cell * celldm1 * bohr2ang
    li += nat
    li += 1                                    # the 'F' label line
    nx, ny, nz = (int(x) for x in lines[li].split())
    li += 1
    conv = ELEVOLT / 1.0e-20 / MASS_PROTON / (1.0e12 * 2.0 * PI) ** 2 * \
        F32(np.float32(13.605662285137) / np.float32(np.float32(0.529177249) ** 2))
    fc = np.zeros((3, 3, nat, nat, nx, ny, nz))
    for _ in range((3 * nat) ** 2):
        al, be, ia, ja = (int(x) for x in lines[li].split())
        li += 1
        for _ in range(nx * ny * nz):
            t = lines[li].split()
            li += 1
            fc[al - 1, be - 1, ia - 1, ja - 1, int(t[0]) - 1, int(t[1]) - 1, int(t[2]) - 1] = float(t[3]) * conv
    cell_sc = np.array([nx * cell[0], ny * cell[1], nz * cell[2]])
    # cutoff radius (update_nbr = 1 in the example): zero pairs beyond r_cutoff
    m = np.array([(m1, m2, m3) for m1 in range(-2, 3) for m2 in range(-2, 3) for m3 in range(-2, 3)], float)
    for ia in range(nat):
        for ja in range(nat):
            for j in range(nx):
                for k in range(ny):
                    for l in range(nz):
                        dr = pos[ja] - pos[ia] - np.array([j, k, l], float) @ cell + m @ cell_sc
                        rmin = np.sqrt((dr * dr).sum(axis=1)).min()
                        if rmin > cfg["r_cutoff"]:
                            fc[:, :, ia, ja, j, k, l] = 0.0
    # acoustic sum rule on the on-site blocks (fc.f90:271-287)
    for i in range(3):
        for j in range(i, 3):
            for a in range(nat):
                fc[i, j, a, a, 0, 0, 0] = -fc[i, j, a, :, :, :, :].sum() + fc[i, j, a, a, 0, 0, 0]
    for i in range(3):
        for j in range(i + 1):
            for a in range(nat):
                fc[i, j, a, a, 0, 0, 0] = fc[j, i, a, a, 0, 0, 0]
    # Wigner-Seitz vectors of the IFC supercell (fc.f90:290-306)
    rws = []
    for m1 in range(-2, 3):
        for m2 in range(-2, 3):
            for m3 in range(-2, 3):
                if m1 == 0 and m2 == 0 and m3 == 0:
                    continue
                rws.append(cell_sc[0] * m1 + cell_sc[1] * m2 + cell_sc[2] * m3)
    rws = np.array(rws)
    rd = 0.5 * (rws * rws).sum(axis=1)
    # device supercell: n_bloch_x x n_bloch_y x nz copies, z outermost (fc.f90:315-323)
    nbx, nby, nzd = cfg["n_bloch_x"], cfg["n_bloch_y"], cfg["nz"]
    cell_all = np.array([nbx * cell[0], nby * cell[1], nzd * cell[2]])
    pos_all, idx_all = [], []
    for m1 in range(nzd):
        for m2 in range(nby):
            for m3 in range(nbx):
                for i in range(nat):
                    pos_all.append(pos[i] + np.array([m3, m2, m1], float) @ cell)
                    idx_all.append(i)
    mass_id = [int(x) for x in open(os.path.join(os.path.dirname(path), "mass_profile.dat")).read().split()[1:]]
    # species 2 -> atomic number 14 -> element_mass = 28.0855 (a REAL(4) literal, input.f90:419-420)
    masses = {1: F32(72.64), 2: F32(28.0855)}
    mass_all = np.array([masses[m] for m in mass_id[:len(pos_all)]])
    return dict(cell=cell, pos=pos, nat=nat, nx=nx, ny=ny, nz=nz, fc=fc, rws=rws, rd=rd, cell_all=cell_all,
                pos_all=np.array(pos_all), idx_all=np.array(idx_all), mass_all=mass_all)


# --------------------------------------------------------------- H(k)
class TdbAssembler:
    """gen_dyn_qe_tdb_k (negf/dyn.f90:1482-2080): every block is the same
    Wigner-Seitz-weighted Fourier sum over atom pairs (i, j) and in-plane
    supercell images (a, b); the k-independent part (weights, IFCs, mass
    factors, dr) is tabulated once per block, the phase is applied per k."""

    def __init__(self, qe, cfg):
        self.qe, self.cfg = qe, cfg
        nat = qe["nat"]
        self.natm_layer = nat * cfg["n_bloch_x"] * cfg["n_bloch_y"]
        self.nl = 3 * self.natm_layer
        self.ivcell = np.linalg.inv(qe["cell"])
        self.cell_reci = 2 * PI * np.linalg.inv(qe["cell"]).T
        self.nscx = -(-qe["nx"] // cfg["n_bloch_x"])
        self.nscy = -(-qe["ny"] // cfg["n_bloch_y"])
        self.nlayer = cfg["nz"] - 6
        per, buf = cfg["period"], cfg["buffer"]
        natm_all = len(qe["pos_all"])
        rng = lambda a, b: np.arange(a, b)       # 0-based atom index ranges
        dev0 = buf + 2 * per
        nla = self.natm_layer
        R = natm_all - buf
        self.ranges = {
            "Haml": (rng(buf, buf + per), rng(buf, buf + per)),
            "Hamlv": (rng(buf, buf + per), rng(buf + per, buf + 2 * per)),
            "Hamr": (rng(R - per, R), rng(R - per, R)),
            "Hamrv": (rng(R - per, R), rng(R - 2 * per, R - per)),
            "Hamld": (rng(buf + per, buf + 2 * per), rng(dev0, dev0 + nla)),
            "Hamrd": (rng(R - 2 * per, R - per), rng(dev0 + (self.nlayer - 1) * nla, dev0 + self.nlayer * nla)),
        }
        for il in range(self.nlayer):
            s = dev0 + il * nla
            self.ranges[f"Hamd{il}"] = (rng(s, s + nla), rng(s, s + nla))
            if il + 1 < self.nlayer:
                self.ranges[f"Hamdv{il}"] = (rng(s, s + nla), rng(s + nla, s + 2 * nla))
        self.tables = {k: self._table(*v) for k, v in self.ranges.items()}

    def _table(self, I, J):
        qe = self.qe
        pos_all, pos, idx, mass = qe["pos_all"], qe["pos"], qe["idx_all"], qe["mass_all"]
        rws, rd = qe["rws"], qe["rd"]
        nx, ny, nz = qe["nx"], qe["ny"], qe["nz"]
        rows, cols, coef, drs = [], [], [], []
        for bi, i in enumerate(I):
            for bj, j in enumerate(J):
                for a in range(-self.nscx, self.nscx + 1):
                    for b in range(-self.nscy, self.nscy + 1):
                        dr = -(a * qe["cell_all"][0] + b * qe["cell_all"][1]) + pos_all[j] - pos_all[i]
                        df = rws @ dr - rd
                        if (df > 1.0e-5).any():
                            continue
                        nreq = 1 + int((np.abs(df) <= 1.0e-5).sum())
                        w = 1.0 / nreq
                        ixyz = (dr - (pos[idx[j]] - pos[idx[i]])) @ self.ivcell
                        aa, bb, cc = (-int(np.rint(x)) for x in ixyz)
                        # Fortran 1-based mod(aa+1, nx) wrapped into [1, nx] is aa mod nx, 0-based
                        t1, t2, t3 = aa % nx, bb % ny, cc % nz
                        f = w * qe["fc"][:, :, idx[i], idx[j], t1, t2, t3] / np.sqrt(mass[i] * mass[j])
                        rows.append(bi)
                        cols.append(bj)
                        coef.append(f)
                        drs.append(dr)
        return (np.array(rows), np.array(cols), np.array(coef), np.array(drs), len(I), len(J))

    def _block(self, name, kcart):
        rows, cols, coef, drs, ni, nj = self.tables[name]
        H = np.zeros((3 * ni, 3 * nj), np.complex128)
        ph = np.exp(1j * (drs @ kcart))
        for r, c, f, e in zip(rows, cols, coef, ph):
            H[3 * r:3 * r + 3, 3 * c:3 * c + 3] += f * e
        return H

    def blocks(self, kgrid):
        """tdb blocks at transverse k (crystal coordinates of config.f90:156-175)."""
        kcart = np.asarray(kgrid, float) @ self.cell_reci
        b = {k: self._block(k, kcart) for k in ("Haml", "Hamlv", "Hamr", "Hamrv", "Hamld", "Hamrd")}
        b["Hamd"] = np.stack([self._block(f"Hamd{i}", kcart) for i in range(self.nlayer)])
        b["Hamdv"] = (np.stack([self._block(f"Hamdv{i}", kcart) for i in range(self.nlayer - 1)])
                      if self.nlayer > 1 else np.zeros((0, self.nl, self.nl), np.complex128))
        return b


def k_grid(cfg):
    """negf/config.f90:156-175 (path_mode = 2)."""
    n1, n2 = cfg["nk1"], cfg["nk2"]
    ks = []
    for i in range(n1):
        for j in range(n2):
            ks.append([i / n1 - 0.5 * (n1 - 1) / n1, j / n2 - 0.5 * (n2 - 1) / n2, 0.0])
    ks = np.array(ks)
    if n1 * n2 == 1:
        ks[0, :2] = 0.0
    return ks


def e_grid(cfg):
    """negf/config.f90:201-211: omega^2 on a uniform omega grid, first row emin^2."""
    ne, emax = cfg["ne"], cfg["emax"]
    E = np.array([((i) / (ne - 1) * emax) ** 2 for i in range(ne)])
    E[0] = cfg["emin"] ** 2
    return E


def golden_T(which="transl"):
    """(ne, nk) table of the committed reference output."""
    tab = np.loadtxt(os.path.join(FIXTURE, f"{which}_tot.dat"))
    return tab[:, 0], tab[:, 1:]


def golden_modes(ik):
    """Per-mode rows of trl_specular_param (negf/write_data.f90:3-75) at
    k-point ik: dict omega -> array of (k_z of the incident left mode,
    transmission summed over propagating right modes), rows in file order."""
    tab = np.loadtxt(os.path.join(FIXTURE, "trl_specular_param_%06d.dat" % (ik + 1)))
    out = {}
    for row in np.atleast_2d(tab):
        out.setdefault(round(row[1], 8), []).append((row[4], row[12]))
    return {k: np.array(v) for k, v in out.items()}


if __name__ == "__main__":
    make_fixture()
    print("fixture written to", FIXTURE)
