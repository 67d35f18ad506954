"""The reference's Quantum ESPRESSO force-constant path (tdb branch) on the host.

    read_config         the namelists the path uses (negf/config.f90)
    read_qe_fc          negf/fc.f90:103-325   (IFCs -> THz^2 * amu, cutoff, ASR,
                                              Wigner-Seitz vectors, supercell)
    TdbAssembler        gen_dyn_qe_tdb_k (negf/dyn.f90:1482-2080): WS-weighted
                        Fourier sums into Haml, Hamlv, Hamr, Hamrv, Hamd, Hamdv,
                        Hamld, Hamrd.  The k-independent part (pairs, images,
                        weights, IFC/mass factors, dr) is tabulated once; the
                        phase sum for every k runs on the GPU (upload) or in
                        numpy (blocks).
    k_grid, e_grid      negf/config.f90:156-211 (k grid, Egrid = omega^2)

Constants written as single-precision literals in the reference (masses,
Ry/bohr^2 conversion, bohr2ang) are rounded to float32 as the Fortran
compiler does.
"""
from __future__ import annotations

import os

import numpy as np

F32 = lambda x: float(np.float32(x))        # a REAL(4) literal promoted to REAL(8)

# negf/util.f90:5-8
PI = 3.141592653589793238
ELEVOLT = 1.609e-19
MASS_PROTON = 1.6726219e-27



# mass of each species of the structure files (ChemicalSpeciesLabel of the
# reference's trans_input.fdf: 1 Ge, 2 Si) through element_mass
# (negf/input.f90:410-440; REAL(4) literals, rounded as the compiler does)
SPECIES_MASS = {1: F32(72.64), 2: F32(28.0855)}


def read_config(path):
    """The namelist values the path uses (negf/config.f90)."""
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
                period=int(vals["period_left"]), buffer=int(vals["buffer_left"]),
                lambda_eta=g("lambda_eta") if "lambda_eta" in vals else 5.0e-3,
                # Buttiker probes (negf/config.f90:28-30, defaults :91-98)
                buttiker=int(vals.get("buttiker", 0)), lifetime=int(vals.get("lifetime", 0)),
                t_left=g("t_left") if "t_left" in vals else 300.0,
                t_right=g("t_right") if "t_right" in vals else 300.0,
                B_probe=g("B_probe") if "B_probe" in vals else 5.0e-20 * (2.0 * 3.1415926 * 1.0e12),
                T_probe=g("T_probe") if "T_probe" in vals else 430.0,
                # &kpath (negf/config.f90:64-68): path_mode 2 = full k mesh
                path_mode=int(vals.get("path_mode", 2)), kpath_file=vals.get("kpath_file"))


def read_qe_fc(cfg, path, mass_profile=None):
    """negf/fc.f90:103-325 for the tdb branch."""
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
    mass_profile = mass_profile or os.path.join(os.path.dirname(path), "mass_profile.dat")
    mass_id = [int(x) for x in open(mass_profile).read().split()[1:]]
    mass_all = np.array([SPECIES_MASS[m] for m in mass_id[:len(pos_all)]])
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

    # ------------------------------------------------------------ GPU path
    def _device_tables(self, torch, dev):
        """k-independent tables as device arrays, one per block name:
        (seg_off, seg_rc, coef (nent x 9), dr (nent x 3), nseg, nent, rows, cols)."""
        key = str(dev)
        cache = getattr(self, "_dev_tables", {})
        if key in cache:
            return cache[key]
        out = {}
        for name, (rows, cols, coef, drs, ni, nj) in self.tables.items():
            nent = len(rows)
            if nent:
                rc = np.stack([rows, cols], axis=1)
                change = np.ones(nent, bool)
                change[1:] = np.any(rc[1:] != rc[:-1], axis=1)
                starts = np.nonzero(change)[0]
                seg_off = np.append(starts, nent).astype(np.int32)
                seg_rc = rc[starts].astype(np.int32).reshape(-1)
                c9 = np.ascontiguousarray(coef.reshape(nent, 9), dtype=np.float64)
                d3 = np.ascontiguousarray(drs, dtype=np.float64)
            else:
                seg_off = np.zeros(1, np.int32)
                seg_rc = np.zeros(0, np.int32)
                c9 = np.zeros((0, 9))
                d3 = np.zeros((0, 3))
            t = lambda a: torch.as_tensor(a, device=dev)
            out[name] = (t(seg_off), t(seg_rc), t(c9), t(d3), len(seg_off) - 1, nent, 3 * ni, 3 * nj)
        cache[key] = out
        self._dev_tables = cache
        return out

    def upload(self, engine, kgrids):
        """Device-resident tdb blocks of several k-points (the layout of
        Engine.upload_blocks), assembled on the GPU."""
        import ctypes as C
        import torch
        dev = torch.device("cuda", engine.device)
        tabs = self._device_tables(torch, dev)
        kcart = torch.as_tensor(np.asarray(kgrids, float) @ self.cell_reci, dtype=torch.float64,
                                device=dev).contiguous()
        nk, L, nl = kcart.shape[0], self.nlayer, self.nl
        b = nl
        nr = 3 * len(self.ranges["Hamr"][0])
        shapes = dict(Haml=(nk, nl, nl), Hamlv=(nk, nl, nl), Hamr=(nk, nr, nr), Hamrv=(nk, nr, nr),
                      Hamd=(nk, L, b, b), Hamdv=(nk, max(L - 1, 1), b, b), Hamld=(nk, b, nl), Hamrd=(nk, b, nr))
        out = {k: torch.zeros(v, dtype=torch.complex128, device=dev) for k, v in shapes.items()}
        ptr = lambda t: C.c_void_p(t.data_ptr())

        def run(name, dst, off, kstride):
            seg_off, seg_rc, c9, d3, nseg, nent, rows, cols = tabs[name]
            st = engine._L.nt_hk_assemble_dev(engine.ctx, nk, ptr(kcart), nseg, ptr(seg_off),
                                              ptr(seg_rc) if nseg else None, nent, ptr(c9) if nent else None,
                                              ptr(d3) if nent else None, rows, cols,
                                              C.c_void_p(dst.data_ptr() + 16 * off), rows, kstride)
            engine._check(st, "hk_assemble " + name)
        for name in ("Haml", "Hamlv", "Hamr", "Hamrv", "Hamld", "Hamrd"):
            _, _, _, _, _, _, rows, cols = tabs[name]
            run(name, out[name], 0, rows * cols)
        for il in range(L):
            run(f"Hamd{il}", out["Hamd"], il * b * b, L * b * b)
            if il + 1 < L:
                run(f"Hamdv{il}", out["Hamdv"], il * b * b, max(L - 1, 1) * b * b)
        engine.sync()
        out.update(nl=nl, nr=nr, b=b, L=L, nk=nk)
        return out

    def blocks(self, kgrid):
        """tdb blocks at transverse k (crystal coordinates of config.f90:156-175)."""
        kcart = np.asarray(kgrid, float) @ self.cell_reci
        b = {k: self._block(k, kcart) for k in ("Haml", "Hamlv", "Hamr", "Hamrv", "Hamld", "Hamrd")}
        b["Hamd"] = np.stack([self._block(f"Hamd{i}", kcart) for i in range(self.nlayer)])
        b["Hamdv"] = (np.stack([self._block(f"Hamdv{i}", kcart) for i in range(self.nlayer - 1)])
                      if self.nlayer > 1 else np.zeros((0, self.nl, self.nl), np.complex128))
        return b


# ------------------------------------------------ primitive cell (find_k_pc)
def read_pc_fdf(path, species_mass=None):
    """read_pc_qe (negf/input.f90:580-747) for one lead: the lattice vectors
    (scaled by LatticeConstant, rotated by pi/4 about z), the atoms
    (Cartesian), their masses, the number of atoms inside the cell
    (fractional coordinates in (-0.001, 0.999)) and 2 pi inv(latvec)^T."""
    species_mass = species_mass or SPECIES_MASS
    txt = open(path).read().splitlines()
    nat, scale, lat, atoms = None, 1.0, None, []
    for i, line in enumerate(txt):
        words = line.split()
        if not words:
            continue
        if words[0] == "NumberOfAtoms":
            nat = int(words[1])
        elif words[0] == "LatticeConstant":
            scale = float(words[1])
        elif line.strip().startswith("%block LatticeVectors"):
            lat = np.array([[float(x) for x in txt[i + r].split()[:3]] for r in (1, 2, 3)])
        elif line.strip().startswith("%block AtomicScaledCoordinatesAndAtomicSpecies"):
            atoms = [txt[i + 1 + r].split() for r in range(nat)]
    c, s = np.cos(PI / 4.0), np.sin(PI / 4.0)
    rot = lambda v: np.stack([v[:, 0] * c - v[:, 1] * s, v[:, 0] * s + v[:, 1] * c, v[:, 2]], axis=1)
    latvec = rot(lat * scale)
    frac = np.array([[float(x) for x in a[:3]] for a in atoms])
    pos = frac @ latvec
    mass = np.array([species_mass[int(a[3])] for a in atoms])
    inv = np.linalg.inv(latvec)
    cr = pos @ inv
    npc = int(((cr < 0.999) & (cr > -0.001)).all(axis=1).sum())
    return dict(latvec=latvec, positions=pos, mass=mass, natoms_pc=npc, recivec=2 * PI * inv.T)


def pc_dynamics_table(qe, natoms_pc, mass):
    """gen_dyn_pc_qe_k (negf/dyn.f90:2537-2644) as a phase-sum table in the
    format of nt_hk_assemble_dev: one segment per 3x3 block (primitive atom i,
    QE atom j mod natoms_pc), entries in the reference's summation order (j,
    then the images a, b, c in [-2 n, 2 n]), coefficient w Phi_ij / sqrt(m_i m_j)
    with w = 1 / (number of Wigner-Seitz faces the image lies on + 1), zero
    outside the Wigner-Seitz cell of the force-constant supercell.
    Returns (seg_off, seg_rc, coef (nent x 9), dr (nent x 3))."""
    pos, cell, fc = qe["pos"], qe["cell"], qe["fc"]
    nx, ny, nz = qe["nx"], qe["ny"], qe["nz"]
    grid = np.stack(np.meshgrid(np.arange(-2 * nx, 2 * nx + 1), np.arange(-2 * ny, 2 * ny + 1),
                                np.arange(-2 * nz, 2 * nz + 1), indexing="ij"), axis=-1).reshape(-1, 3)
    shift = grid @ cell
    blocks = {}
    for i in range(natoms_pc):
        for j in range(len(pos)):
            dr = pos[j] - pos[i] - shift
            df = dr @ qe["rws"].T - qe["rd"]
            inside = (df <= 1.0e-5).all(axis=1)
            faces = (np.abs(df) <= 1.0e-5).sum(axis=1)
            # fc index of the image: -nint of its lattice offset, periodic in the supercell
            t = (grid % np.array([nx, ny, nz]))
            for q in np.nonzero(inside)[0]:
                f = fc[:, :, i, j, t[q, 0], t[q, 1], t[q, 2]] / (faces[q] + 1.0) / np.sqrt(mass[i] * mass[j])
                blocks.setdefault((i, j % natoms_pc), []).append((f.reshape(9), dr[q]))
    seg_off, seg_rc, coef, drs = [0], [], [], []
    for (i, jm) in sorted(blocks):
        for f, d in blocks[(i, jm)]:
            coef.append(f)
            drs.append(d)
        seg_rc += [i, jm]
        seg_off.append(len(coef))
    return (np.array(seg_off, np.int32), np.array(seg_rc, np.int32), np.array(coef, float).reshape(-1, 9),
            np.array(drs, float).reshape(-1, 3))


def k_grid(cfg, base_dir=None):
    """negf/config.f90:156-190: the centred mesh for path_mode = 2, else the
    points of kpath_file (first line: their number; then one point per line,
    crystal coordinates) with the transport component set to 0
    (config.f90:118-127, 186-187); base_dir resolves a relative kpath_file."""
    if cfg.get("path_mode", 2) != 2:
        path = cfg["kpath_file"]
        if base_dir and not os.path.isabs(path):
            path = os.path.join(base_dir, path)
        lines = [ln.split() for ln in open(path) if ln.strip()]
        n = int(lines[0][0])
        ks = np.array([[float(x) for x in lines[1 + i][:3]] for i in range(n)])
        ks[:, 2] = 0.0
        return ks
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


