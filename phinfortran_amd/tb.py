"""Tight-binding H(k) of the reference's infortran path, assembled on the GPU.

    read_param   load_param (tb/param.f90:18-172): onsite energies, hopping
                 parameters V and their power laws Vn per species pair and
                 bond type, the equilibrium distances r_d, the orbital lists
    read_fdf     the structure part of load_input (tb/input.f90:50-131)
    sk_blocks    onsite (tb/hamiltonian.f90:57-72) and V_sk
                 (tb/hamiltonian.f90:74-417) for many bonds at once
    TbAssembler  gen_ham (tb/hamiltonian.f90:444-519): the k-independent
                 table of every atom pair and cell image within r_cutoff is
                 built once on the host; H(k) of any number of k-points is
                 assembled on the GPU (nt_hk_assemble_orb_dev) and
                 hermitised, (H + H^+)/2, as the reference does.

Slater-Koster table.  Orbital o on shell s(o) in {s, p, d, s*} (principal
numbers 1, 2, 3, 4).  For the canonical shell pairs (s,s) (s,p) (s,d) (p,p)
(p,d) (d,d) (s*,s*) (s,s*) (s*,p) (s*,d) the element between o1 on atom i and
o2 on atom j is  sum_t f_t(o1, o2; l, m, n) V_t(i, j),  t = sigma, pi, delta,
V_t(i, j) = V(i, j, s1, s2, t) (r_d(i, j) / d)^Vn(i, j, s1, s2, t).  The
element of a reversed pair is (-1)^(l1 + l2) f_t(o2, o1) V_t(j, i).  Within
the d shell the reference takes V(i, j) above the diagonal (order dxy, dyz,
dxz, dx2-y2, dz2) and on it, V(j, i) below it and for dz2-dz2; within p, s,
s* it always takes V(i, j).  sqrt(3) enters as the single-precision value of
the reference's sqrt(3.0).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional

import numpy as np

ANG = 1.0e-10
SQRT3 = float(np.sqrt(np.float32(3.0)))       # the reference's default-real sqrt(3.0)

SHELL = {"s": 1, "px": 2, "py": 2, "pz": 2, "dxy": 3, "dyz": 3, "dxz": 3, "dx2-y2": 3, "dz2": 3, "s*": 4}
COMP = {"s": 0, "s*": 0, "px": 0, "py": 1, "pz": 2, "dxy": 0, "dyz": 1, "dxz": 2, "dx2-y2": 3, "dz2": 4}
LQ = {1: 0, 2: 1, 3: 2, 4: 0}                 # angular momentum of each shell
CANONICAL = {(1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3), (4, 4), (1, 4), (4, 2), (4, 3)}


# ------------------------------------------------------------------ input
def read_param(path: str) -> Dict:
    """The tight-binding parameter file (tb/param.f90:18-172).

    Returns V, Vn (nspecies, nspecies, 4, 4, 4): indices species, species,
    principal number - 1 of each orbital, bond type - 1 (onsite, sigma, pi,
    delta); r_d (nspecies, nspecies) in metres; orbitals: the orbital names
    per species; orb_num.  Shell code 5 in the file is s* (principal 4).
    Onsite blocks fill V(i, i, ., ., onsite) species by species in file
    order; hopping blocks name their species pair 0-based."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    pos = 0

    def upto(tag):
        nonlocal pos
        while pos < len(lines) and tag not in lines[pos]:
            pos += 1
        if pos == len(lines):
            raise ValueError(f"{path}: no {tag}")

    onsite: List = []
    upto("%onsite")
    while "%endonsite" not in lines[pos]:
        if "%block" in lines[pos]:
            head = lines[pos + 1].split()
            nprinc, norb = int(head[1]), int(head[2])
            names = lines[pos + 2].split()[:norb]
            rows = [[float(x) for x in lines[pos + 3 + q].split()[:4]] for q in range(nprinc)]
            onsite.append((names, rows))
            pos += 3 + nprinc
        else:
            pos += 1
    ns = len(onsite)
    V = np.zeros((ns, ns, 4, 4, 4))
    Vn = np.zeros((ns, ns, 4, 4, 4))
    shell = lambda x: 3 if int(x) == 5 else int(x)          # file code -> principal number - 1
    for s, (names, rows) in enumerate(onsite):
        for r in rows:
            if int(r[0]) == 5:
                V[s, s, 3, 3, int(r[2])] = r[3]
            else:
                V[s, s, int(r[0]), int(r[1]), int(r[2])] = r[3]
    upto("%hopping")
    while "%endhopping" not in lines[pos]:
        if "%block" in lines[pos]:
            s1, s2 = (int(x) for x in lines[pos + 1].split()[:2])
            pos += 2
            while "%endblock" not in lines[pos]:
                r = [float(x) for x in lines[pos].split()[:5]]
                a, b, t = shell(r[0]), shell(r[1]), int(r[2]) + 1
                V[s1, s2, a, b, t] = r[3]
                Vn[s1, s2, a, b, t] = r[4]
                pos += 1
        pos += 1
    upto("%rd")
    r_d = np.array([[float(x) for x in lines[pos + 1 + s].split()[:ns]] for s in range(ns)])
    orbitals = [names for names, _ in onsite]
    return dict(V=V, Vn=Vn, r_d=r_d, orbitals=orbitals, orb_num=[len(o) for o in orbitals])


def read_fdf(path: str) -> Dict:
    """latvec (rows, Angstrom), positions (natoms x 3, Angstrom) and 1-based
    species of the structure file (tb/input.f90:50-131); both scaled by
    LatticeConstant."""
    with open(path) as f:
        lines = f.read().splitlines()
    natoms, scale = 0, 1.0
    for ln in lines:
        t = ln.split()
        if len(t) > 1 and t[0] == "NumberOfAtoms":
            natoms = int(t[1])
        elif len(t) > 1 and t[0] == "LatticeConstant":
            scale = float(t[1])
    latvec, pos, sp = None, np.zeros((natoms, 3)), np.zeros(natoms, np.int64)
    for i, ln in enumerate(lines):
        if "%block LatticeVectors" in ln:
            latvec = np.array([[float(x) for x in lines[i + 1 + r].split()[:3]] for r in range(3)]) * scale
        elif "%block AtomicCoordinatesAndAtomicSpecies" in ln:
            for a in range(natoms):
                t = lines[i + 1 + a].split()
                pos[a] = [float(x) * scale for x in t[:3]]
                sp[a] = int(float(t[3]))
    if latvec is None:
        raise ValueError(f"{path}: no LatticeVectors block")
    return dict(latvec=latvec, positions=pos, species=sp)


def diamond_superlattice(a: float, nx: int, ny: int, nz: int, split: bool = True) -> Dict:
    """The structure of the reference's example (examples/tb/gen_struct.py):
    the four-atom tetragonal diamond cell (a/2, -a/2, 0), (a/2, a/2, 0),
    (0, 0, a) repeated nx x ny x nz, atoms ordered by cell (x fastest, then
    y, then z) and basis atom; with split, the first half of the atoms are
    species 2 and the rest species 1, else all species 1."""
    cell = np.array([[a / 2, -a / 2, 0.0], [a / 2, a / 2, 0.0], [0.0, 0.0, a]])
    basis = np.array([[0.0, 0.0, 0.0], [a / 4, a / 4, a / 4], [a / 2, 0.0, a / 2], [a / 4, -a / 4, 3 * a / 4]])
    idx = np.array([(i, j, k) for k in range(nz) for j in range(ny) for i in range(nx)], float)
    pos = (idx @ cell)[:, None, :] + basis[None, :, :]
    pos = pos.reshape(-1, 3)
    nat = len(pos)
    sp = np.ones(nat, np.int64)
    if split:
        sp[: nat // 2] = 2
    return dict(latvec=np.diag([nx, ny, nz]).astype(float) @ cell, positions=pos, species=sp)


def read_confige(path: str) -> Dict:
    """The namelists of the tight-binding driver's configuration file
    (`confige`, tb/config.f90:19-80: configlist, setup, convergelist,
    unfolding, kpath, saving, misori).  Keys are lower-cased; array keys may
    be written with or without "(:)"; several assignments may share a line."""
    import re
    vals: Dict[str, str] = {}
    pat = re.compile(r"([A-Za-z_]\w*)(?:\(:\))?\s*=\s*(.*?)(?=,?\s*[A-Za-z_]\w*(?:\(:\))?\s*=|$)")
    with open(path) as f:
        for line in f:
            line = line.split("!")[0].strip()
            if not line or line.startswith("&") or line == "/":
                continue
            for k, v in pat.findall(line):
                vals[k.lower()] = v.strip().rstrip(",").strip().strip('"').strip("'")
    num = lambda k, d=None: float(vals[k].replace("d", "e").replace("D", "e")) if k in vals else d
    ints = lambda k, d=None: [int(x) for x in vals[k].replace(",", " ").split()] if k in vals else d
    mesh = ints("k_trans_mesh", [1, 1, 1])
    return dict(nk1=mesh[0], nk2=mesh[1], ne=int(vals.get("ne", 1)), emin=num("emin"), emax=num("emax"),
                r_cutoff=num("r_cutoff"), filename_input=vals.get("filename_input"),
                left_primitive_cell_input=vals.get("left_primitive_cell_input"),
                right_primitive_cell_input=vals.get("right_primitive_cell_input"),
                period_left=int(vals.get("period_left", 0)), period_right=int(vals.get("period_right", 0)),
                buffer_left=int(vals.get("buffer_left", 0)), buffer_right=int(vals.get("buffer_right", 0)),
                eta=num("eta"), convergence=num("convergence"),
                n_bloch_x=int(vals.get("n_bloch_x", 1)), n_bloch_y=int(vals.get("n_bloch_y", 1)),
                circular_flag=int(vals.get("circular_flag", 0)),
                path_mode=int(vals.get("path_mode", 2)), kpath_file=vals.get("kpath_file"),
                output_dir=vals.get("output_dir"), verbosity=int(vals.get("verbosity", 0)),
                theta=num("theta", 0.0))


def k_grid(cfg: Dict, base_dir: Optional[str] = None) -> np.ndarray:
    """tb/config.f90:83-186: the centred k_trans_mesh for path_mode = 2
    (the single point of a 1 x 1 mesh at 0), otherwise the points of
    kpath_file (first line: their number) with the third component 0."""
    if cfg["path_mode"] != 2:
        path = cfg["kpath_file"]
        if base_dir and not os.path.isabs(path):
            path = os.path.join(base_dir, path)
        rows = [ln.split() for ln in open(path) if ln.strip()]
        n = int(rows[0][0])
        ks = np.array([[float(x) for x in rows[1 + i][:3]] for i in range(n)])
        ks[:, 2] = 0.0
        return ks
    n1, n2 = cfg["nk1"], cfg["nk2"]
    ks = np.array([[(i - 1) / n1 - 0.5 * (n1 - 1) / n1, (j - 1) / n2 - 0.5 * (n2 - 1) / n2, 0.0]
                   for i in range(1, n1 + 1) for j in range(1, n2 + 1)])
    if n1 * n2 == 1:
        ks[0, :2] = 0.0
    return ks


def k_grid_in_ws(ks: np.ndarray, recivec: np.ndarray) -> np.ndarray:
    """k_grid_in_ws (tb/input.f90:340-382): each k moved by the in-plane
    reciprocal vector n1 b1 + n2 b2 (n in [-2, 2]^2, n1 slowest, first
    minimum) that brings its in-plane part closest to the origin; the
    shift is applied to all three crystal components."""
    a = np.array(recivec, float)
    a[2, :] = 0.0
    shifts = np.array([(i1, i2, 0.0) for i1 in range(-2, 3) for i2 in range(-2, 3)], float)
    out = np.array(ks, float).copy()
    for i in range(len(out)):
        kp = out[i] @ a
        d = kp[None, :] - shifts @ a
        dist = np.einsum("ij,ij->i", d, d)
        out[i] = out[i] - shifts[int(np.argmin(dist))]
    return out


def e_grid(cfg: Dict) -> np.ndarray:
    """tb/config.f90:197-211: ne energies from emin to emax (emax alone for ne = 1)."""
    ne, emin, emax = cfg["ne"], cfg["emin"], cfg["emax"]
    if not emax > emin:
        raise ValueError("emax must exceed emin")
    if ne == 1:
        return np.array([emax])
    return np.array([emin + (i - 1) / (ne - 1) * (emax - emin) for i in range(1, ne + 1)])


def lead_orbitals(P: Dict, species, cfg: Dict) -> Dict:
    """get_periodic_lead_orb_num (tb/surface.f90:11-91): orbitals of the left
    buffer (the first buffer_left atoms), of the right buffer (the last
    buffer_left atoms, as the reference counts them), of one left period
    and one right period of the leads."""
    onum = [P["orb_num"][s - 1] for s in species]
    nat = len(onum)
    bl, br, pl, pr = cfg["buffer_left"], cfg["buffer_right"], cfg["period_left"], cfg["period_right"]
    return dict(n_buffer_l=sum(onum[:bl]), n_buffer_r=sum(onum[nat - bl:]) if bl else 0,
                nl=sum(onum[bl:bl + pl]), nr=sum(onum[nat - br - pr:nat - br]))


def lead_periods(positions, cfg: Dict):
    """The leads' period lengths along z of get_periodic_lead_orb_num
    (tb/surface.f90:59-86): a_z_L = (lowest z of the second left period) -
    (lowest z of the first, atoms buffer_left + 1 .. buffer_left +
    period_left); a_z_R = (highest z of the atoms period_left beyond the
    right lead's inner period) - (highest z of that period), the reference
    offsetting the right lead by period_left."""
    z = np.asarray(positions, float)[:, 2]
    nat = len(z)
    bl, pl, br, pr = cfg["buffer_left"], cfg["period_left"], cfg["buffer_right"], cfg["period_right"]
    i = np.arange(bl, bl + pl)
    a_z_L = z[i + pl].min() - z[i].min()
    j = np.arange(nat - br - 2 * pr, nat - br - pr)
    a_z_R = z[j + pl].max() - z[j].max()
    return float(a_z_L), float(a_z_R)


def unfold_geometry(positions, latvec, species, P: Dict, cfg: Dict):
    """The lead geometry the mode-resolved path's unfolding needs
    (transverse_uc_setup, tb/velocity_mat.f90:14-174, with the lattices of
    tb/input.f90:128-142 and tb/surface.f90:69-90): the transverse unit
    cell latvec_uc (the first two lattice vectors over n_bloch_x and
    n_bloch_y), each lead's cell with its z vector's z component replaced by
    the lead period (reci_uc_l / reci_uc_r), the transverse cells of the
    second left period and of the right lead's inner period (tolerance 1e-3
    and 1e-5) with the orbitals of every atom.  Returns (UnfoldGeometry,
    a_z_L, a_z_R)."""
    from . import unfold as U
    latvec = np.asarray(latvec, float)
    nbx, nby = cfg["n_bloch_x"], cfg["n_bloch_y"]
    latvec_uc = latvec.copy()
    latvec_uc[0] /= nbx
    latvec_uc[1] /= nby
    a_z_L, a_z_R = lead_periods(positions, cfg)
    reci = []
    for a_z in (a_z_L, a_z_R):
        lat = latvec_uc.copy()
        lat[2, 2] = a_z
        reci.append(2 * np.pi * np.linalg.inv(lat).T)
    onum = np.array([P["orb_num"][s - 1] for s in species], int)
    nat = len(onum)
    bl, pl, br, pr = cfg["buffer_left"], cfg["period_left"], cfg["buffer_right"], cfg["period_right"]
    left = U.transverse_cells(positions, latvec_uc, bl + pl, pl, nbx, nby, 1e-3, ndims=onum)
    right = U.transverse_cells(positions, latvec_uc, nat - br - 2 * pr, pr, nbx, nby, 1e-5, ndims=onum)
    recivec = 2 * np.pi * np.linalg.inv(latvec).T
    return U.UnfoldGeometry(nbx, nby, recivec, reci[0], reci[1], left, right), a_z_L, a_z_R


def write_positions(output_dir: str, positions, species, orb_num, layer_list, n_buffer_l: int, nl: int, nk: int):
    """write_positions (tb/write_data.f90:340-360) as blockwise_setup calls it
    (tb/device.f90:232-247): nkpoints.dat (nk, f20.10) and positions_all.dat,
    one row per orbital: the atom's x, y, z, species and the 1-based device
    layer (0 outside the device), format 5f20.10."""
    rows = []
    for a in range(len(positions)):
        for _ in range(orb_num[species[a] - 1]):
            rows.append([*positions[a], float(species[a]), 0.0])
    rows = np.array(rows)
    start = n_buffer_l + 2 * nl
    for i, n in enumerate(layer_list):
        rows[start:start + n, 4] = i + 1
        start += n
    os.makedirs(output_dir, exist_ok=True)
    with open(os.path.join(output_dir, "nkpoints.dat"), "w") as f:
        f.write("%20.10f\n" % float(nk))
    with open(os.path.join(output_dir, "positions_all.dat"), "w") as f:
        for r in rows:
            f.write("".join("%20.10f" % x for x in r) + "\n")


def write_surface_ldos(output_dir: str, sdosl, sdosr):
    """write_surface_ldos (tb/write_data.f90:278-308): sdosl.dat / sdosr.dat,
    one row per energy of f20.10 values (one per k-point) ending in a blank,
    appended to existing files as the reference opens them."""
    os.makedirs(output_dir, exist_ok=True)
    for name, a in (("sdosl.dat", sdosl), ("sdosr.dat", sdosr)):
        with open(os.path.join(output_dir, name), "a") as f:
            for row in np.asarray(a, float):
                f.write("".join("%20.10f" % v for v in row) + " \n")


def write_ldos(output_dir: str, k_index: int, E, ldos):
    """write_ldos (tb/write_data.f90:244-276) for every energy of one k-point:
    ldos_<k_index as I6.6>.dat, one row per energy: E, then ldos(i, j) with the
    layer i outer and the orbital j inner (f20.10 each), ending in a blank;
    appended as the reference opens it.  ldos: (ne, nlayer, maxlen)."""
    os.makedirs(output_dir, exist_ok=True)
    with open(os.path.join(output_dir, "ldos_%06d.dat" % int(k_index)), "a") as f:
        for e, l in zip(np.asarray(E, float), np.asarray(ldos, float)):
            f.write("%20.10f" % e + "".join("%20.10f" % v for v in l.ravel()) + " \n")


def write_device_dos(output_dir: str, ddos):
    """write_device_dos (tb/write_data.f90:310-335): ddos.dat, one row per
    energy of f20.10 values (one per k-point) ending in a blank, appended."""
    os.makedirs(output_dir, exist_ok=True)
    with open(os.path.join(output_dir, "ddos.dat"), "a") as f:
        for row in np.asarray(ddos, float):
            f.write("".join("%20.10f" % v for v in row) + " \n")


# ------------------------------------------------------------------ Slater-Koster
def _angular(A: int, B: int, l, m, n):
    """f_t for the canonical shell pair (A, B): dict t -> array
    (comp(A), comp(B), nbonds); t = 1 sigma, 2 pi, 3 delta (0-based type - 1)."""
    one = np.ones_like(l)
    l2, m2, n2 = l * l, m * m, n * n
    lmn = l * m * n
    dl = l2 - m2
    z2 = n2 - (l2 + m2) / 2
    h, q = SQRT3 / 2, SQRT3 / 4
    # s-d: the d "direction functions" of the sigma bond
    g = [SQRT3 * (l * m), SQRT3 * (m * n), SQRT3 * (l * n), h * dl, z2]
    if (A, B) in ((1, 1), (4, 4), (1, 4)):
        return {1: np.array([[one]])}
    if (A, B) in ((1, 2), (4, 2)):
        return {1: np.array([[l, m, n]])}
    if (A, B) in ((1, 3), (4, 3)):
        return {1: np.array([g])}
    if (A, B) == (2, 2):
        u = [l, m, n]
        sig = np.array([[u[a] * u[b] if a != b else (l2, m2, n2)[a] for b in range(3)] for a in range(3)])
        pi = np.array([[-(u[a] * u[b]) if a != b else 1 - (l2, m2, n2)[a] for b in range(3)] for a in range(3)])
        return {1: sig, 2: pi}
    if (A, B) == (2, 3):
        sig = np.array([
            [SQRT3 * (l2 * m), SQRT3 * lmn, SQRT3 * (l2 * n), h * l * dl, l * z2],
            [SQRT3 * (m2 * l), SQRT3 * (m2 * n), SQRT3 * lmn, h * m * dl, m * z2],
            [SQRT3 * lmn, SQRT3 * (n2 * m), SQRT3 * (n2 * l), h * n * dl, n * z2]])
        pi = np.array([
            [m * (1 - 2 * l2), -(2 * lmn), n * (1 - 2 * l2), l * (1 - l2 + m2), -(SQRT3 * l * n2)],
            [l * (1 - 2 * m2), n * (1 - 2 * m2), -(2 * lmn), -(m * (1 + l2 - m2)), -(SQRT3 * m * n2)],
            [-(2 * lmn), m * (1 - 2 * n2), l * (1 - 2 * n2), -(n * dl), SQRT3 * n * (l2 + m2)]])
        return {1: sig, 2: pi}
    if (A, B) == (3, 3):
        S = [[None] * 5 for _ in range(5)]
        P = [[None] * 5 for _ in range(5)]
        D = [[None] * 5 for _ in range(5)]

        def put(a, b, s, p, d):
            S[a][b] = S[b][a] = s
            P[a][b] = P[b][a] = p
            D[a][b] = D[b][a] = d
        put(0, 0, 3 * l2 * m2, l2 + m2 - 4 * l2 * m2, n2 + l2 * m2)
        put(1, 1, 3 * m2 * n2, m2 + n2 - 4 * m2 * n2, l2 + m2 * n2)
        put(2, 2, 3 * l2 * n2, l2 + n2 - 4 * l2 * n2, m2 + l2 * n2)
        put(0, 1, 3 * l * m2 * n, l * n * (1 - 4 * m2), l * n * (m2 - 1))
        put(0, 2, 3 * m * l2 * n, m * n * (1 - 4 * l2), m * n * (l2 - 1))
        put(1, 2, 3 * m * n2 * l, m * l * (1 - 4 * n2), m * l * (n2 - 1))
        put(0, 3, 1.5 * l * m * dl, 2 * l * m * (m2 - l2), 0.5 * l * m * dl)
        put(1, 3, 1.5 * m * n * dl, -(m * n * (1 + 2 * dl)), m * n * (1 + dl / 2))
        put(2, 3, 1.5 * l * n * dl, l * n * (1 - 2 * dl), -(l * n * (1 - dl / 2)))
        put(0, 4, SQRT3 * l * m * z2, -((2 * SQRT3) * l * m * n2), h * l * m * (1 + n2))
        put(1, 4, SQRT3 * m * n * z2, SQRT3 * m * n * (l2 + m2 - n2), -(h * m * n * (l2 + m2)))
        put(2, 4, SQRT3 * l * n * z2, SQRT3 * l * n * (l2 + m2 - n2), -(h * l * n * (l2 + m2)))
        put(3, 3, 0.75 * dl ** 2, l2 + m2 - dl ** 2, n2 + dl ** 2 / 4)
        put(3, 4, h * dl * z2, SQRT3 * n2 * (m2 - l2), q * (1 + n2) * dl)
        put(4, 4, z2 ** 2, 3 * n2 * (l2 + m2), 0.75 * (l2 + m2) ** 2)
        return {1: np.array(S), 2: np.array(P), 3: np.array(D)}
    raise KeyError((A, B))


def _vterm(P, si, sj, A, B, t, d):
    """V(si, sj, A, B, t) (r_d(si, sj) / d)^Vn(...), d in Angstrom; species 0-based"""
    return P["V"][si, sj, A - 1, B - 1, t] * (P["r_d"][si, sj] / d / ANG) ** P["Vn"][si, sj, A - 1, B - 1, t]


def sk_blocks(P: Dict, si: int, sj: int, r: np.ndarray) -> np.ndarray:
    """Hopping blocks (nbonds, norb_i, norb_j) between an atom of species si
    and one of species sj (both 1-based) at bond vectors r (nbonds x 3,
    Angstrom, from atom i to atom j); elements of orbital pairs the table has
    no case for are zero."""
    r = np.asarray(r, float).reshape(-1, 3)
    d = np.sqrt(r[:, 0] ** 2 + r[:, 1] ** 2 + r[:, 2] ** 2)
    l, m, n = r[:, 0] / d, r[:, 1] / d, r[:, 2] / d
    oi, oj = P["orbitals"][si - 1], P["orbitals"][sj - 1]
    out = np.zeros((len(r), len(oi), len(oj)))
    ang: Dict = {}
    vt: Dict = {}

    def angular(A, B):
        if (A, B) not in ang:
            ang[(A, B)] = _angular(A, B, l, m, n)
        return ang[(A, B)]

    def V(a, b, A, B, t):
        key = (a, b, A, B, t)
        if key not in vt:
            vt[key] = _vterm(P, a, b, A, B, t, d)
        return vt[key]
    for x, o1 in enumerate(oi):
        for y, o2 in enumerate(oj):
            A, B = SHELL.get(o1), SHELL.get(o2)
            if A is None or B is None:
                continue
            c1, c2 = COMP[o1], COMP[o2]
            if (A, B) in CANONICAL:
                f = angular(A, B)
                sign = 1.0
                swap = A == 3 and B == 3 and (c1 > c2 or (c1 == 4 and c2 == 4))
                a, b, i1, i2 = (sj - 1, si - 1, c1, c2) if swap else (si - 1, sj - 1, c1, c2)
                ps = (A, B)
            else:
                f = angular(B, A)
                sign = -1.0 if (LQ[A] + LQ[B]) % 2 else 1.0
                a, b, i1, i2 = sj - 1, si - 1, c2, c1
                ps = (B, A)
            val = np.zeros(len(r))
            for t in sorted(f):
                val = val + f[t][i1, i2] * V(a, b, ps[0], ps[1], t)
            out[:, x, y] = sign * val
    return out


def onsite_block(P: Dict, s: int) -> np.ndarray:
    """onsite (tb/hamiltonian.f90:57-72): diagonal V(s, s, p, p, onsite)."""
    names = P["orbitals"][s - 1]
    return np.diag([P["V"][s - 1, s - 1, SHELL[o] - 1, SHELL[o] - 1, 0] if o in SHELL else 0.0 for o in names])


# ------------------------------------------------------------------ primitive cell
def misorientation_rotation(theta_deg: float = 0.0) -> np.ndarray:
    """rot_all of tb/input.f90:173-182: rot_z^T rot_misori rot_z, rot_z
    holding the default-real sqrt(2.0) / 2.0 (so the rotation of theta = 0 is
    the identity up to single-precision rounding, as in the reference)."""
    c = float(np.float32(np.sqrt(np.float32(2.0))) / np.float32(2.0))
    rz = np.array([[c, -c, 0.0], [c, c, 0.0], [0.0, 0.0, 1.0]])
    th = np.deg2rad(theta_deg)
    rm = np.array([[1.0, 0.0, 0.0], [0.0, np.cos(th), -np.sin(th)], [0.0, np.sin(th), np.cos(th)]])
    return rz.T @ rm @ rz


def read_pc_fdf(path: str, theta_deg: float = 0.0) -> Dict:
    """The primitive-cell part of load_input (tb/input.f90:144-234) for one
    lead: NumberOfAtoms and LatticeConstant, the lattice vectors (scaled,
    rotated by misorientation_rotation), the listed atoms' scaled
    coordinates and species; the atoms whose crystal coordinates lie in
    (-0.001, 0.999) form the primitive cell.  Returns dict(latvec_pc (rows),
    positions_uc (n x 3), species_uc, natoms_uc, positions_pc, species_pc,
    recivec_pc = 2 pi inv(latvec_pc)^T)."""
    lines = open(path).read().splitlines()
    nat, scale = 0, 1.0
    for ln in lines:
        t = ln.split()
        if t and t[0] == "NumberOfAtoms":
            nat = int(t[1])
        elif t and t[0] == "LatticeConstant":
            scale = float(t[1])
    R = misorientation_rotation(theta_deg)
    lat, rows = None, []
    for i, ln in enumerate(lines):
        if "%block LatticeVectors" in ln:
            lat = np.array([R @ (np.array([float(x) for x in lines[i + 1 + r].split()[:3]]) * scale)
                            for r in range(3)])
        elif "%block AtomicScaledCoordinatesAndAtomicSpecies" in ln:
            rows = [[float(x) for x in lines[i + 1 + r].split()[:4]] for r in range(nat)]
    if lat is None or len(rows) != nat:
        raise ValueError(f"{path}: lattice vectors or the {nat} scaled coordinates are missing")
    rows = np.array(rows)
    pos = rows[:, :3] @ lat
    sp = rows[:, 3].astype(int)
    frac = pos @ np.linalg.inv(lat)
    inside = ((frac > -0.001) & (frac < 0.999)).all(axis=1)
    return dict(latvec_pc=lat, positions_uc=pos, species_uc=sp, natoms_uc=nat, positions_pc=pos[inside],
                species_pc=sp[inside], recivec_pc=2 * np.pi * np.linalg.inv(lat).T)


def pc_table(P: Dict, positions_pc, species_pc, latvec_pc, r_cutoff: float):
    """gen_ham_pc_l (tb/hamiltonian.f90:918-989) as an element-wise table:
    for atom pairs (i, j) and images n in [-2, 2]^3 (n1 slowest) with
    r = x_j - x_i + n . latvec_pc shorter than r_cutoff, the block
    onsite (i = j) or V_sk(r) contributes V exp(i k.r) to every element of
    the pair.  Returns (el_off (n^2 + 1; element r + c n column-major),
    coef (nent), dr (nent x 3), n): the input of nt_unfold_pc_tb_dev, which
    also forms gen_hf_pc_l's sums V r_d exp(i k.r) from it."""
    pos = np.asarray(positions_pc, float)
    sp = [int(x) for x in species_pc]
    lat = np.asarray(latvec_pc, float)
    start = np.concatenate([[0], np.cumsum([P["orb_num"][s - 1] for s in sp])]).astype(int)
    n = int(start[-1])
    shifts = np.array([(i1, i2, i3) for i1 in range(-2, 3) for i2 in range(-2, 3) for i3 in range(-2, 3)], float)
    elems: List[List] = [[] for _ in range(n * n)]
    for ia in range(len(pos)):
        for ja in range(len(pos)):
            rs = pos[ja] - pos[ia] + shifts @ lat
            keep = np.sqrt((rs * rs).sum(axis=1)) < r_cutoff
            if not keep.any():
                continue
            rs = rs[keep]
            blocks = (np.repeat(onsite_block(P, sp[ia])[None], len(rs), axis=0) if ia == ja
                      else sk_blocks(P, sp[ia], sp[ja], rs))
            for q in range(len(rs)):
                for x in range(blocks.shape[1]):
                    for y in range(blocks.shape[2]):
                        elems[(start[ia] + x) + (start[ja] + y) * n].append((blocks[q, x, y], rs[q]))
    el_off = np.zeros(n * n + 1, np.int32)
    coef, dr = [], []
    for e, lst in enumerate(elems):
        for v, r in lst:
            coef.append(v)
            dr.append(r)
        el_off[e + 1] = len(coef)
    return el_off, np.array(coef, float), np.array(dr, float).reshape(-1, 3), n


# ------------------------------------------------------------------ assembly
class TbAssembler:
    """gen_ham (tb/hamiltonian.f90:444-519) for one structure.

    Atom pairs (i, j) in atom order, cell images (i1, i2, i3) in [-1, 1]^3
    (i1 slowest), r = x_j - x_i + i1 a1 + i2 a2 + i3 a3 kept when
    |r| < r_cutoff; the block is onsite for i = j (every image of the atom
    itself, as the reference) and V_sk otherwise.  Orbitals run atom by
    atom in the species' orbital order; H(k) = sum block exp(i k.r) with
    k = kpoint . recivec, recivec = 2 pi inv(latvec)^T, then (H + H^+)/2."""

    def __init__(self, P: Dict, latvec, positions, species, r_cutoff: float):
        self.P = P
        self.latvec = np.asarray(latvec, float)
        self.positions = np.asarray(positions, float)
        self.species = np.asarray(species, np.int64)
        self.r_cutoff = float(r_cutoff)
        self.recivec = 2 * np.pi * np.linalg.inv(self.latvec).T
        nat = len(self.positions)
        onum = np.array([P["orb_num"][s - 1] for s in self.species], np.int64)
        self.orb_off = np.concatenate([[0], np.cumsum(onum)]).astype(np.int64)
        self.norb = int(self.orb_off[-1])
        self.cstride = int(max(P["orb_num"]) ** 2)
        shifts = np.array([(i1, i2, i3) for i1 in (-1, 0, 1) for i2 in (-1, 0, 1) for i3 in (-1, 0, 1)], float)
        rs = shifts[:, 0:1] * self.latvec[0] + shifts[:, 1:2] * self.latvec[1] + shifts[:, 2:3] * self.latvec[2]
        # candidate pairs within the cutoff (a little wider; the reference's test decides below)
        from scipy.spatial import cKDTree
        tree = cKDTree(self.positions)
        ent_i, ent_j, ent_s, ent_r = [], [], [], []
        for si in range(27):
            pairs = tree.query_ball_point(self.positions - rs[si], self.r_cutoff * (1 + 1e-9) + 1e-9)
            for i, js in enumerate(pairs):
                for j in js:
                    r = self.positions[j] - self.positions[i]
                    r = r + rs[si]
                    if np.sqrt(r[0] ** 2 + r[1] ** 2 + r[2] ** 2) < self.r_cutoff:
                        ent_i.append(i)
                        ent_j.append(j)
                        ent_s.append(si)
                        ent_r.append(r)
        ent_i, ent_j, ent_s = np.array(ent_i, np.int64), np.array(ent_j, np.int64), np.array(ent_s, np.int64)
        ent_r = np.array(ent_r, float).reshape(-1, 3)
        order = np.lexsort((ent_s, ent_j, ent_i))
        ent_i, ent_j, ent_s, ent_r = ent_i[order], ent_j[order], ent_s[order], ent_r[order]
        nent = len(ent_i)
        coef = np.zeros((nent, self.cstride))
        spi, spj = self.species[ent_i], self.species[ent_j]
        same = ent_i == ent_j
        for a in np.unique(spi):
            for b in np.unique(spj):
                sel = np.nonzero((spi == a) & (spj == b) & ~same)[0]
                if len(sel):
                    blk = sk_blocks(P, int(a), int(b), ent_r[sel])
                    coef[sel, :blk.shape[1] * blk.shape[2]] = blk.reshape(len(sel), -1)
        for a in np.unique(spi):
            sel = np.nonzero((spi == a) & same)[0]
            if len(sel):
                ob = onsite_block(P, int(a)).reshape(-1)
                coef[sel, :len(ob)] = ob
        # one segment per atom pair
        key = ent_i * nat + ent_j
        change = np.ones(nent, bool)
        change[1:] = key[1:] != key[:-1]
        starts = np.nonzero(change)[0]
        self.seg_off = np.append(starts, nent).astype(np.int32)
        si_, sj_ = ent_i[starts], ent_j[starts]
        self.seg_geom = np.stack([self.orb_off[si_], self.orb_off[sj_], onum[si_], onum[sj_]],
                                 axis=1).astype(np.int32).reshape(-1)
        self.coef = coef
        self.dr = ent_r
        self.nseg = len(starts)
        self.nent = nent
        self._dev: Dict = {}

    @classmethod
    def from_files(cls, param_path: str, fdf_path: str, r_cutoff: float) -> "TbAssembler":
        s = read_fdf(fdf_path)
        return cls(read_param(param_path), s["latvec"], s["positions"], s["species"], r_cutoff)

    def kcart(self, kpoints) -> np.ndarray:
        """k-points in crystal coordinates of recivec -> Cartesian (1/Angstrom)"""
        return np.asarray(kpoints, float).reshape(-1, 3) @ self.recivec

    def _device_tables(self, torch, dev):
        key = str(dev)
        if key not in self._dev:
            t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
            self._dev[key] = (t(self.seg_off, torch.int32), t(self.seg_geom, torch.int32),
                              t(self.coef, torch.float64), t(self.dr, torch.float64))
        return self._dev[key]

    def assemble(self, engine, kpoints, out=None, hermitize: bool = True):
        """H(k) of every k-point (crystal coordinates, nk x 3) on the engine's
        device: complex128 tensor (nk, norb, norb) holding column-major
        matrices (element (r, c) of k at [k, c, r]); `out` may supply that
        tensor.  Use .transpose(-1, -2) for row-major indexing."""
        import torch
        dev = torch.device("cuda", engine.device)
        seg_off, seg_geom, coef, dr = self._device_tables(torch, dev)
        kc = torch.as_tensor(self.kcart(kpoints), dtype=torch.float64, device=dev).contiguous()
        nk, N = kc.shape[0], self.norb
        if out is None:
            out = torch.empty((nk, N, N), dtype=torch.complex128, device=dev)
        elif (tuple(out.shape) != (nk, N, N) or out.dtype != torch.complex128 or not out.is_contiguous()
              or out.device != dev):
            raise ValueError(f"out must be a contiguous complex128 ({nk}, {N}, {N}) tensor on {dev}")
        p = lambda x: C.c_void_p(x.data_ptr())
        has = self.nseg > 0
        st = engine._L.nt_hk_assemble_orb_dev(engine.ctx, nk, p(kc), self.nseg, p(seg_off),
                                              p(seg_geom) if has else None, self.nent, self.cstride,
                                              p(coef) if has else None, p(dr) if has else None, N, N, p(out), N,
                                              N * N, 1 if hermitize else 0)
        engine._check(st, "hk_assemble_orb")
        return out
