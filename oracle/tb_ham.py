"""CPU restatement of the tight-binding H(k) builder (test infrastructure only:
tests/ import it, the product path never does).

    read_param   load_param (tb/param.f90:18-...): onsite and hopping blocks,
                 V(species, species, n1, n2, type), Vn, r_d
    read_fdf     the structure part of load_input (tb/input.f90:50-112)
    v_term       V_term (tb/hamiltonian.f90:28-55): V (d0 / d)^Vn
    onsite       tb/hamiltonian.f90:57-72
    v_sk         V_sk (tb/hamiltonian.f90:74-417): the reference's
                 Slater-Koster cases, one by one, with its default-real
                 literals (sqrt(3.0) is a single-precision value)
    gen_ham      gen_ham (tb/hamiltonian.f90:444-519): sum over atom pairs and
                 the 27 cell images within r_cutoff of V exp(i r.k), then
                 hermitian = (H + H^+)/2 (tb/util.f90:709-718)

Indices follow the reference: species 1-based, principal numbers 1 = s,
2 = p, 3 = d, 4 = s*, bond types 1 = onsite, 2 = sigma, 3 = pi, 4 = delta.
Pinned (tests/test_tb_ham_oracle.py) by the covariance of the whole table
under rotations of the bond (every pair block transforms with the s, p, d
rotation matrices), the transposition symmetry V(o1 o2; r) = V(o2 o1; -r)
with the species swapped, the power law, and the bulk band edges of Si with
the example's parameters.
"""
import numpy as np

ANG = 1.0e-10
SQ3 = float(np.sqrt(np.float32(3.0)))          # sqrt(3.0): a default (single precision) real


def read_param(path):
    """tb/param.f90: returns dict(V, Vn (ns, ns, 4, 4, 4), r_d (ns, ns),
    orbitals (list per species), orb_num (per species))."""
    lines = open(path).read().splitlines()
    i = 0
    onsite = []          # per species: (orbital names, [(n1, n2, type, value)])
    while not lines[i].strip().startswith("%endonsite"):
        if lines[i].strip().startswith("%block"):
            sp, nprinc, norb = (int(x) for x in lines[i + 1].split()[:3])
            names = lines[i + 2].split()[:norb]
            rows = [lines[i + 3 + k].split() for k in range(nprinc)]
            onsite.append((sp, names, rows))
            i += 3 + nprinc
        else:
            i += 1
    ns = len(onsite)
    V = np.zeros((ns, ns, 4, 4, 4))
    Vn = np.zeros((ns, ns, 4, 4, 4))
    orbitals = [None] * ns
    for sp, names, rows in onsite:
        orbitals[sp - 1] = names
        for r in rows:
            n1, n2, ty, val = int(float(r[0])), int(float(r[1])), int(float(r[2])), float(r[3])
            a = 3 if n1 == 5 else n1
            b = 3 if n1 == 5 else n2
            V[sp - 1, sp - 1, a, b, ty] = val
    while not lines[i].strip().startswith("%endhopping"):
        if lines[i].strip().startswith("%block"):
            s1, s2 = (int(x) for x in lines[i + 1].split()[:2])
            k = i + 2
            while not lines[k].strip().startswith("%endblock"):
                r = lines[k].split()
                n1, n2, ty = int(float(r[0])), int(float(r[1])), int(float(r[2]))
                a = 3 if n1 == 5 else n1
                b = 3 if n2 == 5 else n2
                V[s1, s2, a, b, ty + 1] = float(r[3])
                Vn[s1, s2, a, b, ty + 1] = float(r[4])
                k += 1
            i = k
        i += 1
    while not lines[i].strip().startswith("%rd"):
        i += 1
    r_d = np.array([[float(x) for x in lines[i + 1 + s].split()[:ns]] for s in range(ns)])
    return dict(V=V, Vn=Vn, r_d=r_d, orbitals=orbitals, orb_num=[len(o) for o in orbitals])


def read_fdf(path):
    """latvec (rows, Angstrom), positions (N x 3, Angstrom) and 1-based
    species of the structure file (tb/input.f90:50-112)."""
    lines = open(path).read().splitlines()
    scale, natoms = 1.0, 0
    for ln in lines:
        t = ln.split()
        if t and t[0] == "NumberOfAtoms":
            natoms = int(t[1])
        if t and t[0] == "LatticeConstant":
            scale = float(t[1])
    lat, pos, sp = None, [], []
    for i, ln in enumerate(lines):
        if "%block LatticeVectors" in ln:
            lat = np.array([[float(x) for x in lines[i + 1 + r].split()[:3]] for r in range(3)]) * scale
        if "%block AtomicCoordinatesAndAtomicSpecies" in ln:
            for r in range(natoms):
                t = lines[i + 1 + r].split()
                pos.append([float(x) * scale for x in t[:3]])
                sp.append(int(float(t[3])))
    return dict(latvec=lat, positions=np.array(pos), species=np.array(sp))


def principal(orb):
    """get_principal_number (tb/hamiltonian.f90:6-26)"""
    o = orb.strip()
    if o.startswith("s*"):
        return 4
    return {"s": 1, "p": 2, "d": 3}[o[0]]


def v_term(P, i, j, n1, n2, ty, dist):
    """V(i, j, n1, n2, ty) (d0 / dist)^Vn, dist in Angstrom, d0 = r_d(i, j) in metres"""
    return P["V"][i - 1, j - 1, n1 - 1, n2 - 1, ty - 1] * (P["r_d"][i - 1, j - 1] / dist / ANG) ** \
        P["Vn"][i - 1, j - 1, n1 - 1, n2 - 1, ty - 1]


def onsite(P, i, j, o1, o2):
    if o1 != o2:
        return 0.0
    return P["V"][i - 1, j - 1, principal(o1) - 1, principal(o2) - 1, 0]


def v_sk(P, i, j, o1, o2, r):
    """The reference's Slater-Koster case for orbital o1 of species i and o2
    of species j separated by r (Angstrom); 0 for a pair it has no case for."""
    d = float(np.sqrt(r[0] ** 2 + r[1] ** 2 + r[2] ** 2))
    l, m, n = r[0] / d, r[1] / d, r[2] / d

    def T(n1, n2, ty):             # V_term(i, j, ...)
        return v_term(P, i, j, n1, n2, ty, d)

    def R(n1, n2, ty):             # V_term(j, i, ...)
        return v_term(P, j, i, n1, n2, ty, d)

    h = SQ3 / 2
    q = SQ3 / 4
    z2 = n ** 2 - (l ** 2 + m ** 2) / 2
    c = {
        # s, p
        "ss": lambda: T(1, 1, 2),
        "spx": lambda: l * T(1, 2, 2), "pxs": lambda: -l * R(1, 2, 2),
        "spy": lambda: m * T(1, 2, 2), "pys": lambda: -m * R(1, 2, 2),
        "spz": lambda: n * T(1, 2, 2), "pzs": lambda: -n * R(1, 2, 2),
        "pxpx": lambda: l ** 2 * T(2, 2, 2) + (1 - l ** 2) * T(2, 2, 3),
        "pypy": lambda: m ** 2 * T(2, 2, 2) + (1 - m ** 2) * T(2, 2, 3),
        "pzpz": lambda: n ** 2 * T(2, 2, 2) + (1 - n ** 2) * T(2, 2, 3),
        "pxpy": lambda: l * m * T(2, 2, 2) - l * m * T(2, 2, 3),
        "pypx": lambda: l * m * T(2, 2, 2) - l * m * T(2, 2, 3),
        "pypz": lambda: m * n * T(2, 2, 2) - m * n * T(2, 2, 3),
        "pzpy": lambda: m * n * T(2, 2, 2) - m * n * T(2, 2, 3),
        "pxpz": lambda: l * n * T(2, 2, 2) - l * n * T(2, 2, 3),
        "pzpx": lambda: l * n * T(2, 2, 2) - l * n * T(2, 2, 3),
        # s, d
        "sdxy": lambda: SQ3 * (l * m) * T(1, 3, 2), "dxys": lambda: SQ3 * (l * m) * R(1, 3, 2),
        "sdyz": lambda: SQ3 * (m * n) * T(1, 3, 2), "dyzs": lambda: SQ3 * (m * n) * R(1, 3, 2),
        "sdxz": lambda: SQ3 * (l * n) * T(1, 3, 2), "dxzs": lambda: SQ3 * (l * n) * R(1, 3, 2),
        "sdx2-y2": lambda: h * (l ** 2 - m ** 2) * T(1, 3, 2),
        "dx2-y2s": lambda: h * (l ** 2 - m ** 2) * R(1, 3, 2),
        "sdz2": lambda: z2 * T(1, 3, 2), "dz2s": lambda: z2 * R(1, 3, 2),
        # p, d
        "pxdxy": lambda: SQ3 * (l ** 2 * m) * T(2, 3, 2) + m * (1 - 2 * l ** 2) * T(2, 3, 3),
        "dxypx": lambda: -SQ3 * (l ** 2 * m) * R(2, 3, 2) - m * (1 - 2 * l ** 2) * R(2, 3, 3),
        "pxdyz": lambda: SQ3 * (l * m * n) * T(2, 3, 2) - 2 * (l * m * n) * T(2, 3, 3),
        "dyzpx": lambda: -SQ3 * (l * m * n) * R(2, 3, 2) + 2 * (l * m * n) * R(2, 3, 3),
        "pxdxz": lambda: SQ3 * (l ** 2 * n) * T(2, 3, 2) + n * (1 - 2 * l ** 2) * T(2, 3, 3),
        "dxzpx": lambda: -SQ3 * (l ** 2 * n) * R(2, 3, 2) - n * (1 - 2 * l ** 2) * R(2, 3, 3),
        "pydxy": lambda: SQ3 * (m ** 2 * l) * T(2, 3, 2) + l * (1 - 2 * m ** 2) * T(2, 3, 3),
        "dxypy": lambda: -SQ3 * (m ** 2 * l) * R(2, 3, 2) - l * (1 - 2 * m ** 2) * R(2, 3, 3),
        "pydyz": lambda: SQ3 * (m ** 2 * n) * T(2, 3, 2) + n * (1 - 2 * m ** 2) * T(2, 3, 3),
        "dyzpy": lambda: -SQ3 * (m ** 2 * n) * R(2, 3, 2) - n * (1 - 2 * m ** 2) * R(2, 3, 3),
        "pydxz": lambda: SQ3 * (l * m * n) * T(2, 3, 2) - 2 * (l * m * n) * T(2, 3, 3),
        "dxzpy": lambda: -SQ3 * (l * m * n) * R(2, 3, 2) + 2 * (l * m * n) * R(2, 3, 3),
        "pzdxy": lambda: SQ3 * (l * m * n) * T(2, 3, 2) - 2 * (l * m * n) * T(2, 3, 3),
        "dxypz": lambda: -SQ3 * (l * m * n) * R(2, 3, 2) + 2 * (l * m * n) * R(2, 3, 3),
        "pzdyz": lambda: SQ3 * (n ** 2 * m) * T(2, 3, 2) + m * (1 - 2 * n ** 2) * T(2, 3, 3),
        "dyzpz": lambda: -SQ3 * (n ** 2 * m) * R(2, 3, 2) - m * (1 - 2 * n ** 2) * R(2, 3, 3),
        "pzdxz": lambda: SQ3 * (n ** 2 * l) * T(2, 3, 2) + l * (1 - 2 * n ** 2) * T(2, 3, 3),
        "dxzpz": lambda: -SQ3 * (n ** 2 * l) * R(2, 3, 2) - l * (1 - 2 * n ** 2) * R(2, 3, 3),
        "pxdx2-y2": lambda: h * l * (l ** 2 - m ** 2) * T(2, 3, 2) + l * (1 - l ** 2 + m ** 2) * T(2, 3, 3),
        "dx2-y2px": lambda: -h * l * (l ** 2 - m ** 2) * R(2, 3, 2) - l * (1 - l ** 2 + m ** 2) * R(2, 3, 3),
        "pydx2-y2": lambda: h * m * (l ** 2 - m ** 2) * T(2, 3, 2) - m * (1 + l ** 2 - m ** 2) * T(2, 3, 3),
        "dx2-y2py": lambda: -h * m * (l ** 2 - m ** 2) * R(2, 3, 2) + m * (1 + l ** 2 - m ** 2) * R(2, 3, 3),
        "pzdx2-y2": lambda: h * n * (l ** 2 - m ** 2) * T(2, 3, 2) - n * (l ** 2 - m ** 2) * T(2, 3, 3),
        "dx2-y2pz": lambda: -h * n * (l ** 2 - m ** 2) * R(2, 3, 2) + n * (l ** 2 - m ** 2) * R(2, 3, 3),
        "pxdz2": lambda: l * z2 * T(2, 3, 2) - SQ3 * l * n ** 2 * T(2, 3, 3),
        "dz2px": lambda: -l * z2 * R(2, 3, 2) + SQ3 * l * n ** 2 * R(2, 3, 3),
        "pydz2": lambda: m * z2 * T(2, 3, 2) - SQ3 * m * n ** 2 * T(2, 3, 3),
        "dz2py": lambda: -m * z2 * R(2, 3, 2) + SQ3 * m * n ** 2 * R(2, 3, 3),
        "pzdz2": lambda: n * z2 * T(2, 3, 2) + SQ3 * n * (l ** 2 + m ** 2) * T(2, 3, 3),
        "dz2pz": lambda: -n * z2 * R(2, 3, 2) - SQ3 * n * (l ** 2 + m ** 2) * R(2, 3, 3),
        # d, d
        "dxydxy": lambda: 3 * l ** 2 * m ** 2 * T(3, 3, 2) + (l ** 2 + m ** 2 - 4 * l ** 2 * m ** 2) * T(3, 3, 3)
        + (n ** 2 + l ** 2 * m ** 2) * T(3, 3, 4),
        "dyzdyz": lambda: 3 * m ** 2 * n ** 2 * T(3, 3, 2) + (m ** 2 + n ** 2 - 4 * m ** 2 * n ** 2) * T(3, 3, 3)
        + (l ** 2 + m ** 2 * n ** 2) * T(3, 3, 4),
        "dxzdxz": lambda: 3 * l ** 2 * n ** 2 * T(3, 3, 2) + (l ** 2 + n ** 2 - 4 * l ** 2 * n ** 2) * T(3, 3, 3)
        + (m ** 2 + l ** 2 * n ** 2) * T(3, 3, 4),
        "dxydyz": lambda: 3 * l * m ** 2 * n * T(3, 3, 2) + l * n * (1 - 4 * m ** 2) * T(3, 3, 3)
        + l * n * (m ** 2 - 1) * T(3, 3, 4),
        "dyzdxy": lambda: 3 * l * m ** 2 * n * R(3, 3, 2) + l * n * (1 - 4 * m ** 2) * R(3, 3, 3)
        + l * n * (m ** 2 - 1) * R(3, 3, 4),
        "dxydxz": lambda: 3 * m * l ** 2 * n * T(3, 3, 2) + m * n * (1 - 4 * l ** 2) * T(3, 3, 3)
        + m * n * (l ** 2 - 1) * T(3, 3, 4),
        "dxzdxy": lambda: 3 * m * l ** 2 * n * R(3, 3, 2) + m * n * (1 - 4 * l ** 2) * R(3, 3, 3)
        + m * n * (l ** 2 - 1) * R(3, 3, 4),
        "dyzdxz": lambda: 3 * m * n ** 2 * l * T(3, 3, 2) + m * l * (1 - 4 * n ** 2) * T(3, 3, 3)
        + m * l * (n ** 2 - 1) * T(3, 3, 4),
        "dxzdyz": lambda: 3 * m * n ** 2 * l * R(3, 3, 2) + m * l * (1 - 4 * n ** 2) * R(3, 3, 3)
        + m * l * (n ** 2 - 1) * R(3, 3, 4),
        "dxydx2-y2": lambda: 1.5 * l * m * (l ** 2 - m ** 2) * T(3, 3, 2) + 2 * l * m * (m ** 2 - l ** 2) * T(3, 3, 3)
        + 0.5 * l * m * (l ** 2 - m ** 2) * T(3, 3, 4),
        "dx2-y2dxy": lambda: 1.5 * l * m * (l ** 2 - m ** 2) * R(3, 3, 2) + 2 * l * m * (m ** 2 - l ** 2) * R(3, 3, 3)
        + 0.5 * l * m * (l ** 2 - m ** 2) * R(3, 3, 4),
        "dyzdx2-y2": lambda: 1.5 * m * n * (l ** 2 - m ** 2) * T(3, 3, 2) - m * n * (1 + 2 * (l ** 2 - m ** 2)) * T(3, 3, 3)
        + m * n * (1 + (l ** 2 - m ** 2) / 2) * T(3, 3, 4),
        "dx2-y2dyz": lambda: 1.5 * m * n * (l ** 2 - m ** 2) * R(3, 3, 2) - m * n * (1 + 2 * (l ** 2 - m ** 2)) * R(3, 3, 3)
        + m * n * (1 + (l ** 2 - m ** 2) / 2) * R(3, 3, 4),
        "dxzdx2-y2": lambda: 1.5 * l * n * (l ** 2 - m ** 2) * T(3, 3, 2) + l * n * (1 - 2 * (l ** 2 - m ** 2)) * T(3, 3, 3)
        - l * n * (1 - (l ** 2 - m ** 2) / 2) * T(3, 3, 4),
        "dx2-y2dxz": lambda: 1.5 * l * n * (l ** 2 - m ** 2) * R(3, 3, 2) + l * n * (1 - 2 * (l ** 2 - m ** 2)) * R(3, 3, 3)
        - l * n * (1 - (l ** 2 - m ** 2) / 2) * R(3, 3, 4),
        "dxydz2": lambda: SQ3 * l * m * z2 * T(3, 3, 2) - (2 * SQ3) * l * m * n ** 2 * T(3, 3, 3)
        + h * l * m * (1 + n ** 2) * T(3, 3, 4),
        "dz2dxy": lambda: SQ3 * l * m * z2 * R(3, 3, 2) - (2 * SQ3) * l * m * n ** 2 * R(3, 3, 3)
        + h * l * m * (1 + n ** 2) * R(3, 3, 4),
        "dyzdz2": lambda: SQ3 * m * n * z2 * T(3, 3, 2) + SQ3 * m * n * (l ** 2 + m ** 2 - n ** 2) * T(3, 3, 3)
        - h * m * n * (l ** 2 + m ** 2) * T(3, 3, 4),
        "dz2dyz": lambda: SQ3 * m * n * z2 * R(3, 3, 2) + SQ3 * m * n * (l ** 2 + m ** 2 - n ** 2) * R(3, 3, 3)
        - h * m * n * (l ** 2 + m ** 2) * R(3, 3, 4),
        "dxzdz2": lambda: SQ3 * l * n * z2 * T(3, 3, 2) + SQ3 * l * n * (l ** 2 + m ** 2 - n ** 2) * T(3, 3, 3)
        - h * l * n * (l ** 2 + m ** 2) * T(3, 3, 4),
        "dz2dxz": lambda: SQ3 * l * n * z2 * R(3, 3, 2) + SQ3 * l * n * (l ** 2 + m ** 2 - n ** 2) * R(3, 3, 3)
        - h * l * n * (l ** 2 + m ** 2) * R(3, 3, 4),
        "dx2-y2dx2-y2": lambda: 0.75 * (l ** 2 - m ** 2) ** 2 * T(3, 3, 2)
        + (l ** 2 + m ** 2 - (l ** 2 - m ** 2) ** 2) * T(3, 3, 3) + (n ** 2 + (l ** 2 - m ** 2) ** 2 / 4) * T(3, 3, 4),
        "dx2-y2dz2": lambda: h * (l ** 2 - m ** 2) * z2 * T(3, 3, 2) + SQ3 * n ** 2 * (m ** 2 - l ** 2) * T(3, 3, 3)
        + q * (1 + n ** 2) * (l ** 2 - m ** 2) * T(3, 3, 4),
        "dz2dx2-y2": lambda: h * (l ** 2 - m ** 2) * z2 * R(3, 3, 2) + SQ3 * n ** 2 * (m ** 2 - l ** 2) * R(3, 3, 3)
        + q * (1 + n ** 2) * (l ** 2 - m ** 2) * R(3, 3, 4),
        "dz2dz2": lambda: z2 ** 2 * R(3, 3, 2) + 3 * n ** 2 * (l ** 2 + m ** 2) * R(3, 3, 3)
        + 0.75 * (l ** 2 + m ** 2) ** 2 * R(3, 3, 4),
        # s*
        "s*s*": lambda: T(4, 4, 2), "ss*": lambda: T(1, 4, 2), "s*s": lambda: R(1, 4, 2),
        "s*px": lambda: l * T(4, 2, 2), "pxs*": lambda: -l * R(4, 2, 2),
        "s*py": lambda: m * T(4, 2, 2), "pys*": lambda: -m * R(4, 2, 2),
        "s*pz": lambda: n * T(4, 2, 2), "pzs*": lambda: -n * R(4, 2, 2),
        "s*dxy": lambda: SQ3 * (l * m) * T(4, 3, 2), "dxys*": lambda: SQ3 * (l * m) * R(4, 3, 2),
        "s*dyz": lambda: SQ3 * (m * n) * T(4, 3, 2), "dyzs*": lambda: SQ3 * (m * n) * R(4, 3, 2),
        "s*dxz": lambda: SQ3 * (l * n) * T(4, 3, 2), "dxzs*": lambda: SQ3 * (l * n) * R(4, 3, 2),
        "s*dx2-y2": lambda: h * (l ** 2 - m ** 2) * T(4, 3, 2), "dx2-y2s*": lambda: h * (l ** 2 - m ** 2) * R(4, 3, 2),
        "s*dz2": lambda: z2 * T(4, 3, 2), "dz2s*": lambda: z2 * R(4, 3, 2),
    }
    f = c.get(o1.strip() + o2.strip())
    return 0.0 if f is None else float(f())


def gen_ham(P, positions, species, latvec, kpoint, r_cutoff):
    """Dense H(k) of the structure: kpoint in crystal coordinates of
    recivec = 2 pi inv(latvec)^T; orbitals atom by atom in the order of the
    species' orbital list."""
    recivec = 2 * np.pi * np.linalg.inv(latvec).T
    kc = np.asarray(kpoint) @ recivec
    orbs = [(a, o) for a in range(len(positions)) for o in P["orbitals"][species[a] - 1]]
    N = len(orbs)
    H = np.zeros((N, N), complex)
    shifts = [(i1, i2, i3) for i1 in (-1, 0, 1) for i2 in (-1, 0, 1) for i3 in (-1, 0, 1)]
    for ia in range(len(positions)):
        for ja in range(len(positions)):
            ti, tj = species[ia], species[ja]
            for s in shifts:
                r = positions[ja] - positions[ia] + np.array(s, float) @ latvec
                if not np.sqrt(r @ r) < r_cutoff:
                    continue
                ph = np.exp(1j * (r @ kc))
                for io, o1 in enumerate(P["orbitals"][ti - 1]):
                    row = sum(P["orb_num"][species[a] - 1] for a in range(ia)) + io
                    for jo, o2 in enumerate(P["orbitals"][tj - 1]):
                        col = sum(P["orb_num"][species[a] - 1] for a in range(ja)) + jo
                        v = onsite(P, ti, tj, o1, o2) if ia == ja else v_sk(P, ti, tj, o1, o2, r)
                        H[row, col] += ph * v
    return 0.5 * (H + H.conj().T)
