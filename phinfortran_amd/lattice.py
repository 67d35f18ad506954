"""Seeded block-Hamiltonian builders for the transport workloads.

These produce the per-k block arrays that phinfortran's driver hands to its
per-(omega,k) solve (the tdb layout of negf/phinfortran.f90:1188-1250):

    Haml, Hamlv   left-lead H00 and V_{-1,0}            (nl x nl)
    Hamr, Hamrv   right-lead H00 and V_{N+2,N+1}        (nr x nr)
    Hamd          device diagonal blocks                 (L, b, b)
    Hamdv         device couplings V_{i,i+1}             (L-1, b, b)
    Hamld         V_{0,1}  left lead -> first layer      (nl x b)
    Hamrd         V_{N+1,N} right lead -> last layer     (nr x b)

The reference cannot express these inputs in its decks (it has no spring-model
generator and no per-site or per-bond disorder), so they are built here from
the physical description of each workload.

Phonons: diamond lattice (a = 5.431 A), principal layer = one conventional
cubic cell along z, nearest-neighbour central (alpha) plus transverse
(beta) springs, D = Phi / sqrt(m_i m_j) in THz^2 (the reference's unit,
negf/dyn.f90:138: E = f^2 with f in THz).

Electrons: p_z tight binding on an armchair graphene nanoribbon, principal
layer = one armchair unit cell (3 a_cc long, 2 N_dimer orbitals).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

AMU_KG = 1.66053906660e-27
THZ2_PER_N_M_AMU = 1.0 / AMU_KG / (2.0 * math.pi * 1.0e12) ** 2   # (N/m)/amu -> THz^2
MASS_SI = 28.0855
MASS_GE = 72.63
A_SI = 5.431

_DIAMOND = np.array([
    [0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0],
    [0.25, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.25, 0.75], [0.75, 0.75, 0.25]])
_NN = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=float) * 0.25


@dataclasses.dataclass
class Blocks:
    """tdb block set for one k-point (numpy complex128, C-contiguous per block)."""
    Haml: np.ndarray
    Hamlv: np.ndarray
    Hamr: np.ndarray
    Hamrv: np.ndarray
    Hamd: np.ndarray
    Hamdv: np.ndarray
    Hamld: np.ndarray
    Hamrd: np.ndarray

    @property
    def nl(self):
        return self.Haml.shape[0]

    @property
    def nr(self):
        return self.Hamr.shape[0]

    @property
    def b(self):
        return self.Hamd.shape[1]

    @property
    def L(self):
        return self.Hamd.shape[0]

    def as_dict(self):
        return dataclasses.asdict(self)

    def full_hamiltonian(self):
        """Dense block-tridiagonal H over [lead layer 0, device 1..L, lead layer N+1]."""
        nl, nr, b, L = self.nl, self.nr, self.b, self.L
        n = nl + L * b + nr
        H = np.zeros((n, n), np.complex128)
        H[:nl, :nl] = self.Haml
        H[:nl, nl:nl + b] = self.Hamld
        H[nl:nl + b, :nl] = self.Hamld.conj().T
        for i in range(L):
            s = nl + i * b
            H[s:s + b, s:s + b] = self.Hamd[i]
            if i + 1 < L:
                H[s:s + b, s + b:s + 2 * b] = self.Hamdv[i]
                H[s + b:s + 2 * b, s:s + b] = self.Hamdv[i].conj().T
        s = nl + (L - 1) * b
        H[n - nr:, s:s + b] = self.Hamrd
        H[s:s + b, n - nr:] = self.Hamrd.conj().T
        H[n - nr:, n - nr:] = self.Hamr
        return H


class DiamondLayer:
    """Geometry and bond topology of one principal layer (nx x ny conventional
    cells, one cell thick).  periodic=False gives a free-standing wire,
    periodic=True a film with in-plane supercell nx*a x ny*a."""

    def __init__(self, nx: int, ny: int, periodic: bool, a: float = A_SI):
        self.nx, self.ny, self.a, self.periodic = nx, ny, a, periodic
        cells = [(i, j) for i in range(nx) for j in range(ny)]
        pos = np.array([(np.array([i, j, 0.0]) + f) for (i, j) in cells for f in _DIAMOND])
        self.frac = pos                       # in units of a
        self.natom = len(pos)
        self.ndof = 3 * self.natom
        key = {self._key(p): k for k, p in enumerate(pos)}
        bonds = []   # (i, j, dlayer, Rx, Ry, e)
        for i, p in enumerate(pos):
            sub_a = (round(p[0] * 4) + round(p[1] * 4) + round(p[2] * 4)) % 4 == 0 and \
                    round((p[0] % 1) * 4) % 2 == 0
            for d in _NN:
                q = p + (d if sub_a else -d)
                dl = int(math.floor(q[2] + 1e-9))
                q2 = q.copy()
                q2[2] -= dl
                Rx = int(math.floor(q2[0] / nx + 1e-9))
                Ry = int(math.floor(q2[1] / ny + 1e-9))
                if (Rx or Ry) and not periodic:
                    continue
                q2[0] -= Rx * nx
                q2[1] -= Ry * ny
                j = key.get(self._key(q2))
                if j is None:
                    raise RuntimeError("diamond neighbour not found")
                e = (d if sub_a else -d) / np.linalg.norm(d)
                bonds.append((i, j, dl, Rx, Ry, e))
        self.bonds = bonds
        self.nbond = len(bonds)

    @staticmethod
    def _key(p):
        return tuple(int(round(x * 4)) for x in p)

    def phi(self, e, alpha, beta):
        """Force-constant tensor of one bond (N/m), Phi_ij = -(alpha e e^T + beta (1 - e e^T))."""
        ee = np.outer(e, e)
        return -(alpha * ee + beta * (np.eye(3) - ee))


def _phase(R, kfrac):
    return np.exp(2j * np.pi * (R[0] * kfrac[0] + R[1] * kfrac[1]))


def bond_owner(layer: DiamondLayer, L: int, l: int, bi: int):
    """(owning device layer, owning bond index) of directed bond bi of layer l
    (1-based layers), or None for a bond touching a lead layer.  A bond
    between two device layers is owned by the lower layer (its dl=+1 entry),
    an in-layer pair by the smaller of its two directed indices."""
    dl = layer.bonds[bi][2]
    l2 = l + dl
    if min(l, l2) <= 0 or max(l, l2) >= L + 1:
        return None
    mb = _mirror(layer, bi)
    if dl > 0:
        return (l, bi)
    if dl < 0:
        return (l2, mb)
    return (l, min(bi, mb))


def phonon_blocks(layer: DiamondLayer, masses, springs, lead_mass_l, lead_mass_r,
                  kfrac=(0.0, 0.0), alpha=48.5, beta=13.8, beta_bond=None):
    """Dynamical-matrix tdb blocks (THz^2) for a device of L layers.

    masses   (L, natom) device masses in amu
    springs  (L, nbond) per-bond scale factors.  Bond bi of device layer l
             (1-based l = 1..L) is the directed bond (i in layer l -> j in layer
             l+dl); a bond between two device layers is owned by the lower
             layer (dl=+1 entry), an in-layer pair by the smaller bond index.
             Bonds touching a lead layer keep scale 1, so the lead layers 0 and
             N+1 are exactly the pristine lead principal layers.
    kfrac    transverse k in units of the supercell reciprocal vectors
    beta_bond (L, nbond) optional per-bond angular constant (N/m) with the
             same ownership rule, replacing beta for device bonds
    """
    masses = np.asarray(masses, float)
    springs = np.asarray(springs, float)
    L = masses.shape[0]
    n, nd = layer.natom, layer.ndof
    ml = np.broadcast_to(np.asarray(lead_mass_l, float), (n,))
    mr = np.broadcast_to(np.asarray(lead_mass_r, float), (n,))
    phis = [layer.phi(e, alpha, beta) for (_, _, _, _, _, e) in layer.bonds]
    phases = [_phase((Rx, Ry), kfrac) for (_, _, _, Rx, Ry, _) in layer.bonds]
    beta_bond = None if beta_bond is None else np.asarray(beta_bond, float)

    def mass(l):
        return ml if l <= 0 else (mr if l >= L + 1 else masses[l - 1])

    def phi_s(l, bi):
        own = bond_owner(layer, L, l, bi)
        if own is None:
            return phis[bi]
        lo, bo = own
        ph = phis[bi] if beta_bond is None else layer.phi(layer.bonds[bi][5], alpha, beta_bond[lo - 1, bo])
        return ph * springs[lo - 1, bo]

    def D(l, l2):
        out = np.zeros((nd, nd), np.complex128)
        dl = l2 - l
        for bi, (i, j, d, _, _, _) in enumerate(layer.bonds):
            ph = phi_s(l, bi)
            if l == l2:
                out[3 * i:3 * i + 3, 3 * i:3 * i + 3] -= ph
            if d == dl:
                out[3 * i:3 * i + 3, 3 * j:3 * j + 3] += ph * phases[bi]
        w1 = 1.0 / np.sqrt(np.repeat(mass(l), 3))
        w2 = 1.0 / np.sqrt(np.repeat(mass(l2), 3))
        return out * w1[:, None] * w2[None, :] * THZ2_PER_N_M_AMU

    Haml = D(0, 0)
    Hamlv = D(-1, 0)                     # V_{-1,0}, pristine left lead
    Hamr = D(L + 1, L + 1)
    Hamrv = D(L + 2, L + 1)              # V_{N+2,N+1}, pristine right lead
    Hamd = np.stack([D(l, l) for l in range(1, L + 1)])
    Hamdv = np.stack([D(l, l + 1) for l in range(1, L)]) if L > 1 else np.zeros((0, nd, nd), np.complex128)
    blk = Blocks(Haml=Haml, Hamlv=Hamlv, Hamr=Hamr, Hamrv=Hamrv, Hamd=Hamd, Hamdv=Hamdv,
                 Hamld=D(0, 1), Hamrd=D(L + 1, L))
    _hermitize(blk)
    return blk


def _mirror(layer, bi):
    """Index of the bond that is the reverse of bond bi (j->i, opposite dl, -R)."""
    mir = getattr(layer, "_mirror_idx", None)
    if mir is None:
        idx = {(i, j, dl, Rx, Ry): k for k, (i, j, dl, Rx, Ry, e) in enumerate(layer.bonds)}
        mir = [idx[(j, i, -dl, -Rx, -Ry)] for (i, j, dl, Rx, Ry, e) in layer.bonds]
        layer._mirror_idx = mir
    return mir[bi]


def _hermitize(blk: Blocks):
    blk.Haml[...] = 0.5 * (blk.Haml + blk.Haml.conj().T)
    blk.Hamr[...] = 0.5 * (blk.Hamr + blk.Hamr.conj().T)
    for i in range(blk.L):
        blk.Hamd[i] = 0.5 * (blk.Hamd[i] + blk.Hamd[i].conj().T)


# ----------------------------------------------------------------- workloads

def si_ge_wire(nx=2, ny=2, L=20, ge_fraction=0.5, seed=0, bond_noise=0.05,
               alpha=48.5, beta=13.8):
    """Si-like wire with a random Si/Ge mass alloy and per-bond spring noise,
    pristine Si leads, no transverse k."""
    rng = np.random.default_rng(seed)
    layer = DiamondLayer(nx, ny, periodic=False)
    masses = np.where(rng.random((L, layer.natom)) < ge_fraction, MASS_GE, MASS_SI)
    springs = rng.uniform(1.0 - bond_noise, 1.0 + bond_noise, (L, layer.nbond))
    return phonon_blocks(layer, masses, springs, MASS_SI, MASS_SI, (0.0, 0.0), alpha, beta)


def si_ge_interface_film(nx=4, ny=2, L=40, kfrac=(0.0, 0.0), seed=0, iface_springs=None,
                         alpha=48.5, beta=13.8, n_iface_layers=2):
    """Si|Ge interface film: left half Si, right half Ge, Si and Ge leads,
    periodic in-plane.  iface_springs (optional) overrides beta-scales of the
    bonds starting in the n_iface_layers layers around the interface."""
    layer = DiamondLayer(nx, ny, periodic=True)
    masses = np.full((L, layer.natom), MASS_SI)
    masses[L // 2:] = MASS_GE
    springs = np.ones((L, layer.nbond))
    if iface_springs is not None:
        lo = L // 2 - n_iface_layers // 2
        springs[lo:lo + n_iface_layers] = np.asarray(iface_springs).reshape(n_iface_layers, layer.nbond)
    return phonon_blocks(layer, masses, springs, MASS_SI, MASS_GE, kfrac, alpha, beta)


def rough_interface_masses(natom, L=100, rough_layers=3, seed=0):
    """Masses (L, natom) of a Si|Ge film with a rough interface: layers below
    the interface Si, above it Ge, and `rough_layers` layers centred on it
    with each site Ge with probability (j + 1) / (rough_layers + 1) in the
    j-th of them (a graded, seeded alloy)."""
    rng = np.random.default_rng(seed)
    lo = L // 2 - rough_layers // 2
    masses = np.full((L, natom), MASS_SI)
    masses[lo + rough_layers:] = MASS_GE
    for j in range(rough_layers):
        f = (j + 1) / (rough_layers + 1)
        masses[lo + j] = np.where(rng.random(natom) < f, MASS_GE, MASS_SI)
    return masses


def rough_interface_film(nx=8, ny=4, L=100, rough_layers=3, kfrac=(0.0, 0.0), seed=0, masses=None,
                         alpha=48.5, beta=13.8):
    """Si thin film (nx x ny conventional cells in-plane, periodic) with a
    seeded rough Si/Ge interface of `rough_layers` layers in the middle of
    the device; Si left lead, Ge right lead.  masses (L, natom) overrides the
    seeded mass profile (the interface masses are the large mode-resolved
    workload's trainable parameters)."""
    layer = _layer_cache(nx, ny)
    if masses is None:
        masses = rough_interface_masses(layer.natom, L, rough_layers, seed)
    springs = np.ones((L, layer.nbond))
    return phonon_blocks(layer, masses, springs, MASS_SI, MASS_GE, kfrac, alpha, beta)


def chain_blocks(L=4, k=1.0, m=1.0, masses=None):
    """1-D monatomic chain (b=1), the closed-form case: T = 1 for 0 < E < 4k/m."""
    masses = np.full(L, m) if masses is None else np.asarray(masses, float)
    def d(mi, mj):
        return -k / math.sqrt(mi * mj)
    Haml = np.array([[2 * k / m]], np.complex128)
    Hamlv = np.array([[d(m, m)]], np.complex128)
    Hamd = np.array([[[2 * k / mi]] for mi in masses], np.complex128)
    Hamdv = np.array([[[d(masses[i], masses[i + 1])]] for i in range(L - 1)], np.complex128).reshape(L - 1, 1, 1)
    Hamld = np.array([[d(m, masses[0])]], np.complex128)
    Hamrd = np.array([[d(m, masses[-1])]], np.complex128)
    return Blocks(Haml=Haml, Hamlv=Hamlv, Hamr=Haml.copy(), Hamrv=Hamlv.conj().T.copy(),
                  Hamd=Hamd, Hamdv=Hamdv, Hamld=Hamld, Hamrd=Hamrd)


def random_blocks(nl, b, L, nr=None, seed=0, scale=1.0, hermitian=True):
    """Random Hermitian block-tridiagonal system with periodic leads (tests)."""
    rng = np.random.default_rng(seed)
    nr = nl if nr is None else nr

    def c(m, n):
        return (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))) * scale / math.sqrt(max(m, n))

    def h(n):
        x = c(n, n)
        return 0.5 * (x + x.conj().T) if hermitian else x
    Haml, Hamr = h(nl), h(nr)
    Vl, Vr = c(nl, nl), c(nr, nr)
    return Blocks(Haml=Haml, Hamlv=Vl, Hamr=Hamr, Hamrv=Vr.conj().T.copy(),
                  Hamd=np.stack([h(b) for _ in range(L)]),
                  Hamdv=np.stack([c(b, b) for _ in range(L - 1)]) if L > 1 else np.zeros((0, b, b), np.complex128),
                  Hamld=c(nl, b), Hamrd=c(nr, b))


# ------------------------------------------------------------ tight binding

def agnr_blocks(n_dimer=40, L=60, t=-2.7, hop_noise=0.02, onsite_w=0.5, seed=0):
    """Armchair graphene nanoribbon, p_z orbitals, n_dimer dimer lines wide.
    Principal layer = one armchair cell (2*n_dimer orbitals).  Device: hopping
    t*(1+U(-hop_noise, hop_noise)) per bond and on-site U(-W/2, W/2); pristine
    leads.  Energies in eV."""
    rng = np.random.default_rng(seed)
    nb = 2 * n_dimer
    intra, inter = _agnr_topology(n_dimer)
    H0 = np.zeros((nb, nb), np.complex128)
    for (i, j) in intra:
        H0[i, j] = H0[j, i] = t
    V = np.zeros((nb, nb), np.complex128)
    for (i, j) in inter:
        V[i, j] = t
    Hamd, Hamdv = [], []
    for l in range(L):
        Hl = np.zeros((nb, nb), np.complex128)
        for (i, j) in intra:
            v = t * (1.0 + rng.uniform(-hop_noise, hop_noise))
            Hl[i, j] = Hl[j, i] = v
        Hl[np.diag_indices(nb)] = rng.uniform(-onsite_w / 2, onsite_w / 2, nb)
        Hamd.append(Hl)
    for l in range(L - 1):
        Vl = np.zeros((nb, nb), np.complex128)
        for (i, j) in inter:
            Vl[i, j] = t * (1.0 + rng.uniform(-hop_noise, hop_noise))
        Hamdv.append(Vl)
    return Blocks(Haml=H0.copy(), Hamlv=V.copy(), Hamr=H0.copy(), Hamrv=V.conj().T.copy(),
                  Hamd=np.stack(Hamd), Hamdv=np.stack(Hamdv), Hamld=V.copy(), Hamrd=V.conj().T.copy())


def _agnr_topology(n):
    """Bonds of an armchair ribbon cell with n dimer lines.  The cell has 4 atom
    columns along x (transport) at x = 0, 0.5, 1.5, 2 (in a_cc) and n rows along
    y; sites are numbered row-major over columns [A0, B0, A1, B1] per pair of
    rows, giving 2n orbitals.  Returns (intra-cell bonds, bonds from this cell
    to the next one)."""
    # Build explicit coordinates for a zigzag-edged armchair ribbon.
    acc = 1.0
    sites = []
    for r in range(n):
        y = r * math.sqrt(3) / 2 * acc
        if r % 2 == 0:
            xs = [0.0, acc]
        else:
            xs = [1.5 * acc, 2.5 * acc]
        for x in xs:
            sites.append((x, y))
    sites = np.array(sites)
    period = 3.0 * acc
    intra, inter = [], []
    for i in range(len(sites)):
        for j in range(len(sites)):
            for shift, lst in ((0.0, intra), (period, inter)):
                d = np.linalg.norm(sites[j] + np.array([shift, 0.0]) - sites[i])
                if abs(d - acc) < 1e-6 and (shift > 0 or i < j):
                    lst.append((i, j))
    return intra, inter


# ------------------------------------------------------------ block-circulant leads

def circulant_matrix(cells, nbx, nby):
    """Full matrix of an nby x nbx supercell from its translation blocks.

    cells[(dx, dy)] couples cell (ix, iy) to cell ((ix+dx) % nbx, (iy+dy) % nby);
    cells are ordered y-major (index iy*nbx + ix), which is the nesting that
    negf/circular.f90 block-diagonalises (outer y, inner x)."""
    na = next(iter(cells.values())).shape[0]
    n = nbx * nby * na
    M = np.zeros((n, n), np.complex128)
    for iy in range(nby):
        for ix in range(nbx):
            for jy in range(nby):
                for jx in range(nbx):
                    d = ((jx - ix) % nbx, (jy - iy) % nby)
                    r, c = (iy * nbx + ix) * na, (jy * nbx + jx) * na
                    M[r:r + na, c:c + na] = cells[d]
    return M


def circulant_lead_blocks(na, nbx, nby, L, seed=0, scale=0.3, device_noise=0.1):
    """Random Hermitian system whose leads are nby x nbx supercells with
    block-circulant H00 and V (a lead at a transverse k commensurate with the
    supercell); the device is the lead plus a random Hermitian perturbation."""
    rng = np.random.default_rng(seed)

    def c(m, n):
        return (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))) * scale / math.sqrt(max(m, n))

    keys = [(dx, dy) for dx in range(nbx) for dy in range(nby)]
    K = circulant_matrix({d: c(na, na) for d in keys}, nbx, nby)
    H = 0.5 * (K + K.conj().T)
    V = circulant_matrix({d: c(na, na) for d in keys}, nbx, nby)
    n = H.shape[0]

    def pert():
        x = c(n, n) * device_noise
        return 0.5 * (x + x.conj().T)
    return Blocks(Haml=H.copy(), Hamlv=V.copy(), Hamr=H.copy(), Hamrv=V.conj().T.copy(),
                  Hamd=np.stack([H + pert() for _ in range(L)]),
                  Hamdv=np.stack([V + c(n, n) * device_noise for _ in range(L - 1)]) if L > 1
                  else np.zeros((0, n, n), np.complex128),
                  Hamld=V.copy(), Hamrd=V.conj().T.copy())


def si_ge_superlattice_film(n=1, period=2, L=10, kfrac=(0.0, 0.0), seed=0, bond_noise=0.02,
                            alpha=48.5, beta=13.8):
    """Si/Ge superlattice film: n x n conventional cells in-plane (b = 24 n^2),
    periodic in-plane, device layers alternating `period` Si and `period` Ge
    layers, seeded per-bond spring noise, pristine Si leads."""
    rng = np.random.default_rng(seed)
    layer = _layer_cache(n, n)
    masses = np.array([[MASS_SI if (l // period) % 2 == 0 else MASS_GE] * layer.natom for l in range(L)])
    springs = rng.uniform(1.0 - bond_noise, 1.0 + bond_noise, (L, layer.nbond))
    return phonon_blocks(layer, masses, springs, MASS_SI, MASS_SI, kfrac, alpha, beta)


_LAYERS = {}


def _layer_cache(nx, ny):
    key = (nx, ny)
    if key not in _LAYERS:
        _LAYERS[key] = DiamondLayer(nx, ny, periodic=True)
    return _LAYERS[key]


def superlattice_set(ndev=256, seed=0, sizes=(1, 2, 3, 4), periods=(2, 8), lengths=(10, 80)):
    """Seeded device descriptions of the uneven-batching workload:
    list of dict(n, b, period, L, seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for d in range(ndev):
        n = int(rng.choice(sizes))
        out.append(dict(n=n, b=24 * n * n, period=int(rng.integers(periods[0], periods[1] + 1)),
                        L=int(rng.integers(lengths[0], lengths[1] + 1)), seed=int(rng.integers(1 << 30))))
    return out


def mp_grid(nk1, nk2):
    """Monkhorst-Pack transverse grid centred as negf/config.f90:156-175."""
    return [(i / nk1 - 0.5 * (nk1 - 1) / nk1, j / nk2 - 0.5 * (nk2 - 1) / nk2)
            for i in range(nk1) for j in range(nk2)]
