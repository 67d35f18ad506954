"""Lead geometry of the Brillouin-zone unfolding (host side of
nt_modes_unfold_tdb_batched_dev).

The unfolding labels every propagating lead mode with the in-plane k-shift
of the transverse supercell it belongs to (negf/velocity_mat.f90:2020-2130
and mirrors); the GPU computes the weights, labels, unfolded k and the
specular / non-specular sums.  What it needs from the structure is fixed per
run and built here, as transverse_uc_setup (velocity_mat.f90:14-179) builds
it in the reference:

    transverse_cells   the transverse unit cell of one lead period (L_list),
                       the cell atom of every orbital (pidx) and the orbital
                       positions (pos_l / pos_r)
    UnfoldGeometry     both leads plus the reciprocal lattices; .desc() makes
                       the nt_unfold_desc of a set of k-points
    UnfoldGeometry.from_qe   the QE force-constant path's structure
                       (negf/phinfortran.f90:120-160: positions = the device
                       supercell, latvec_uc = the QE cell)
    PrimitiveCell      what the primitive-cell step (find_k_pc,
                       negf/dyn.f90:770-889; nt_unfold_pc_dev) needs: the
                       primitive cell's dynamical-matrix table and the
                       reciprocal lattices
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np

SPLIT_NAMES = ("transl_s", "transl_ns", "transr_s", "transr_ns", "refll_s", "refll_ns", "reflr_s", "reflr_ns")


class LeadCells(NamedTuple):
    L_list: np.ndarray      # 0-based atom indices of the transverse unit cell
    pidx: np.ndarray        # per orbital of the period: index into L_list
    pos: np.ndarray         # per orbital: atom position (n x 3)
    pos_uc: np.ndarray      # per orbital of the first image: position of its cell atom (n / nb x 3)


def transverse_cells(positions, latvec_uc, first: int, period: int, nbx: int, nby: int, tol: float,
                     ndims=3) -> LeadCells:
    """transverse_uc_setup for one lead period starting at 0-based atom
    `first`: atoms whose in-plane fractional coordinates lie in (-0.001,
    0.999) form L_list (velocity_mat.f90:39-68); every atom maps to the L_list
    atom it equals up to k a1 + l a2 (0 <= k < nbx, 0 <= l < nby) within tol,
    the last such L_list entry winning (:121-134; the reference uses tol 1e-3
    on the left, 1e-5 on the right).  ndims: degrees of freedom per atom, 3
    for phonons, or per-atom orbital counts of the whole structure (the
    tight-binding run's atom_orb_num, tb/velocity_mat.f90:14-174)."""
    positions = np.asarray(positions, float)[:, :3]
    latvec_uc = np.asarray(latvec_uc, float)
    atoms = positions[first:first + period]
    frac = atoms @ np.linalg.inv(latvec_uc)
    inside = (frac[:, 0] < 0.999) & (frac[:, 1] < 0.999) & (frac[:, 0] > -0.001) & (frac[:, 1] > -0.001)
    L_list = first + np.nonzero(inside)[0]
    if len(L_list) == 0:
        raise ValueError("no atom of the lead period lies in the transverse unit cell")
    shifts = np.array([k * latvec_uc[0] + l * latvec_uc[1] for k in range(nbx) for l in range(nby)])
    cell = positions[L_list]
    # dp[a, j, s] = atom a - cell atom j - shift s
    dp = atoms[:, None, None, :] - cell[None, :, None, :] - shifts[None, None, :, :]
    hit = (np.sqrt((dp * dp).sum(axis=-1)) < tol).any(axis=2)           # (atoms, cells)
    pidx_atom = np.empty(len(atoms), int)
    last = None
    for a in range(len(atoms)):
        js = np.nonzero(hit[a])[0]
        if len(js):
            last = int(js[-1])
        if last is None:
            raise ValueError(f"lead atom {first + a} matches no transverse-cell atom")
        pidx_atom[a] = last
    per_atom = np.full(len(positions), ndims, int) if np.isscalar(ndims) else np.asarray(ndims, int)
    counts = per_atom[first:first + period]
    pidx = np.repeat(pidx_atom, counts)
    pos = np.repeat(atoms, counts, axis=0)
    nb = nbx * nby
    n = len(pos)
    if n % nb or n // nb != per_atom[L_list].sum():
        raise ValueError(f"lead orbitals ({n}) are not {nb} images of the {len(L_list)}-atom transverse cell")
    pos_uc = positions[L_list[pidx[:n // nb]]]
    return LeadCells(L_list, pidx, pos, pos_uc)


class UnfoldGeometry:
    """Both leads' cells and the reciprocal lattices (rows = vectors)."""

    def __init__(self, n_bloch_x, n_bloch_y, recivec, reci_uc_l, reci_uc_r, left: LeadCells, right: LeadCells):
        self.n_bloch_x, self.n_bloch_y = int(n_bloch_x), int(n_bloch_y)
        self.recivec = np.asarray(recivec, float)
        self.reci_uc_l = np.asarray(reci_uc_l, float)
        self.reci_uc_r = np.asarray(reci_uc_r, float)
        self.left, self.right = left, right

    @classmethod
    def from_qe(cls, qe, cfg):
        """The QE tdb structure (phinfortran_amd.qe.read_qe_fc): positions =
        the device supercell (pos_all), latvec_uc = the QE cell; the second
        period of each lead (velocity_mat.f90:37, :72)."""
        pos = qe["pos_all"]
        N, per, buf = len(pos), cfg["period"], cfg["buffer"]
        nbx, nby = cfg["n_bloch_x"], cfg["n_bloch_y"]
        left = transverse_cells(pos, qe["cell"], buf + per, per, nbx, nby, 1e-3)
        right = transverse_cells(pos, qe["cell"], N - buf - 2 * per, per, nbx, nby, 1e-5)
        recivec = 2 * np.pi * np.linalg.inv(qe["cell_all"]).T
        reci_uc = 2 * np.pi * np.linalg.inv(qe["cell"]).T
        return cls(nbx, nby, recivec, reci_uc, reci_uc, left, right)

    def desc(self, torch, dev, ksc):
        """(nt_unfold_desc, tensors it points to) for k-points ksc (nk x 3,
        crystal coordinates of recivec)."""
        from . import _native as N
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
        keep = dict(ksc=t(np.atleast_2d(ksc)), pos_l=t(self.left.pos), pos_uc_l=t(self.left.pos_uc),
                    pos_r=t(self.right.pos), pos_uc_r=t(self.right.pos_uc))
        d = N.NtUnfoldDesc()
        d.n_bloch_x, d.n_bloch_y = self.n_bloch_x, self.n_bloch_y
        for name in ("recivec", "reci_uc_l", "reci_uc_r"):
            getattr(d, name)[:] = [float(x) for x in getattr(self, name).reshape(-1)]
        for name, v in keep.items():
            setattr(d, name, v.data_ptr())
        return d, keep


class PrimitiveCell:
    """What find_k_pc needs (negf/dyn.f90:770-889): the primitive cell's
    dynamical-matrix table (phinfortran_amd.qe.pc_dynamics_table), its
    reciprocal vectors, the lead unit cell's, and l = max(n_bloch_x, n_bloch_y)."""

    def __init__(self, table, n_pc, recivec_pc, reci_uc, n_bloch_max):
        self.table = table
        self.n_pc = int(n_pc)
        self.recivec_pc = np.asarray(recivec_pc, float)
        self.reci_uc = np.asarray(reci_uc, float)
        self.l = int(n_bloch_max)
        self._dev = {}

    @classmethod
    def from_qe(cls, qe, cfg, pc_fdf_path, species_mass=None):
        from . import qe as Q
        pc = Q.read_pc_fdf(pc_fdf_path, species_mass)
        table = Q.pc_dynamics_table(qe, pc["natoms_pc"], pc["mass"])
        reci_uc = 2 * np.pi * np.linalg.inv(qe["cell"]).T
        return cls(table, 3 * pc["natoms_pc"], pc["recivec"], reci_uc, max(cfg["n_bloch_x"], cfg["n_bloch_y"]))

    def device_table(self, torch, dev):
        key = str(dev)
        if key not in self._dev:
            self._dev[key] = tuple(torch.as_tensor(np.ascontiguousarray(x), device=dev) for x in self.table)
        return self._dev[key]


class TbPrimitiveCell:
    """What the tight-binding run's find_k_pc_l / _r needs
    (tb/hamiltonian.f90:666-916; nt_unfold_pc_tb_dev): the primitive cell's
    element-wise H table (tb.pc_table), its reciprocal vectors, the lead
    unit cell's (reci_uc_l / reci_uc_r) and l = max(n_bloch_x, n_bloch_y)."""

    tb = True

    def __init__(self, table, recivec_pc, reci_uc, n_bloch_max):
        self.table = table[:3]
        self.n_pc = int(table[3])
        self.recivec_pc = np.asarray(recivec_pc, float)
        self.reci_uc = np.asarray(reci_uc, float)
        self.l = int(n_bloch_max)
        self._dev = {}

    @classmethod
    def from_fdf(cls, P, pc_fdf_path, reci_uc, cfg, r_cutoff, theta_deg=0.0):
        from . import tb as TBM
        pc = TBM.read_pc_fdf(pc_fdf_path, theta_deg)
        table = TBM.pc_table(P, pc["positions_pc"], pc["species_pc"], pc["latvec_pc"], r_cutoff)
        return cls(table, pc["recivec_pc"], reci_uc, max(cfg["n_bloch_x"], cfg["n_bloch_y"]))

    def device_table(self, torch, dev):
        key = str(dev)
        if key not in self._dev:
            self._dev[key] = tuple(torch.as_tensor(np.ascontiguousarray(x), device=dev) for x in self.table)
        return self._dev[key]
