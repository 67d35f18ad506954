"""Differentiable thermal conductance of an interface film (adjoint RGF).

The objective is the phonon thermal conductance per unit area,

    G(T) = 1/(2 pi A nk) sum_k int hbar w T_k(w) dn_B/dT dw        (W m^-2 K^-1)

as examples/negf/si-si/condt.py:14-30 evaluates it from the reference's
transl output (trapezoid rule over w = 2 pi f, f in THz, E = f^2).  The
trainable parameters are the angular spring constants beta_b (N/m) of the
bonds owned by a window of device layers; the dynamical matrix is linear in
them, D = D_fixed + sum_b beta_b J_b, so each J_b is stored as a short list of
(block entry, value) pairs per k-point.

The forward pass writes the window blocks into the device-resident block
stack and runs the batched adjoint solve (nt_solve_tdb_adjoint_batched_dev),
which returns G and dG/dH for the window blocks together; backward contracts
dG/dH with the J_b (dG/dbeta_b = Re sum conj(dG/dH) J_b).
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np

from . import lattice as lt

HBAR = 1.0545718e-34
KB = 1.380649e-23
THZ_TO_RAD = 2.0e12 * math.pi


def conductance_weights(f_thz, temperature: float, area_m2: float, nk: int = 1):
    """Weights w_i with G = sum_i w_i T(f_i): the trapezoid rule of condt.py
    over w = 2 pi f (rad/s) applied to hbar w dn_B/dT / (2 pi A nk)."""
    f = np.asarray(f_thz, float)
    w = f * THZ_TO_RAD
    x = HBAR * w / (KB * temperature)
    with np.errstate(over="ignore"):
        ex = np.exp(x)
        dndT = np.where(x < 700, ex / (ex - 1.0) ** 2 * x / temperature, 0.0)
    integrand = HBAR * w * dndT / (2.0 * math.pi * area_m2 * nk)
    tw = np.zeros_like(w)
    dw = np.diff(w)
    tw[:-1] += 0.5 * dw
    tw[1:] += 0.5 * dw
    return tw * integrand


class InterfaceFilm:
    """Si|Ge interface film (nx x ny conventional cells in-plane, L layers)
    with trainable angular springs on the bonds owned by n_iface layers
    around the interface.  Holds the device-resident blocks of all k-points.

    kpts: transverse k-points (fractions of the supercell reciprocal vectors)
    """

    def __init__(self, engine, nx: int, ny: int, L: int, kpts: Sequence, f_thz, temperature: float = 300.0,
                 n_iface: int = 2, alpha: float = 48.5, beta: float = 13.8, beta0=None, params=None):
        import torch
        from . import transport as T
        self.engine = engine          # None: CPU-only set-up (block algebra, no solves)
        self.dev = torch.device("cuda", engine.device) if engine is not None else torch.device("cpu")
        self.layer = lt.DiamondLayer(nx, ny, periodic=True)
        self.L, self.alpha, self.beta = L, alpha, beta
        self.kpts = [tuple(k) for k in kpts]
        self.nk = len(self.kpts)
        self.masses = np.full((L, self.layer.natom), lt.MASS_SI)
        self.masses[L // 2:] = lt.MASS_GE
        # window: device layers (1-based) lo1 .. lo1 + n_iface - 1 own the trainable bonds
        self.lo1 = L // 2 - n_iface // 2 + 1
        self.n_iface = n_iface
        self.params = [(l, bi) for l in range(self.lo1, self.lo1 + n_iface)
                       for bi in range(self.layer.nbond) if lt.bond_owner(self.layer, L, l, bi) == (l, bi)]
        self.nparam = len(self.params)
        self.beta0 = np.full(self.nparam, beta) if beta0 is None else np.asarray(beta0, float)
        # engine window (0-based device layers): the owned bonds reach one layer up
        self.w_lo = self.lo1 - 1
        self.w_hi = min(self.lo1 - 1 + n_iface, L - 1)
        self.p = params if params is not None else T.params("phonon")
        self.f = np.asarray(f_thz, float)
        self.nf = len(self.f)
        a = lt.A_SI * 1e-10
        self.area = nx * a * ny * a
        w = conductance_weights(self.f, temperature, self.area, self.nk)
        E = self.f ** 2
        self.E = torch.as_tensor(np.tile(E, self.nk), dtype=torch.float64, device=self.dev)
        self.kidx = torch.as_tensor(np.repeat(np.arange(self.nk), self.nf), dtype=torch.int32, device=self.dev)
        mask = np.zeros(self.nf, np.int32)
        mask[0] = 1                                  # the reference's ie = 1 point (lesser.f90:802-807)
        self.mask = torch.as_tensor(np.tile(mask, self.nk), device=self.dev)
        self.weight = torch.as_tensor(np.tile(w, self.nk), dtype=torch.float64, device=self.dev)
        self._build()

    # -------------------------------------------------------------- blocks
    def _beta_bond(self, beta):
        bb = np.full((self.L, self.layer.nbond), self.beta)
        for v, (l, bi) in zip(beta, self.params):
            bb[l - 1, bi] = v
        return bb

    def blocks_numpy(self, beta, k: int):
        """tdb blocks at one k-point for the given beta (numpy; reference path)."""
        springs = np.ones((self.L, self.layer.nbond))
        return lt.phonon_blocks(self.layer, self.masses, springs, lt.MASS_SI, lt.MASS_GE, self.kpts[k],
                                self.alpha, self.beta, beta_bond=self._beta_bond(beta))

    def _build(self):
        """Upload blocks at beta = 0 on the trainable bonds (D_fixed) and the
        derivative entries J_b of the window blocks."""
        import torch
        zero = np.zeros(self.nparam)
        blocks = [self.blocks_numpy(zero, k) for k in range(self.nk)]
        self.dblocks = self.engine.upload_blocks(blocks) if self.engine is not None else None
        b = self.layer.ndof
        self.b = b
        nw = self.w_hi - self.w_lo + 1
        nc = min(self.w_hi, self.L - 2) - self.w_lo + 1
        self.nw, self.nc = nw, nc
        self.Hd_fixed = torch.as_tensor(np.stack([blk.Hamd[self.w_lo:self.w_hi + 1] for blk in blocks]),
                                        device=self.dev)
        self.Hdv_fixed = torch.as_tensor(np.stack([blk.Hamdv[self.w_lo:self.w_lo + nc] for blk in blocks]),
                                         device=self.dev)
        # derivative entries: flat index into (nk, nw, b, b) / (nk, nc, b, b), parameter, value
        di, dp, dv, vi, vp, vv = [], [], [], [], [], []
        pid = {pb: n for n, pb in enumerate(self.params)}
        lay = self.layer
        conv = lt.THZ2_PER_N_M_AMU
        for k in range(self.nk):
            for l in range(self.w_lo + 1, self.w_hi + 2):        # 1-based device layers of the window
                for bi, (i, j, d, Rx, Ry, e) in enumerate(lay.bonds):
                    own = lt.bond_owner(lay, self.L, l, bi)
                    if own is None or own not in pid:
                        continue
                    n = pid[own]
                    dphi = -(np.eye(3) - np.outer(e, e))          # d phi / d beta
                    mi = self.masses[l - 1, i]
                    # on-site block of atom i in layer l
                    if l - 1 <= self.w_hi:
                        base = ((k * nw + (l - 1 - self.w_lo)) * b)
                        for r in range(3):
                            for c in range(3):
                                di.append((base + 3 * i + r) * b + 3 * i + c)
                                dp.append(n)
                                dv.append(-dphi[r, c] * conv / mi)
                    if d == 0 and l - 1 <= self.w_hi:
                        mj = self.masses[l - 1, j]
                        ph = lt._phase((Rx, Ry), self.kpts[k])
                        base = ((k * nw + (l - 1 - self.w_lo)) * b)
                        for r in range(3):
                            for c in range(3):
                                di.append((base + 3 * i + r) * b + 3 * j + c)
                                dp.append(n)
                                dv.append(dphi[r, c] * ph * conv / math.sqrt(mi * mj))
                    if d == 1 and l - 1 - self.w_lo < nc and l <= self.L - 1:
                        mj = self.masses[l, j]
                        ph = lt._phase((Rx, Ry), self.kpts[k])
                        base = ((k * nc + (l - 1 - self.w_lo)) * b)
                        for r in range(3):
                            for c in range(3):
                                vi.append((base + 3 * i + r) * b + 3 * j + c)
                                vp.append(n)
                                vv.append(dphi[r, c] * ph * conv / math.sqrt(mi * mj))
        t = lambda x, dt: torch.as_tensor(np.asarray(x), dtype=dt, device=self.dev)
        self.Jd = (t(di, torch.int64), t(dp, torch.int64), t(dv, torch.complex128))
        self.Jv = (t(vi, torch.int64), t(vp, torch.int64), t(vv, torch.complex128))

    def window_blocks(self, beta):
        """Row-major window blocks (nk, nw, b, b), (nk, nc, b, b) at beta (torch)."""
        import torch
        beta = beta.to(torch.complex128)
        hd = self.Hd_fixed.reshape(-1).clone()
        hd.index_add_(0, self.Jd[0], beta[self.Jd[1]] * self.Jd[2])
        hv = self.Hdv_fixed.reshape(-1).clone()
        if self.nc > 0:
            hv.index_add_(0, self.Jv[0], beta[self.Jv[1]] * self.Jv[2])
        return hd.reshape(self.Hd_fixed.shape), hv.reshape(self.Hdv_fixed.shape)

    def write_window(self, hd, hv):
        """Store window blocks into the device block stack (ABI column-major)."""
        self.dblocks["Hamd"][:, self.w_lo:self.w_hi + 1] = hd.transpose(-1, -2)
        if self.nc > 0:
            self.dblocks["Hamdv"][:, self.w_lo:self.w_lo + self.nc] = hv.transpose(-1, -2)

    def beta_grad(self, gd, gv):
        """dW/dbeta from dW/dH of the window blocks (torch complex convention)."""
        import torch
        g = torch.zeros(self.nparam, dtype=torch.float64, device=self.dev)
        g.index_add_(0, self.Jd[1], (torch.conj(gd.reshape(-1)[self.Jd[0]]) * self.Jd[2]).real)
        if self.nc > 0:
            g.index_add_(0, self.Jv[1], (torch.conj(gv.reshape(-1)[self.Jv[0]]) * self.Jv[2]).real)
        return g

    # ----------------------------------------------------------- objective
    def conductance(self, beta, with_grad: bool = True):
        """(G, dG/dbeta or None, per-point T_LR) at beta (torch float64 tensor)."""
        hd, hv = self.window_blocks(beta.detach())
        self.write_window(hd, hv)
        if with_grad:
            r = self.engine.solve_tdb_adjoint(self.dblocks, self.E, self.p, self.weight, self.w_lo, self.w_hi,
                                              kidx=self.kidx, first_point_mask=self.mask)
            g = self.beta_grad(r["dHamd"], r["dHamdv"])
        else:
            r = self.engine.solve_tdb(self.dblocks, self.E, self.p, kidx=self.kidx, first_point_mask=self.mask)
            g = None
        if not bool((r["status"] == 0).all()):
            from ._native import TransportError
            bad = int((r["status"] != 0).sum())
            raise TransportError(int(r["status"][r["status"] != 0][0]), f"{bad} points failed")
        G = (self.weight * r["T_LR"]).sum()
        return G, g, r["T_LR"]


def _function():
    import torch

    class Conductance(torch.autograd.Function):
        """G(beta) with dG/dbeta from the adjoint RGF solve."""

        @staticmethod
        def forward(ctx, beta, film):
            need = ctx.needs_input_grad[0]
            G, g, _ = film.conductance(beta, with_grad=need)
            if need:
                ctx.save_for_backward(g)
            return G

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g, None

    return Conductance


_Conductance = None


def conductance(beta, film: InterfaceFilm):
    """Differentiable G(beta) (torch autograd)."""
    global _Conductance
    if _Conductance is None:
        _Conductance = _function()
    return _Conductance.apply(beta, film)


def maximize_conductance(film: InterfaceFilm, beta, steps: int = 50, lr: float = 0.2,
                         callback: Optional[callable] = None):
    """Adam ascent of G over the window springs; returns (beta, history of G)."""
    import torch
    x = beta.detach().clone().requires_grad_(True)
    opt = torch.optim.Adam([x], lr=lr)
    hist = []
    for it in range(steps):
        opt.zero_grad()
        G = conductance(x, film)
        (-G).backward()
        opt.step()
        hist.append(float(G.detach()))
        if callback:
            callback(it, hist[-1])
    return x.detach(), hist


class MassFilm:
    """Si thin film with a rough Si/Ge interface (lattice.rough_interface_film)
    whose interface masses are the trainable parameters.

    The dynamical matrix scales with the masses as D_xy = Phi_xy / sqrt(m_x m_y),
    so the window blocks at masses m follow from the blocks at the initial
    masses m0 by the diagonal rescaling D(m) = S D(m0) S, S = diag(sqrt(m0 / m))
    per degree of freedom, and

        dG/dm_a = -1/(2 m_a) sum over the window blocks of the entries in the
                  rows or columns of atom a of Re(conj(dG/dD) D),

    (twice for an entry in both, the on-site block of a), with dG/dD from the
    adjoint solve over device layers lo - 1 .. hi: the on-site blocks of the
    interface layers lo..hi, the couplings into them (lo - 1 -> lo) and out of
    them.  Masses of the layers lo..hi are trainable (rough_layers x natom).
    """

    def __init__(self, engine, nx: int, ny: int, L: int, kpts: Sequence, f_thz, temperature: float = 300.0,
                 rough_layers: int = 3, seed: int = 0, params=None):
        import torch
        from . import transport as T
        self.engine = engine
        self.dev = torch.device("cuda", engine.device)
        self.layer = lt._layer_cache(nx, ny)
        self.L = L
        self.kpts = [tuple(k) for k in kpts]
        self.nk = len(self.kpts)
        self.masses0 = lt.rough_interface_masses(self.layer.natom, L, rough_layers, seed)
        self.lo = L // 2 - rough_layers // 2             # 0-based first interface layer
        self.hi = self.lo + rough_layers - 1
        if self.lo < 1 or self.hi > L - 2:
            raise ValueError("the interface layers must not touch the leads")
        self.nparam = rough_layers * self.layer.natom
        self.p = params if params is not None else T.params("phonon")
        self.f = np.asarray(f_thz, float)
        self.nf = len(self.f)
        a = lt.A_SI * 1e-10
        self.area = nx * a * ny * a
        w = conductance_weights(self.f, temperature, self.area, self.nk)
        self.E = torch.as_tensor(np.tile(self.f ** 2, self.nk), dtype=torch.float64, device=self.dev)
        self.kidx = torch.as_tensor(np.repeat(np.arange(self.nk), self.nf), dtype=torch.int32, device=self.dev)
        mask = np.zeros(self.nf, np.int32)
        mask[0] = 1                                  # the reference's ie = 1 point (lesser.f90:802-807)
        self.mask = torch.as_tensor(np.tile(mask, self.nk), device=self.dev)
        self.weight = torch.as_tensor(np.tile(w, self.nk), dtype=torch.float64, device=self.dev)
        blocks = [lt.rough_interface_film(nx, ny, L, rough_layers, kf, masses=self.masses0) for kf in self.kpts]
        self.b = blocks[0].b
        self.dblocks = engine.upload_blocks(blocks)
        self.w_lo, self.w_hi = self.lo - 1, self.hi        # engine window (0-based layers)
        self.nw = self.w_hi - self.w_lo + 1
        self.nc = min(self.w_hi, L - 2) - self.w_lo + 1
        # row-major window blocks at the initial masses
        self.Hd0 = self.dblocks["Hamd"][:, self.w_lo:self.w_hi + 1].transpose(-1, -2).clone()
        self.Hdv0 = self.dblocks["Hamdv"][:, self.w_lo:self.w_lo + self.nc].transpose(-1, -2).clone()
        self.m0 = torch.as_tensor(self.masses0, dtype=torch.float64, device=self.dev)

    def initial(self):
        """the trainable masses at their seeded values (rough layers, flat)"""
        return self.m0[self.lo:self.hi + 1].reshape(-1).clone()

    def _layer_masses(self, m):
        full = self.m0.clone()
        full[self.lo:self.hi + 1] = m.reshape(self.hi - self.lo + 1, -1)
        return full

    def window_blocks(self, m):
        """row-major window blocks at trainable masses m"""
        full = self._layer_masses(m)
        r = torch_sqrt_ratio(self.m0, full)                           # (L, natom) sqrt(m0 / m)
        rd = r.repeat_interleave(3, dim=1)                            # per dof
        ls = slice(self.w_lo, self.w_hi + 1)
        hd = self.Hd0 * rd[ls][None, :, :, None] * rd[ls][None, :, None, :]
        lc = slice(self.w_lo, self.w_lo + self.nc)
        lc1 = slice(self.w_lo + 1, self.w_lo + 1 + self.nc)
        hv = self.Hdv0 * rd[lc][None, :, :, None] * rd[lc1][None, :, None, :]
        return hd, hv

    def write_window(self, hd, hv):
        self.dblocks["Hamd"][:, self.w_lo:self.w_hi + 1] = hd.transpose(-1, -2)
        if self.nc > 0:
            self.dblocks["Hamdv"][:, self.w_lo:self.w_lo + self.nc] = hv.transpose(-1, -2)

    def mass_grad(self, m, hd, hv, gd, gv):
        """dG/dm of the trainable masses from dG/dD of the window blocks"""
        import torch
        na = self.layer.natom
        full = self._layer_masses(m)
        acc = torch.zeros_like(full)                                   # (L, natom)
        Ad = (torch.conj(gd) * hd).real.sum(0)                         # (nw, b, b), summed over k
        rows = Ad.reshape(self.nw, na, 3, -1).sum(dim=(2, 3))
        cols = Ad.reshape(self.nw, -1, na, 3).sum(dim=(1, 3))
        acc[self.w_lo:self.w_hi + 1] += rows + cols
        if self.nc > 0:
            Av = (torch.conj(gv) * hv).real.sum(0)                     # (nc, b, b)
            acc[self.w_lo:self.w_lo + self.nc] += Av.reshape(self.nc, na, 3, -1).sum(dim=(2, 3))
            acc[self.w_lo + 1:self.w_lo + 1 + self.nc] += Av.reshape(self.nc, -1, na, 3).sum(dim=(1, 3))
        g = -acc / (2.0 * full)
        return g[self.lo:self.hi + 1].reshape(-1)

    def conductance(self, m, with_grad: bool = True):
        """(G, dG/dm or None, per-point T_LR) at trainable masses m (torch float64)"""
        hd, hv = self.window_blocks(m.detach())
        self.write_window(hd, hv)
        if with_grad:
            r = self.engine.solve_tdb_adjoint(self.dblocks, self.E, self.p, self.weight, self.w_lo, self.w_hi,
                                              kidx=self.kidx, first_point_mask=self.mask)
            g = self.mass_grad(m.detach(), hd, hv, r["dHamd"], r["dHamdv"])
        else:
            r = self.engine.solve_tdb(self.dblocks, self.E, self.p, kidx=self.kidx, first_point_mask=self.mask)
            g = None
        if not bool((r["status"] == 0).all()):
            from ._native import TransportError
            raise TransportError(int(r["status"][r["status"] != 0][0]), "mass film: points failed")
        return (self.weight * r["T_LR"]).sum(), g, r["T_LR"]


def torch_sqrt_ratio(m0, m):
    import torch
    return torch.sqrt(m0 / m)
