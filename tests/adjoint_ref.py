"""Dense reference of T_LR and of its gradient with respect to the device
blocks, for small systems (test helper).

The left-to-right RGF sweep of get_G_full_tdb (negf/lesser.f90:556-651)
inverts, block by block, the matrix M of the finite system made of layer 0
(closed by the left surface GF gl), the device layers 1..N and layer N+1
(closed by Sigma_R), all but layer 0 at the device broadening.  Here M is
built whole and inverted densely:
    T_LR = Re Tr[G_{0,N+1} Gamma_R G_{0,N+1}^+ Gamma_L]          (lesser.f90:787-792)
    dT   = 2 Re Tr[dH X],  X = G Gamma_R G^+ Gamma_L G   (Gamma's embedded)
The gradient is returned in torch's convention (dW/dRe + i dW/dIm): 2 X_{ll}^+
for Hamd_l and 2 (X_{l+1,l}^+ + X_{l,l+1}) for Hamdv_l.
"""
import numpy as np

import oracle as O


def eta_of(E, op):
    return E * op.eta + op.eta0 if op.eta_mode == 0 else op.eta


def system(blk, E, op):
    """(M, offsets, lead quantities) of the left-to-right system at E."""
    d = blk.as_dict() if hasattr(blk, "as_dict") else blk
    sg = O.surface_green(d["Haml"], d["Hamlv"], d["Hamr"], d["Hamrv"], E, op, want_plus_minus=False)
    gl = sg["gl_adv"].conj().T
    gam_l = -sg["gml_adv"]
    gam_r = sg["gmr_ret"]
    sr = sg["sr_ret"]
    Hd, Hdv, Hld, Hrd, Hr = d["Hamd"], d["Hamdv"], d["Hamld"], d["Hamrd"], d["Hamr"]
    L, b = Hd.shape[0], Hd.shape[1]
    nl, nr = gl.shape[0], Hr.shape[0]
    off = [0, nl] + [nl + b * (i + 1) for i in range(L)]
    n = nl + L * b + nr
    zp = E + 1j * op.device_eta_factor * eta_of(E, op)
    M = np.zeros((n, n), complex)
    M[:nl, :nl] = np.linalg.inv(gl)
    for i in range(L):
        s = off[1 + i]
        M[s:s + b, s:s + b] = zp * np.eye(b) - Hd[i]
        if i + 1 < L:
            M[s:s + b, s + b:s + 2 * b] = -Hdv[i]
            M[s + b:s + 2 * b, s:s + b] = -Hdv[i].conj().T
    M[:nl, nl:nl + b] = -Hld
    M[nl:nl + b, :nl] = -Hld.conj().T
    sN = off[L]
    sR = off[L + 1]
    M[sR:, sR:] = zp * np.eye(nr) - Hr - sr
    M[sN:sN + b, sR:] = -Hrd.conj().T
    M[sR:, sN:sN + b] = -Hrd
    return M, off, gam_l, gam_r


def T_and_grad(blk, E, op, lo, hi):
    """T_LR and (dHamd[lo..hi], dHamdv[lo..min(hi, L-2)]) at one energy."""
    M, off, gam_l, gam_r = system(blk, E, op)
    d = blk.as_dict() if hasattr(blk, "as_dict") else blk
    L, b = d["Hamd"].shape[0], d["Hamd"].shape[1]
    nl = gam_l.shape[0]
    G = np.linalg.inv(M)
    sR = off[L + 1]
    G0R = G[:nl, sR:]
    T = float(np.real(np.trace(G0R @ gam_r @ G0R.conj().T @ gam_l)))
    A = gam_r @ G0R.conj().T @ gam_l
    X = G[:, sR:] @ A @ G[:nl, :]

    def blkX(i, j):   # device layers i, j (0-based)
        si, sj = off[1 + i], off[1 + j]
        return X[si:si + b, sj:sj + b]
    gd = np.stack([2 * blkX(l, l).conj().T for l in range(lo, hi + 1)])
    hc = min(hi, L - 2)
    gv = np.stack([2 * (blkX(l + 1, l).conj().T + blkX(l, l + 1)) for l in range(lo, hc + 1)]) \
        if hc >= lo else np.zeros((0, b, b), complex)
    return T, gd, gv
