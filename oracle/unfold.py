"""CPU restatement of the Brillouin-zone unfolding that labels lead modes
specular or non-specular (test infrastructure only: tests/ import it, the
product path never does).

    transverse_uc_setup   negf/velocity_mat.f90:14-179 (L_list, pidxL, pos_l)
    k_shift               negf/config.f90:176-183
    move_in_ws2           negf/util.f90:675-706
    unfold_direction      the weight loop, bz_idx and k_xyz of one mode set:
                          velocity_mat.f90:2035-2076 (L_adv), :2321-2355
                          (L_ret), :2578-2614 (R_ret), :2826-2862 (R_adv)
    split_sums            velocity_mat.f90:3032-3099 (transr_s/ns, transl_s/ns,
                          reflectr_s/ns, reflectl_s/ns)
    specular_rows         the per-incident-mode sums of write_specular and
                          write_md (negf/write_data.f90:3-175)

Every function follows the reference's loops one for one (1-based loop
variables become 0-based indices).  Pinned by: the si-si example's
committed transl_s/ns, transr_s/ns, refll_s/ns, reflr_s/ns tables and the
unfolded k_x, k_y columns and specular sums of its per-mode tables
(tests/test_unfold_oracle.py); by the closed form of an exact transverse
Bloch state (weight |lambda|^2 at its own k-shift, 0 at every other one);
and by s + ns = total for every table.
"""
import numpy as np

NOT_PROP = 100.0


def k_shift(nbx, nby):
    """negf/config.f90:176-183: n = (i-1) * n_bloch_y + j, shift (i-1, j-1, 0)."""
    ks = np.zeros((nbx * nby, 3))
    n = 0
    for i in range(nbx):
        for j in range(nby):
            ks[n, 0] = i
            ks[n, 1] = j
            n += 1
    return ks


def move_in_ws2(k_pos, a_vectors):
    """negf/util.f90:675-706: k minus the in-plane reciprocal vector
    n1 a1 + n2 a2 (n in -2..2) nearest to it; the first minimum in loop order."""
    dist = []
    k_all = []
    for i1 in range(-2, 3):
        for i2 in range(-2, 3):
            nd = np.array([i1, i2, 0.0])
            v = k_pos - nd @ a_vectors
            k_all.append(nd)
            dist.append(v @ v)
    i_min = int(np.argmin(dist))          # minloc: first occurrence
    return k_pos - k_all[i_min] @ a_vectors


def transverse_uc_setup(positions, latvec_uc, first, period, nbx, nby, tol, ndims=3):
    """negf/velocity_mat.f90:14-179 for one lead: the atoms of the period
    starting at 0-based atom `first` whose in-plane fractional coordinates
    lie in [0, 1) form the transverse unit cell (L_list); every atom of the
    period maps to the L_list entry it equals up to an in-plane translation
    (pidx, one per orbital, last match wins); pos: the atom position of each
    orbital.  Returns (L_list, pidx, pos) with 0-based indices."""
    inv = np.linalg.inv(latvec_uc)
    L_list = []
    for i in range(first, first + period):
        n = positions[i, :3] @ inv
        if n[0] < 0.999 and n[1] < 0.999 and n[0] > -0.001 and n[1] > -0.001:
            L_list.append(i)
    pos, pidx = [], []
    temp = None
    for i in range(first, first + period):
        p0 = positions[i, :3]
        for j, a in enumerate(L_list):
            p = positions[a, :3]
            for k in range(nbx):
                for l in range(nby):
                    dp = p0 - p - np.array([k, l, 0.0]) @ latvec_uc
                    if np.sqrt(dp @ dp) < tol:
                        temp = j
        for _ in range(ndims):
            pos.append(p0.copy())
            pidx.append(temp)
    return np.array(L_list), np.array(pidx), np.array(pos)


def unfold_direction(U, lam, kz, ksc, recivec, reci_uc, nbx, nby, positions, L_list, pidx, pos, ndims=3):
    """Weights, bz_idx (1-based, 0 for evanescent modes) and k_xyz (n x 3,
    100 for evanescent modes) of one mode set (velocity_mat.f90:2035-2076)."""
    n = U.shape[0]
    nb = nbx * nby
    shifts = k_shift(nbx, nby)
    natm_uc = len(L_list)
    kc = ksc @ recivec
    weight = np.zeros((n, nb))
    bz = np.zeros(n, dtype=int)
    kxyz = np.full((n, 3), NOT_PROP)
    for i in range(n):
        if not kz[i] < 100.0:
            continue
        for j in range(nb):
            kpoint = move_in_ws2((shifts[j] + ksc) @ recivec, reci_uc)
            temp_idx = 0
            for _ in range(natm_uc):
                idx = pidx[temp_idx]
                pos_uc = positions[L_list[idx], :3]
                for it in range(ndims):
                    s = 0.0 + 0.0j
                    for ir in range(nb):
                        o = temp_idx + it + ir * (n // nb)
                        p = pos[o]
                        s += np.exp(-1j * (-(kc @ p) + kpoint @ (p - pos_uc))) * U[o, i] * lam[i]
                    weight[i, j] += abs(s) ** 2 / nb
                temp_idx += ndims
        bz[i] = int(np.argmax(weight[i])) + 1      # maxloc: first maximum
        kxyz[i] = move_in_ws2((ksc + shifts[bz[i] - 1]) @ recivec, reci_uc)
        kxyz[i, 2] = kz[i]
    return weight, bz, kxyz


def _split(M, bz_row, bz_col):
    """sum of |M(row, col)|^2 with equal (s) and different (ns) positive labels."""
    s = ns = 0.0
    for c in range(M.shape[1]):
        for r in range(M.shape[0]):
            if bz_col[c] == bz_row[r] and bz_col[c] > 0:
                s += abs(M[r, c]) ** 2
            elif bz_col[c] != bz_row[r] and bz_row[r] > 0 and bz_col[c] > 0:
                ns += abs(M[r, c]) ** 2
    return s, ns


def split_sums(m, bz):
    """velocity_mat.f90:3032-3099.  m: t/r matrices (trl nr x nl, tlr nl x nr,
    rrr nr x nr, rll nl x nl, rows outgoing, columns incident); bz: the four
    label arrays.  Returns dict transl_s, transl_ns, transr_s, transr_ns,
    refll_s, refll_ns, reflr_s, reflr_ns (the names of the output tables)."""
    out = {}
    out["transr_s"], out["transr_ns"] = _split(m["trl"], bz["R_ret"], bz["L_adv"])
    out["transl_s"], out["transl_ns"] = _split(m["tlr"], bz["L_ret"], bz["R_adv"])
    out["reflr_s"], out["reflr_ns"] = _split(m["rrr"], bz["R_ret"], bz["R_adv"])
    out["refll_s"], out["refll_ns"] = _split(m["rll"], bz["L_ret"], bz["L_adv"])
    return out


def specular_rows(M, k_f, k_b):
    """write_specular's per-row sums (negf/write_data.f90:36-58) before the
    1e-60 clamp: for every incident mode i with k_b(i,3) < 99.9, (total,
    specular) = sums of |M(j,i)|^2 over outgoing modes j with k_f(j,3) < 99.9,
    specular where the in-plane k differ by less than 1e-5 (squared)."""
    rows = []
    for i in range(M.shape[1]):
        if not k_b[i, 2] < 99.9:
            continue
        t = s = 0.0
        for j in range(M.shape[0]):
            if k_f[j, 2] < 99.9:
                t += abs(M[j, i]) ** 2
                d = k_f[j].copy() - k_b[i]
                d[2] = 0.0
                if d @ d < 1.0e-5:
                    s += abs(M[j, i]) ** 2
        rows.append((i, t, s))
    return rows


def unfold_point(m, U, lam, ksc, recivec, reci_uc_l, reci_uc_r, nbx, nby, positions, setup_l, setup_r, ndims=3):
    """The unfolding of one (omega, k) point: m is oracle.modes.vel_mat_tdb's
    result, U / lam the four eigenvector matrices and eigenvalues (keys
    L_adv, L_ret, R_ret, R_adv), setup_l / setup_r transverse_uc_setup's
    results (with the same ndims: 3 for phonons, the orbitals per atom of a
    tight-binding structure whose atoms all carry that many, as in
    tb/velocity_mat.f90:524-557).  Returns bz, k_xyz (dicts by mode set) and
    the split sums."""
    kz = dict(L_adv=m["k_z_L_adv"], L_ret=m["k_z_L_ret"], R_ret=m["k_z_R_ret"], R_adv=m["k_z_R_adv"])
    bz, kxyz = {}, {}
    for key in ("L_adv", "L_ret", "R_ret", "R_adv"):
        Ll, pidx, pos = setup_l if key[0] == "L" else setup_r
        reci = reci_uc_l if key[0] == "L" else reci_uc_r
        _, bz[key], kxyz[key] = unfold_direction(U[key], lam[key], kz[key], ksc, recivec, reci, nbx, nby,
                                                 positions, Ll, pidx, pos, ndims=ndims)
    return bz, kxyz, split_sums(m, bz)
