"""CPU restatement of the mode-resolved transmission (test infrastructure only).

Follows vel_mat_tdb (negf/velocity_mat.f90:1692-3150), without the
Brillouin-zone unfolding that labels modes as specular or not (:2020-2130 and
its three mirrors).  path_mode != 1 is the branch the reference's example
runs (Gram-Schmidt on the propagating modes); path_mode = 1 adds the
degenerate-subspace rotation (:1884-1972 and mirrors): find_degenerate
(negf/util.f90:513-587), eigenH (ZHEEV on the upper triangle, numpy.linalg.eigh
with UPLO='U' here) and lstsq (ZGELSY, numpy.linalg.lstsq here).  The
LAPACK calls of the reference are the LAPACK calls here: ZGEEV through
numpy.linalg.eig (eigen, negf/util.f90:226-262) and ZGEES through
scipy.linalg.schur (sqrtm, negf/util.f90:359-454); the triangular square root,
Gram-Schmidt and log_fbz are restated line by line.

Pinned by: the si-si example's committed per-mode tables
(examples/negf/si-si/output/trl_specular_param_*.dat: incident k_z and
per-mode transmission of every (omega, k)), the closed form of the 1-D chain
(one mode, |t|^2 = T), sum_n T_n = Caroli T, and unitarity of the scattering
matrix of pristine leads (tests/test_modes_oracle.py).  Per-mode rows inside
a degenerate subspace depend on the eigenvector basis ZGEEV returns and are
compared only through basis-invariant sums.
"""
import numpy as np
import scipy.linalg as sl

NOT_PROP = 100.0     # k_z of an evanescent mode (velocity_mat.f90:1856)


def log_fbz(lam, az):
    """negf/util.f90:825-839"""
    gz = 2.0 * np.pi / az
    kz = np.mod(np.real(np.log(lam) / (1j * az)), gz)
    return kz - gz if kz > 0.5 * gz else kz


def gram_schmidt(U1):
    """negf/util.f90:800-823, including its coefficient
    dot_product(U1_i, U2_j) / dot_product(U2_j, U2_j) (conjugate on U1_i)."""
    U2 = np.array(U1, dtype=complex, copy=True)
    n = U1.shape[1]
    for i in range(1, n):
        for j in range(i):
            U2[:, i] = U2[:, i] - np.vdot(U1[:, i], U2[:, j]) / np.vdot(U2[:, j], U2[:, j]) * U2[:, j]
    for i in range(n):
        U2[:, i] = U2[:, i] / np.sqrt(np.vdot(U2[:, i], U2[:, i]))
    return U2


def sqrtm(A):
    """negf/util.f90:359-454: complex Schur form, triangular square root with
    U_ij = 0 where U_ii + U_jj is exactly zero, Q U Q^H."""
    S, Q = sl.schur(np.asarray(A, dtype=complex), output="complex")
    n = S.shape[0]
    U = np.zeros_like(S)
    for i in range(n):
        U[i, i] = np.sqrt(S[i, i])
    for sd in range(1, n):
        for i in range(n - sd):
            j = i + sd
            v = S[i, j] - U[i, i + 1:j] @ U[i + 1:j, j]
            den = U[i, i] + U[j, j]
            U[i, j] = v / den if den != 0 else 0.0
    return Q @ U @ Q.conj().T


def find_degenerate(lam, start, deg_all):
    """negf/util.f90:513-587 with its 1-based indices: the first pair i < j
    (i from start) with |lam_i - lam_j| < 1e-4 and ||lam_j| - 1| < 1e-3 gives
    the group of every index within 1e-4 of lam_j, unless one of them was
    already processed (then [0]).  Returns (group, next_start, found)."""
    ne = len(lam)
    found, itemp, edeg = 0, 0, None
    for i in range(start, ne):
        for j in range(i + 1, ne + 1):
            if abs(lam[i - 1] - lam[j - 1]) < 1.0e-4 and abs(abs(lam[j - 1]) - 1.0) < 1.0e-3:
                found, itemp, edeg = 1, i, lam[j - 1]
            if found:
                break
        if found:
            break
    group = [0]
    if found:
        group = [i for i in range(1, ne + 1) if abs(lam[i - 1] - edeg) < 1.0e-4]
        if any(g in deg_all for g in group):
            group = [0]
    return group, itemp + 1, found


def degenerate_rotation(E, lam, U, prop, H00, V):
    """The path_mode = 1 rotation of one mode set (velocity_mat.f90:1884-1978):
    each degenerate group is re-expressed, by least squares, in the
    eigenvectors of H(lam) = H00 + V / lam + V^+ lam (upper triangle) whose
    eigenvalue lies within 1e-4 of E, divided by lam; then every propagating
    column is normalised.  B has one column per group member: eigenvectors
    within 1e-4 beyond that count are dropped, and missing columns are taken
    from the eigenvectors next closest to E (the reference writes past, or
    leaves uninitialised, the columns of its B in those two cases; a member
    whose lam differs from the group's first by up to 1e-4 can sit that far
    from E in H(lam))."""
    n = len(lam)
    tf = np.eye(n, dtype=complex)
    start, found, deg_all = 1, 1, [0]
    while found == 1:
        group, nxt, found = find_degenerate(lam, start, deg_all)
        if len(group) > 1:
            deg_all = group[:] if deg_all == [0] else deg_all + group
            cols = [g - 1 for g in group]
            ld = lam[cols[0]]
            Hl = H00 + V / ld + V.conj().T * ld
            e, ev = np.linalg.eigh(Hl, UPLO="U")
            A = U[:, cols]
            sel = [i for i in range(n) if abs(E - e[i]) < 1.0e-4][:len(cols)]
            rest = sorted((i for i in range(n) if i not in sel), key=lambda i: (abs(E - e[i]), i))
            sel += rest[:len(cols) - len(sel)]
            B = np.zeros((n, len(cols)), complex)
            for c, i in enumerate(sel):
                B[:, c] = ev[:, i] / ld
            X = np.linalg.lstsq(A, B, rcond=None)[0]
            tf[np.ix_(cols, cols)] = X
        start = nxt
    U = U @ tf
    for i in prop:
        U[:, i] = U[:, i] / np.sqrt(np.vdot(U[:, i], U[:, i]))
    return U


def _modes(F, lambda_eta, az, sign, rot=None):
    """eigen + propagating selection + Gram-Schmidt of one lead direction
    (velocity_mat.f90:1839-1882 and mirrors): (lambda, U, k_z, n_prop).
    rot = (E, H00, V) applies the path_mode = 1 rotation."""
    lam, U = np.linalg.eig(F)
    kz = np.full(lam.shape[0], NOT_PROP)
    prop = [i for i in range(lam.shape[0]) if abs(abs(lam[i]) - 1.0) < lambda_eta]
    for i in prop:
        kz[i] = sign * log_fbz(lam[i], az)
    if prop:
        UL = gram_schmidt(U[:, prop])
        U = U.copy()
        U[:, prop] = UL
    if rot is not None:
        U = degenerate_rotation(rot[0], lam, U, prop, rot[1], rot[2])
    return lam, U, kz, len(prop)


def vel_mat_tdb(E, gl_adv, gr_ret, gl_plus, gr_minus, gml_adv, gmr_ret, G_D_RL, G_D_LR, G_L, G_R,
                Haml, Hamr, Hamlv, Hamrv, eta, eta0, lambda_eta, a_z_L, a_z_R, first_point=False, tb=False,
                path_mode=2):
    """Mode-resolved transmission and reflection of one (omega, k) point.

    tb=True follows tb/velocity_mat.f90: velocity matrices without the
    i a/(2 sqrt E) factor (:467-468), t and r with prefactor i (:1352-1370),
    constant lead broadening eta in Q (:1315).

    Returns dict with trl (nr x nl, left -> right), tlr, rrr, rll, the four
    k_z arrays (NOT_PROP for evanescent modes), the propagating counts, the
    per-mode sums t_mode_L[i] = sum_j |trl(j,i)|^2 and t_mode_R[j]
    (velocity_mat.f90:3036-3041)."""
    # velocity factors s_L, s_R and t/r prefactors (negf: i a/(2 sqrt E) and
    # 2 i sqrt(E/(a a')); tb: 1 and i)
    sL = 1.0 if tb else 1j * a_z_L / (2.0 * np.sqrt(E))
    sR = 1.0 if tb else 1j * a_z_R / (2.0 * np.sqrt(E))
    f = (lambda a, b: 1j) if tb else (lambda a, b: 1j * 2.0 * np.sqrt(E / a / b))

    # left advanced, incident from the left (:1835-2012)
    rl = (E, Haml, Hamlv) if path_mode == 1 else None
    rr = (E, Hamr, Hamrv) if path_mode == 1 else None
    lam_al, U_al, kz_al, n_al = _modes(gl_adv @ Hamlv, lambda_eta, a_z_L, -1.0, rl)
    V_al = -(U_al.conj().T @ gml_adv @ U_al) * sL
    inv_U_al = np.linalg.inv(U_al.conj().T)
    Vh_al = sqrtm(V_al)
    # left retarded, leaving to the left (:2131-2303)
    lam_rl, U_rl, kz_rl, n_rl = _modes(gl_adv.conj().T @ Hamlv, lambda_eta, a_z_L, -1.0, rl)
    V_rl = -(U_rl.conj().T @ gml_adv.conj().T @ U_rl) * sL
    inv_U_rl = np.linalg.inv(U_rl)
    Vh_rl = sqrtm(V_rl)
    # right retarded, leaving to the right (:2394-2576)
    lam_rr, U_rr, kz_rr, n_rr = _modes(gr_ret @ Hamrv, lambda_eta, a_z_R, 1.0, rr)
    V_rr = (U_rr.conj().T @ gmr_ret @ U_rr) * sR
    inv_U_rr = np.linalg.inv(U_rr)
    Vh_rr = sqrtm(V_rr)
    # right advanced, incident from the right (:2657-2824)
    lam_ar, U_ar, kz_ar, n_ar = _modes(gr_ret.conj().T @ Hamrv, lambda_eta, a_z_R, 1.0, rr)
    V_ar = (U_ar.conj().T @ gmr_ret.conj().T @ U_ar) * sR
    inv_U_ar = np.linalg.inv(U_ar.conj().T)
    Vh_ar = sqrtm(V_ar)
    # bulk GFs of the leads (:2893-2938), lead eta
    z = E + 1j * (eta if tb else E * eta + eta0)
    nl, nr = Haml.shape[0], Hamr.shape[0]
    Q_L = z * np.eye(nl) - Haml - Hamlv.conj().T @ gl_adv.conj().T @ Hamlv - Hamlv @ gl_plus @ Hamlv.conj().T
    Q_R = z * np.eye(nr) - Hamr - Hamrv.conj().T @ gr_ret @ Hamrv - Hamrv @ gr_minus.conj().T @ Hamrv.conj().T
    # t and r (:2945-2963)
    trl = f(a_z_R, a_z_L) * (Vh_rr @ (inv_U_rr @ (G_D_RL @ (inv_U_al @ Vh_al))))
    tlr = f(a_z_L, a_z_R) * (Vh_rl @ (inv_U_rl @ (G_D_LR @ (inv_U_ar @ Vh_ar))))
    rrr = f(a_z_R, a_z_R) * (Vh_rr @ (inv_U_rr @ ((G_R - np.linalg.inv(Q_R)) @ (inv_U_ar @ Vh_ar))))
    rll = f(a_z_L, a_z_L) * (Vh_rl @ (inv_U_rl @ ((G_L - np.linalg.inv(Q_L)) @ (inv_U_al @ Vh_al))))
    if first_point:                         # :2965-2970
        trl, tlr, rrr, rll = (np.zeros_like(x) for x in (trl, tlr, rrr, rll))
    a2 = lambda x: np.abs(x) ** 2
    return dict(trl=trl, tlr=tlr, rrr=rrr, rll=rll,
                k_z_L_adv=kz_al, k_z_L_ret=kz_rl, k_z_R_ret=kz_rr, k_z_R_adv=kz_ar,
                lam_L_adv=lam_al, lam_L_ret=lam_rl, lam_R_ret=lam_rr, lam_R_adv=lam_ar,
                U_L_adv=U_al, U_L_ret=U_rl, U_R_ret=U_rr, U_R_adv=U_ar,
                n_prop=dict(L_adv=n_al, L_ret=n_rl, R_ret=n_rr, R_adv=n_ar),
                t_mode_L=a2(trl).sum(axis=0), t_mode_R=a2(trl).sum(axis=1),
                T_modes_RL=float(a2(trl).sum()), T_modes_LR=float(a2(tlr).sum()),
                R_modes_R=float(a2(rrr).sum()), R_modes_L=float(a2(rll).sum()))


def solve_point_modes(blk, E, params, lambda_eta, a_z_L, a_z_R, first_point=False, tb=False, path_mode=2):
    """Full oracle chain for one point: surface GFs and get_G_full_tdb of the
    C oracle, then vel_mat_tdb with the driver's argument mapping
    (negf/phinfortran.f90:1238-1250)."""
    import oracle as O
    d = blk.as_dict() if hasattr(blk, "as_dict") else blk
    sg = O.surface_green(d["Haml"], d["Hamlv"], d["Hamr"], d["Hamrv"], E, params, want_plus_minus=True)
    r = O.rgf(list(d["Hamd"]), list(d["Hamdv"]), d["Hamld"], d["Hamrd"], d["Haml"], d["Hamr"],
              sg["gl_adv"].conj().T, sg["gr_ret"], -sg["gml_adv"], sg["gmr_ret"], sg["sl_adv"].conj().T,
              sg["sr_ret"], E, params)
    m = vel_mat_tdb(E, sg["gl_adv"], sg["gr_ret"], sg["gl_plus"], sg["gr_minus"], sg["gml_adv"], sg["gmr_ret"],
                    r["G_D_RL"], r["G_D_LR"], r["G_L"], r["G_R"], d["Haml"], d["Hamr"], d["Hamlv"], d["Hamrv"],
                    params.eta, params.eta0, lambda_eta, a_z_L, a_z_R, first_point, tb, path_mode)
    m["T_LR"], m["T_RL"] = r["T_LR"], r["T_RL"]
    return m
