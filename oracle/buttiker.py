"""CPU restatement of the Buttiker-probe (lesser Green's function) path of the
tdb solve (test infrastructure only: tests/ import it, the product path never
does).

    sigma_bp            set_sigma_bp_w (negf/lesser.f90:12-45): probe
                        self-energy -2 i omega / tau(T, E) of every layer and
                        its temperature derivative
    bose, dbose         negf/util.f90:987-1013
    simpson             negf/util.f90:1015-1043
    full_tdb            get_G_full_tdb (negf/lesser.f90:475-809) with the probe
                        self-energies (eyed) in every device layer of both
                        sweeps: T_LR, T_RL and the retarded blocks the lesser
                        sweep needs (GLeft, G_full, G_q_qp)
    lesser_tdb          get_G_lesser_tdb (negf/lesser.f90:1199-1503): the
                        left-connected and full lesser functions, their
                        temperature derivatives, probe fluxes, Jacobian,
                        contact fluxes
    integrate_jacobian, update_temperature, get_ph_temperature
                        negf/lesser.f90:1506-1623

Every function follows the reference's statements in order; matrices are
numpy arrays, layers 0-based.  The Jacobian is the reference's: its diagonal
differentiates the probe broadening gamma_i(T_i) in the A_i term only, not in
-Tr(gamma_i Gn_i) (negf/lesser.f90:1478-1489).  Pinned (tests/test_buttiker_oracle.py) by:
full_tdb's T against the C oracle without probes; equilibrium (all
temperatures equal) giving zero probe and contact fluxes; conservation of
the energy current over the probes and both contacts; the Jacobian against
central differences of the fluxes in the probe temperatures at fixed
retarded functions; Simpson's rule on polynomials.
"""
import numpy as np

PI = 3.141592653589793238
KB = 1.380649e-23
HBAR = 1.0545718e-34


def bose(freq, T):
    x = np.exp(HBAR * freq * 1.0e12 * 2 * PI / KB / T)
    return 1.0 / (x - 1.0)


def dbose(freq, T):
    x = np.exp(HBAR * freq * 1.0e12 * 2 * PI / KB / T)
    return x / (x - 1.0) ** 2 * HBAR * freq * 1.0e12 * 2 * PI / KB / T ** 2


def sigma_bp(E, T_layers, B, C, lifetime=True):
    """sigma_w (L x ne) and its T derivative (negf/lesser.f90:12-45); E: the
    energy grid Egrid(:, 1)."""
    E = np.asarray(E, float)
    L = len(T_layers)
    sw = np.zeros((L, len(E)), complex)
    dsw = np.zeros((L, len(E)), complex)
    for j in range(L):
        T = T_layers[j]
        tau = 1.0 / (B * E * T * np.exp(-C / T))
        sw[j] = -2.0j * np.sqrt(E) / tau
        dsw[j] = -2.0j * np.sqrt(E) * (B * E * np.exp(-C / T) * (1.0 + C / T))
    if not lifetime:
        sw[:] = 0.0
    return sw, dsw


def simpson(x, y):
    """negf/util.f90:1015-1043 (uniform spacing x(n) - x(n-1), odd n)."""
    n = len(x)
    if (n - 1) % 2:
        raise ValueError("ne should be an odd number")
    dx = x[n - 1] - x[n - 2]
    mm = (n - 1) // 2
    f = y[0] + 4.0 * sum(y[2 * i - 1] for i in range(1, mm + 1)) + 2.0 * sum(y[2 * i] for i in range(1, mm)) + y[n - 1]
    return f * dx / 3.0


def full_tdb(d, E, eta, eta0, gl, gr, gml, gmr, sl, sr, sig, first_point=False):
    """get_G_full_tdb with probes: d = tdb blocks (Hamd, Hamdv, Hamld, Hamrd,
    Haml, Hamr), gl..sr the retarded lead quantities as the batched solve
    passes them (gl = gl_adv^+, gml = -gml_adv, sl = sl_adv^+), sig (L): the
    probe self-energy of every layer at this point.  Also returns what
    vel_mat_tdb receives: G_D_LR, G_D_RL and the closing-layer G_L, G_R."""
    Hd, Hdv, Hld, Hrd = d["Hamd"], d["Hamdv"], d["Hamld"], d["Hamrd"]
    L, b = len(Hd), Hd[0].shape[0]
    nl, nr = gl.shape[0], gr.shape[0]
    z = E + 10.0 * (E * eta + eta0) * 1j
    I = np.eye(b)
    inv = np.linalg.inv
    GL = [None] * L
    G = inv(z * I - Hd[0] - Hld.conj().T @ gl @ Hld - sig[0] * I)
    GL[0] = G
    G0N = gl @ Hld @ G
    GN1 = None
    for i in range(1, L):
        V = Hdv[i - 1]
        Gi = inv(z * I - Hd[i] - V.conj().T @ G @ V - sig[i] * I)
        if i == L - 1:
            Vc1 = Hrd.conj().T
            GN1 = inv(z * I - Hd[i] - V.conj().T @ G @ V - sig[i] * I - Vc1 @ gr @ Vc1.conj().T)
        GL[i] = Gi
        G0N = G0N @ V @ Gi
        G = Gi
    # closing right lead layer N+1
    Vc = Hrd.conj().T
    GR = inv(z * np.eye(nr) - d["Hamr"] - sr - Vc.conj().T @ G @ Vc)
    G_full = [None] * L
    GL[L - 1] = GN1
    G_full[L - 1] = GN1
    G_DLR = G0N @ Vc @ GR
    Gq = [None] * (L - 1)
    for i in range(L - 2, -1, -1):
        V = Hdv[i]
        Gqp_q = G_full[i + 1] @ V.conj().T @ GL[i]
        Gq[i] = GL[i] @ V @ G_full[i + 1]
        G_full[i] = GL[i] + GL[i] @ V @ Gqp_q
    # right-to-left sweep for T_RL
    G = inv(z * I - Hd[L - 1] - Hrd.conj().T @ gr @ Hrd - sig[L - 1] * I)
    GN0 = gr @ Hrd @ G
    for i in range(L - 1, 0, -1):
        V = Hdv[i - 1].conj().T
        Gn = inv(z * I - Hd[i - 1] - V.conj().T @ G @ V - sig[i - 1] * I)
        GN0 = GN0 @ V @ Gn
        G = Gn
    GLl = inv(z * np.eye(nl) - d["Haml"] - sl - Hld @ G @ Hld.conj().T)
    G_DRL = GN0 @ Hld.conj().T @ GLl
    TLR = np.real(np.trace(G_DLR @ gmr @ G_DLR.conj().T @ gml))
    TRL = np.real(np.trace(G_DRL @ gml @ G_DRL.conj().T @ gmr))
    TLR = 0.0 if TLR < 1e-60 or first_point else TLR
    TRL = 0.0 if TRL < 1e-60 or first_point else TRL
    return dict(GLeft=GL, G_full=G_full, G_q_qp=Gq, T_LR=TLR, T_RL=TRL, G_D_LR=G_DLR, G_D_RL=G_DRL, G_L=GLl,
                G_R=GR)


def lesser_tdb(d, E, f, gl, gr, sig, dsig, T_layers, T_left, T_right, first_point=False):
    """get_G_lesser_tdb: f = full_tdb's result at this point.  Returns
    flux_i (L), jacobian (L x L), flux_left, flux_right, A_sum (L), Gn_sum (L)
    (complex traces, as the reference keeps them)."""
    Hdv, Hld, Hrd = d["Hamdv"], d["Hamld"], d["Hamrd"]
    GL, Gf, Gq = f["GLeft"], f["G_full"], f["G_q_qp"]
    L, b = len(GL), GL[0].shape[0]
    I = np.eye(b)
    H = lambda X: X.conj().T
    tr = np.trace
    freq = np.sqrt(E)
    eyed = [sig[i] * I for i in range(L)]
    deyed = [dsig[i] * I for i in range(L)]
    gam = [1j * (eyed[i] - H(eyed[i])) for i in range(L)]
    dgam = [1j * (deyed[i] - H(deyed[i])) for i in range(L)]
    A = [1j * (Gf[i] - H(Gf[i])) for i in range(L)]
    sigma_first = H(Hld) @ gl @ Hld
    gamma_first = 1j * (sigma_first - H(sigma_first))
    flux_left = tr(gamma_first @ A[0]) * bose(freq, T_left)
    GnL = [None] * L
    dGnL = [[np.zeros((b, b), complex) for _ in range(L)] for _ in range(L)]
    sin0 = gamma_first * bose(freq, T_left) + gam[0] * bose(freq, T_layers[0])
    GnL[0] = GL[0] @ sin0 @ H(GL[0])
    ds = gam[0] * dbose(freq, T_layers[0]) + dgam[0] * bose(freq, T_layers[0])
    dGnL[0][0] = GL[0] @ ds @ H(GL[0])
    for i in range(1, L - 1):
        V = Hdv[i - 1]
        sprev = H(V) @ GnL[i - 1] @ V
        sin = gam[i] * bose(freq, T_layers[i])
        GnL[i] = GL[i] @ (sin + sprev) @ H(GL[i])
        for j in range(i):
            ds = H(V) @ dGnL[i - 1][j] @ V
            dGnL[i][j] = GL[i] @ ds @ H(GL[i])
        ds = gam[i] * dbose(freq, T_layers[i]) + dgam[i] * bose(freq, T_layers[i])
        dGnL[i][i] = GL[i] @ ds @ H(GL[i])
    # last layer
    V = Hdv[L - 2]
    sprev = H(V) @ GnL[L - 2] @ V
    Vc1 = H(Hrd)
    sigma_last = Vc1 @ gr @ H(Vc1)
    gamma_last = 1j * (sigma_last - H(sigma_last))
    flux_right = tr(gamma_last @ A[L - 1]) * bose(freq, T_right)
    sin = gam[L - 1] * bose(freq, T_layers[L - 1]) + gamma_last * bose(freq, T_right)
    GnL[L - 1] = GL[L - 1] @ (sin + sprev) @ H(GL[L - 1])
    for j in range(L - 1):
        ds = H(V) @ dGnL[L - 2][j] @ V
        dGnL[L - 1][j] = GL[L - 1] @ ds @ H(GL[L - 1])
    ds = gam[L - 1] * dbose(freq, T_layers[L - 1]) + dgam[L - 1] * bose(freq, T_layers[L - 1])
    dGnL[L - 1][L - 1] = GL[L - 1] @ ds @ H(GL[L - 1])
    # full lesser function
    Gn = [None] * L
    dGn = [[None] * L for _ in range(L)]
    Gn[L - 1] = GnL[L - 1]
    flux_right = (flux_right - tr(gamma_last @ Gn[L - 1])) * freq
    for j in range(L):
        dGn[L - 1][j] = dGnL[L - 1][j]
    for i in range(L - 2, -1, -1):
        V = Hdv[i]
        Gn[i] = (GnL[i] + GL[i] @ V @ Gn[i + 1] @ H(V) @ H(GL[i]) + GnL[i] @ V @ H(Gq[i])
                 + Gq[i] @ H(V) @ GnL[i])
        for j in range(L):
            dGn[i][j] = (dGnL[i][j] + GL[i] @ V @ dGn[i + 1][j] @ H(V) @ H(GL[i]) + dGnL[i][j] @ V @ H(Gq[i])
                         + Gq[i] @ H(V) @ dGnL[i][j])
    flux_left = (flux_left - tr(gamma_first @ Gn[0])) * freq
    if first_point:
        Gn = [np.zeros_like(x) for x in Gn]
        dGn = [[np.zeros_like(x) for x in row] for row in dGn]
        A = [np.zeros_like(x) for x in A]
    flux = np.zeros(L, complex)
    jac = np.zeros((L, L), complex)
    for i in range(L):
        flux[i] = -tr(gam[i] @ Gn[i]) + tr(gam[i] @ A[i]) * bose(freq, T_layers[i])
        for j in range(L):
            jac[i, j] = -tr(gam[i] @ dGn[i][j])
            if i == j:
                jac[i, j] += (tr(gam[i] @ A[i] * dbose(freq, T_layers[j]))
                              + tr(dgam[i] @ A[i] * bose(freq, T_layers[j])))
    A_sum = np.array([tr(A[i]) for i in range(L)])
    Gn_sum = np.array([tr(Gn[i]) for i in range(L)])
    return dict(flux=flux, jac=jac, flux_left=flux_left, flux_right=flux_right, A_sum=A_sum, Gn_sum=Gn_sum)


def integrate_jacobian(jac_all, flux_all, E):
    """negf/lesser.f90:1506-1529: jac_all (nk, ne, L, L), flux_all (nk, ne, L)."""
    nk = flux_all.shape[0]
    w = np.sqrt(E)
    L = flux_all.shape[2]
    flux = np.array([simpson(w, flux_all[:, :, i].sum(axis=0) * w / nk) for i in range(L)])
    jac = np.array([[simpson(w, jac_all[:, :, i, j].sum(axis=0) * w / nk) for j in range(L)] for i in range(L)])
    return jac, flux


def update_temperature(T_layers, jac, flux):
    """negf/lesser.f90:1531-1557: one Newton step; all temperatures reset to 1
    when one turns negative.  Returns (new temperatures, rms change)."""
    T_old = np.array(T_layers, float)
    T_new = T_old - np.real(np.linalg.inv(jac) @ flux)
    if T_new.min() < 0.0:
        T_new[:] = 1.0
    return T_new, float(np.sqrt(((T_new - T_old) ** 2).sum() / len(T_new)))


def get_ph_temperature(A_sum, Gn_sum, E, T_ph, max_steps=200):
    """negf/lesser.f90:1588-1623: A_sum, Gn_sum (ne, L); Newton per layer until
    |dT| <= 1e-4 (bounded here)."""
    w = np.sqrt(E)
    T_ph = np.array(T_ph, float)
    for nn in range(A_sum.shape[1]):
        told = -100.0
        for _ in range(max_steps):
            if not abs(told - T_ph[nn]) > 1.0e-4:
                break
            told = T_ph[nn]
            dfl = np.array([dbose(x, T_ph[nn]) for x in w])
            fl = np.array([bose(x, T_ph[nn]) for x in w])
            jdiff = np.real(simpson(w, -dfl * E * A_sum[:, nn]))
            diff = np.real(simpson(w, Gn_sum[:, nn] * E - fl * E * A_sum[:, nn]))
            T_ph[nn] = told - diff / jdiff
    return T_ph
