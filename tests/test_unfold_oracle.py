"""The unfolding oracle (oracle/unfold.py) against the closed form of exact
transverse Bloch states and against the reference's committed si-si
specular / non-specular tables and per-mode tables.

Inside a degenerate set of propagating modes (equal k_z) whose members
unfold to different in-plane k, the eigenvectors are any basis of the
subspace; the reference's labels then follow the basis its ZGEEV returned
and the split into specular and non-specular parts is that basis's.  Such
points are compared through s + ns (basis-invariant) only; every other point
is compared table by table and mode by mode."""
import math
import os
import sys

import numpy as np
import pytest

import oracle as O
from oracle import modes as M
from oracle import unfold as UF

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import si_si as S  # noqa: E402

SETS = ("L_adv", "L_ret", "R_ret", "R_adv")


def test_k_shift_and_move_in_ws2():
    ks = UF.k_shift(2, 3)
    assert ks.tolist() == [[0, 0, 0], [0, 1, 0], [0, 2, 0], [1, 0, 0], [1, 1, 0], [1, 2, 0]]
    a = np.array([[2.0, 0.0, 0.0], [0.5, 1.5, 0.0], [0.0, 0.0, 7.0]])
    for k in ([2.6, 0.1, 0.3], [-3.1, 2.05, -1.0], [0.2, -0.1, 0.0]):
        k = np.array(k)
        out = UF.move_in_ws2(k, a)
        # brute force over a wider range: the nearest in-plane lattice shift
        best = min((np.linalg.norm(k - np.array([i, j, 0.0]) @ a), i, j)
                   for i in range(-6, 7) for j in range(-6, 7))
        assert np.allclose(out, k - np.array([best[1], best[2], 0.0]) @ a)
        assert out[2] == k[2]


@pytest.fixture(scope="module")
def si():
    cfg = S.read_config()
    qe = S.read_qe_fc(cfg)
    pos = qe["pos_all"]
    N, per, buf = len(pos), cfg["period"], cfg["buffer"]
    nbx, nby = cfg["n_bloch_x"], cfg["n_bloch_y"]
    setl = UF.transverse_uc_setup(pos, qe["cell"], buf + per, per, nbx, nby, 1e-3)
    setr = UF.transverse_uc_setup(pos, qe["cell"], N - buf - 2 * per, per, nbx, nby, 1e-5)
    recivec = 2 * np.pi * np.linalg.inv(qe["cell_all"]).T
    reci_uc = 2 * np.pi * np.linalg.inv(qe["cell"]).T
    return dict(cfg=cfg, qe=qe, pos=pos, setl=setl, setr=setr, recivec=recivec, reci_uc=reci_uc,
                asm=S.TdbAssembler(qe, cfg), ks=S.k_grid(cfg), E=S.e_grid(cfg), nbx=nbx, nby=nby)


def test_transverse_setup_si_si(si):
    Ll, pidx, pos = si["setl"]
    # the four atoms of the (0, 0) transverse cell of the second lead period
    assert Ll.tolist() == [32, 33, 34, 35]
    # every orbital maps to the cell atom it translates to; images follow in blocks of 12
    assert pidx.tolist() == [a for a in range(4) for _ in range(3)] * 4
    assert pos.shape == (48, 3)


@pytest.mark.parametrize("ik", [10, 24])
def test_bloch_states_unfold_to_their_shift(si, ik):
    """U columns that are exact transverse Bloch states of the k-shift j
    (u_o exp(i G_j . r) on the orbital at r, G_j = k_shift(j) . recivec) get
    weight |lambda|^2 at j and zero elsewhere."""
    rng = np.random.default_rng(ik)
    Ll, pidx, pos = si["setl"]
    n, nb = pos.shape[0], si["nbx"] * si["nby"]
    nuc = n // nb
    shifts = UF.k_shift(si["nbx"], si["nby"])
    cols, expect = [], []
    for rep in range(3):
        for j in range(nb):
            G = shifts[j] @ si["recivec"]
            c = rng.standard_normal(nuc) + 1j * rng.standard_normal(nuc)
            u = np.array([c[o % nuc] * np.exp(1j * (G @ pos[o])) for o in range(n)])
            cols.append(u / np.linalg.norm(u))
            expect.append(j)
    U = np.zeros((n, n), complex)
    U[:, :len(cols)] = np.array(cols).T
    lam = np.exp(1j * rng.uniform(-3, 3, n)) * (1 + 1e-3 * rng.standard_normal(n))
    kz = np.full(n, UF.NOT_PROP)
    kz[:len(cols)] = rng.uniform(-0.5, 0.5, len(cols))
    w, bz, kxyz = UF.unfold_direction(U, lam, kz, si["ks"][ik], si["recivec"], si["reci_uc"], si["nbx"], si["nby"],
                                      si["pos"], Ll, pidx, pos)
    for i, j in enumerate(expect):
        ref = np.zeros(nb)
        ref[j] = abs(lam[i]) ** 2
        assert np.abs(w[i] - ref).max() < 1e-12, (i, w[i], ref)
        assert bz[i] == j + 1
        want = UF.move_in_ws2((si["ks"][ik] + shifts[j]) @ si["recivec"], si["reci_uc"])
        assert np.allclose(kxyz[i, :2], want[:2]) and kxyz[i, 2] == kz[i]
    assert (bz[len(cols):] == 0).all() and (kxyz[len(cols):] == UF.NOT_PROP).all()


def degenerate(m, tol=1e-6):
    """True when a mode set holds two propagating modes with equal k_z."""
    for key in SETS:
        kz = np.sort(m["k_z_" + key][m["k_z_" + key] < 99.9])
        if len(kz) > 1 and (np.diff(kz) < tol).any():
            return True
    return False


def unfold(si, m, ik):
    U = {k: m["U_" + k] for k in SETS}
    lam = {k: m["lam_" + k] for k in SETS}
    return UF.unfold_point(m, U, lam, si["ks"][ik], si["recivec"], si["reci_uc"], si["reci_uc"], si["nbx"],
                           si["nby"], si["pos"], si["setl"], si["setr"])


def md_rows(t, r, k_ft, k_f, k_b):
    """write_md's row numbers (negf/write_data.f90:113-170) before the clamp:
    for every propagating incident mode (k_b(1:3), r_s, r_t, t_s, t_t)."""
    rr = {i: (tt, s) for i, tt, s in UF.specular_rows(r, k_f, k_b)}
    out = []
    for i, tt, s in UF.specular_rows(t, k_ft, k_b):
        out.append(list(k_b[i]) + [rr[i][1], rr[i][0], s, tt])
    return np.array(out)


def compare_md(ours, gold):
    key = lambda a: np.lexsort((np.round(a[:, 1], 6), np.round(a[:, 0], 6), np.round(a[:, 2], 6)))
    a, g = ours[key(ours)], gold[key(gold)]
    assert np.abs(a[:, :3] - g[:, :3]).max() < 1e-7
    assert np.abs(a[:, 3:] - g[:, 3:]).max() < 2e-6


@pytest.mark.parametrize("ik", list(S.MODE_KPOINTS))
def test_si_si_split_and_mode_tables(si, ik):
    """transl_s/ns, transr_s/ns, refll_s/ns, reflr_s/ns of the committed
    tables; the unfolded (k_x, k_y, k_z) and the specular sums of every row
    of the trl (all MODE_KPOINTS) and mdl / mdr (MD_KPOINTS) tables."""
    cfg, E = si["cfg"], si["E"]
    Eg, split = S.golden_split()
    j = S.MODE_KPOINTS.index(ik)
    blk = si["asm"].blocks(si["ks"][ik])
    p = O.OrParams(cfg["eta"], cfg["eta0"], cfg["convergence"], 10.0, O.ETA_RELATIVE, 500)
    gold_trl = S.golden_rows("trl", ik)
    md = ik in S.MD_KPOINTS
    gold_md = {d: S.golden_rows(d, ik) for d in ("mdl", "mdr")} if md else None
    step = 1 if md else 3
    n_split = n_sum = n_rows = 0
    for ie in range(0, len(E), step):
        m = M.solve_point_modes(blk, E[ie], p, 5e-3, si["qe"]["cell"][2, 2], si["qe"]["cell"][2, 2],
                                first_point=(ie == 0))
        bz, kxyz, sp = unfold(si, m, ik)
        deg = degenerate(m)
        for base in ("transl", "transr", "refll", "reflr"):
            s, ns = sp[base + "_s"], sp[base + "_ns"]
            gs, gns = split[base + "_s"][ie, j], split[base + "_ns"][ie, j]
            scale = max(1.0, gs + gns)
            assert abs((s + ns) - (gs + gns)) < 2e-6 * scale, (ik, ie, base, s, ns, gs, gns)
            if not deg:
                assert abs(s - gs) < 2e-6 * scale and abs(ns - gns) < 2e-6 * scale, (ik, ie, base, s, ns, gs, gns)
        n_sum += 1
        n_split += not deg
        w = round(math.sqrt(E[ie]), 8)
        rows = gold_trl.get(w)
        inc = np.where(kxyz["L_adv"][:, 2] < 99.9)[0]
        if rows is None:
            assert len(inc) == 0
            continue
        if deg:
            continue
        ours = np.array([list(kxyz["L_adv"][i]) + [s, t] for i, t, s in
                         UF.specular_rows(m["trl"], kxyz["R_ret"], kxyz["L_adv"])])
        compare_md(ours, rows[:, [1, 2, 3, 10, 11]])
        n_rows += len(ours)
        if md:
            compare_md(md_rows(m["trl"], m["rll"], kxyz["R_ret"], kxyz["L_ret"], kxyz["L_adv"]),
                       gold_md["mdl"][w][:, [1, 2, 3, 10, 11, 12, 13]])
            compare_md(md_rows(m["tlr"], m["rrr"], kxyz["L_ret"], kxyz["R_ret"], kxyz["R_adv"]),
                       gold_md["mdr"][w][:, [1, 2, 3, 10, 11, 12, 13]])
    print(f"k-point {ik}: {n_sum} points (s + ns), {n_split} split-compared, {n_rows} trl rows")
    assert n_split >= n_sum // 2 and n_rows > 0
