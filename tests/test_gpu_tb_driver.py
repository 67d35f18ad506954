"""The tight-binding transport run (driver.run_tb: tb/infortran.f90:77-630)
on the GPU against the oracle chain at every (E, k) point of a Si/Ge
superlattice run: oracle gen_ham at the Wigner-Seitz-shifted k, the
reference's layer partition and slices, the C oracle's decimation and RGF.
The batched and the point-by-point solves agree; a pristine Si run gives
integer channel counts; the output files have the reference's shape."""
import os

import numpy as np
import pytest

import oracle as O
from oracle import tb_ham as TB
from phinfortran_amd import dense as D
from phinfortran_amd import driver
from phinfortran_amd import tb
from tests import tb_run_inputs as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Tm():
    from phinfortran_amd import transport
    return transport


def oracle_T(Tm, P, lat, pos, sp, k, E, eta, layers, nl, nr):
    H = TB.gen_ham(P, pos, sp, lat, k, 2.6)
    s = D.slices(H, layers, nl, nr, 0, 0, variant="tb")
    p = Tm.params("tb", eta=eta, convergence=1e-10)
    op = O.OrParams(p.eta, p.eta0, p.convergence, p.device_eta_factor, p.eta_mode, p.max_iter)
    r = O.surface_green(s.Haml, s.Hamlv, s.Hamr, s.Hamrv, E, op, want_plus_minus=False)
    o = O.rgf(s.Hd, s.Hdv, s.Hld, s.Hrd, s.Haml, s.Hamr, r["gl_adv"].conj().T, r["gr_ret"], -r["gml_adv"],
              r["gmr_ret"], r["sl_adv"].conj().T, r["sr_ret"], E, op, Hdv_back=s.Hdv_back, Hrd_fwd=s.Hrd_fwd)
    oracle_T.sdos = (np.trace(r["gl_adv"]).imag / np.pi, -np.trace(r["gr_ret"]).imag / np.pi)
    oracle_T.ldos = O.get_ldos(s.Hd, s.Hdv, s.Hld, s.Hrd, r["gl_adv"].conj().T, r["gr_ret"], E, op,
                               Hdv_back=s.Hdv_back, Hrd_fwd=s.Hrd_fwd)
    return o["T_LR"], o["T_RL"]


def test_run_tb_matches_oracle(engine, Tm, tmp_path):
    cfg, lat, pos, sp = R.write_run(str(tmp_path))
    print("run_tb batched", flush=True)
    r = driver.run_tb(cfg, engine=engine)
    assert r["norb"] == 320 and r["nl"] == r["nr"] == 40 and len(set(r["layer_list"])) == 1
    assert (r["status"] == 0).all()
    P = tb.read_param(R.PARAM)
    for ik, k in enumerate(r["k"]):
        for ie, E in enumerate(r["E"]):
            tl, tr = oracle_T(Tm, P, lat, pos, sp, k, E, 1e-5, r["layer_list"], 40, 40)
            assert abs(r["transl"][ie, ik] - tl) <= 1e-8 * max(1.0, tl), (ik, E)
            assert abs(r["transr"][ie, ik] - tr) <= 1e-8 * max(1.0, tr), (ik, E)
            sl_, sr_ = oracle_T.sdos                     # surface DOS: +-Im Tr g / pi, no sqrt(E)
            assert abs(r["sdos_l"][ie, ik] - sl_) <= 1e-8 * max(1.0, abs(sl_)), (ik, E)
            assert abs(r["sdos_r"][ie, ik] - sr_) <= 1e-8 * max(1.0, abs(sr_)), (ik, E)
            lo, do = oracle_T.ldos                       # get_ldos: layer-resolved device DOS
            assert np.abs(r["ldos"][ik, ie] - lo).max() <= 1e-8 * max(1.0, np.abs(lo).max()), (ik, E)
            assert abs(r["ddos"][ie, ik] - do) <= 1e-8 * max(1.0, abs(do)), (ik, E)
    assert (r["sdos_l"] > 0).any() and (r["sdos_l"] >= -1e-8).all()
    assert (r["ddos"] > 0).any() and (r["ldos"] >= -1e-8).all()
    assert r["transl"].max() > 0.5 and r["transl"][2].max() < 1e-10            # E = 0.35 eV lies in the gap
    out = os.path.join(str(tmp_path), "out")
    rows = open(os.path.join(out, "transl_tot.dat")).read().splitlines()
    assert len(rows) == 5 and all(len(x) == 20 * 5 for x in rows)
    assert float(rows[3][20:40]) == pytest.approx(r["transl"][3, 0], rel=1e-9)
    assert len(open(os.path.join(out, "positions_all.dat")).read().splitlines()) == 320
    srows = open(os.path.join(out, "sdosl.dat")).read().splitlines()
    assert len(srows) == 5 and all(len(x) == 20 * 4 + 1 for x in srows)
    assert float(srows[0][:20]) == pytest.approx(r["sdos_l"][0, 0], abs=1e-9)
    assert open(os.path.join(out, "nkpoints.dat")).read() == "        4.0000000000\n"
    drows = open(os.path.join(out, "ddos.dat")).read().splitlines()
    assert len(drows) == 5 and all(len(x) == 20 * 4 + 1 for x in drows)
    assert float(drows[1][20:40]) == pytest.approx(r["ddos"][1, 1], abs=1e-9)
    nlay, maxlen = r["ldos"].shape[2:]
    lrows = open(os.path.join(out, "ldos_000002.dat")).read().splitlines()
    assert len(lrows) == 5 and all(len(x) == 20 * (1 + nlay * maxlen) + 1 for x in lrows)
    assert float(lrows[3][:20]) == pytest.approx(r["E"][3], abs=1e-9)
    assert float(lrows[3][20 * (1 + maxlen + 2):20 * (2 + maxlen + 2)]) == pytest.approx(r["ldos"][1, 3, 1, 2],
                                                                                       abs=1e-9)


@pytest.mark.parametrize("cells,buffer", [((1, 1, 8), 0), ((2, 1, 8), 0), ((1, 1, 10), 4)])
def test_run_tb_pointwise_equals_batched(engine, tmp_path, cells, buffer):
    """equal one-atom layers, unequal layers [90, 40, 80, 110] padded to 110
    in the batched solve, and one lead cell of buffer on each side"""
    cfg, *_ = R.write_run(str(tmp_path), nk=(1, 2), ne=3, cells=cells, buffer=buffer)
    print("run_tb batched vs point by point", cells, flush=True)
    a = driver.run_tb(cfg, engine=engine, output_dir=str(tmp_path / "a"))
    b = driver.run_tb(cfg, engine=engine, output_dir=str(tmp_path / "b"), batched=False)
    if cells == (2, 1, 8):
        assert a["layer_list"] == [90, 40, 80, 110]
    if buffer:
        assert sum(a["layer_list"]) == a["norb"] - 2 * 40 - 2 * 2 * 40
    assert (a["status"] == 0).all() and a["transl"].max() > 0.5
    assert np.abs(a["transl"] - b["transl"]).max() <= 1e-9 * max(1.0, np.abs(a["transl"]).max())
    assert np.abs(a["transr"] - b["transr"]).max() <= 1e-9 * max(1.0, np.abs(a["transr"]).max())
    assert np.abs(a["sdos_l"] - b["sdos_l"]).max() <= 1e-9 * max(1.0, np.abs(a["sdos_l"]).max())
    assert np.abs(a["sdos_r"] - b["sdos_r"]).max() <= 1e-9 * max(1.0, np.abs(a["sdos_r"]).max())
    # device DOS: batched (padded layers) against the per-point get_ldos
    assert np.abs(a["ldos"] - b["ldos"]).max() <= 1e-9 * max(1.0, np.abs(b["ldos"]).max())
    assert np.abs(a["ddos"] - b["ddos"]).max() <= 1e-9 * max(1.0, np.abs(b["ddos"]).max())
    for i, n in enumerate(a["layer_list"]):
        assert (a["ldos"][:, :, i, n:] == 0).all() and (b["ldos"][:, :, i, n:] == 0).all()


@pytest.mark.parametrize("cells", [(2, 1, 8), (2, 2, 6)])
def test_run_tb_circular_leads(engine, Tm, tmp_path, cells):
    """circular_flag = 1 (tb/surface.f90:225-340 -> circular_g,
    tb/circular.f90:43-108): the lead Green's functions by the block-circulant
    decimation over the n_bloch_x x n_bloch_y transverse copies.  H(k) with
    the phases of the actual bond vectors is block-circulant over the copies
    at every k, so T and the surface DOS equal the direct decimation's (both
    batched and point by point) and the oracle's."""
    runs = {}
    for circ in (0, 1):
        folder = tmp_path / f"c{circ}"
        folder.mkdir()
        cfg, lat, pos, sp = R.write_run(str(folder), cells=cells, nk=(2, 1), ne=4, circular=circ)
        print(f"run_tb {cells} circular_flag={circ}", flush=True)
        runs[circ] = driver.run_tb(cfg, engine=engine, dos=False)
    runs["pt"] = driver.run_tb(cfg, engine=engine, dos=False, batched=False)
    a, nl = runs[0], runs[0]["nl"]
    assert nl == 40 * cells[0] * cells[1]
    for key in (1, "pt"):
        b = runs[key]
        assert (b["status"] == 0).all() and b["layer_list"] == a["layer_list"]
        for k in ("transl", "transr", "sdos_l", "sdos_r"):
            assert np.abs(a[k] - b[k]).max() <= 1e-9 * max(1.0, np.abs(a[k]).max()), (key, k)
    assert a["transl"].max() > 0.5
    P = tb.read_param(R.PARAM)
    for ik, k in enumerate(runs[1]["k"]):
        for ie in (0, 3):
            E = runs[1]["E"][ie]
            tl, tr = oracle_T(Tm, P, lat, pos, sp, k, E, 1e-5, runs[1]["layer_list"], nl, nl)
            assert abs(runs[1]["transl"][ie, ik] - tl) <= 1e-8 * max(1.0, tl), (ik, E)
            assert abs(runs[1]["transr"][ie, ik] - tr) <= 1e-8 * max(1.0, tr), (ik, E)


def _degenerate(kz_sets, tol=1e-6):
    for kz in kz_sets:
        kz = np.sort(kz[kz < 99.9])
        if len(kz) > 1 and (np.diff(kz) < tol).any():
            return True
    return False


@pytest.mark.parametrize("defect", [None, 30])
def test_run_tb_modes_and_unfolding(engine, Tm, tmp_path, defect):
    """The loop's vel_mat_mpi (tb/velocity_mat.f90:222-1560) in run_tb on the
    Si/Ge superlattice with a 2 x 1 transverse supercell (n_bloch_x = 2):
    the specular and non-specular parts add up to the Caroli transmission,
    transmission plus reflection is the number of incident propagating modes
    (unitarity), the leads' period is the cell height a = 5.5 A, and at a
    sample of points the split equals the oracle chain's (C oracle lead GFs
    and RGF, oracle.modes.vel_mat_tdb with the tight-binding conventions and
    |lambda| within 1e-3 of 1, oracle.unfold with 10 orbitals per atom on
    the transverse cells and reciprocal lattices of the cell, written out
    here from the structure's closed form).  Without a defect the structure
    has the transverse period of the cell, which conserves the unfolded
    in-plane k: the non-specular parts vanish.  One swapped atom in the
    device (defect) breaks that period and scatters into the other k."""
    from oracle import modes as M
    from oracle import unfold as UF
    a = 5.5
    cfg, lat, pos, sp = R.write_run(str(tmp_path), cells=(2, 1, 8), nk=(2, 1), ne=4, n_bloch=(2, 1),
                                    defect=defect)
    print(f"run_tb with modes, n_bloch 2 x 1, defect {defect}", flush=True)
    r = driver.run_tb(cfg, engine=engine, dos=False)
    assert (r["status"] == 0).all()
    P = tb.read_param(R.PARAM)
    geom, azl, azr = tb.unfold_geometry(pos, lat, sp, P, tb.read_confige(cfg))
    assert azl == pytest.approx(a, abs=1e-12) and azr == pytest.approx(a, abs=1e-12)
    # the mode sums against the Caroli T and unitarity hold up to the lead
    # broadening eta = 1e-5 (the modes counted as propagating are those with
    # |lambda| within 1e-3 of 1; slow modes near a band edge turn eta into
    # O(eta / v) errors, up to 1e-3 in the reflections here, in the oracle
    # chain as on the GPU: the split below equals the oracle's to 2e-6)
    sc = np.maximum(1.0, np.abs(r["transr"]))
    assert np.abs(r["transr_s"] + r["transr_ns"] - r["mode_sums"][..., 0]).max() <= 1e-12 * sc.max()
    assert np.abs(r["transl_s"] + r["transl_ns"] - r["mode_sums"][..., 1]).max() <= 1e-12 * sc.max()
    assert np.abs(r["mode_sums"][..., 0] - r["transr"]).max() <= 1e-4 * sc.max()
    assert np.abs(r["mode_sums"][..., 1] - r["transl"]).max() <= 1e-4 * sc.max()
    npl, npr = r["n_prop"][..., 0], r["n_prop"][..., 3]
    assert np.abs(r["mode_sums"][..., 0] + r["mode_sums"][..., 3] - npl).max() < 5e-3
    assert np.abs(r["mode_sums"][..., 1] + r["mode_sums"][..., 2] - npr).max() < 5e-3
    assert (npl > 0).any() and r["transr_s"].max() > 0.5
    if defect is not None:
        assert r["transr_ns"].max() > 1e-4 and r["transl_ns"].max() > 1e-4
    out = os.path.join(str(tmp_path), "out")
    rows = open(os.path.join(out, "transl_s.dat")).read().splitlines()
    assert len(rows) == 4 and all(len(x) == 20 * 3 + 1 for x in rows)
    assert float(rows[2][20:40]) == pytest.approx(r["transl_s"][2, 0], rel=1e-9, abs=1e-12)
    # reflr_ns.dat holds reflr_s, as the reference writes it
    rns = open(os.path.join(out, "reflr_ns.dat")).read().splitlines()
    assert float(rns[1][40:60]) == pytest.approx(r["reflr_s"][1, 1], rel=1e-9, abs=1e-12)
    # the oracle chain at a sample of points
    cell = np.array([[a / 2, -a / 2, 0.0], [a / 2, a / 2, 0.0], [0.0, 0.0, a]])
    reci_uc = 2 * np.pi * np.linalg.inv(cell).T
    recivec = 2 * np.pi * np.linalg.inv(lat).T
    nl = r["nl"]
    per = 8
    setl = UF.transverse_uc_setup(pos, cell, per, per, 2, 1, 1e-3, ndims=10)
    setr = UF.transverse_uc_setup(pos, cell, len(pos) - 2 * per, per, 2, 1, 1e-5, ndims=10)
    p = Tm.params("tb", eta=1e-5, convergence=1e-10)
    op = O.OrParams(p.eta, p.eta0, p.convergence, p.device_eta_factor, p.eta_mode, p.max_iter)
    names = ("transl_s", "transl_ns", "transr_s", "transr_ns", "refll_s", "refll_ns", "reflr_s", "reflr_ns")
    checked = 0
    for ik, k in enumerate(r["k"]):
        H = TB.gen_ham(P, pos, sp, lat, k, 2.6)
        s = D.slices(H, r["layer_list"], nl, nl, 0, 0, variant="tb")
        for ie in (0, 1, 3):
            E = r["E"][ie]
            g = O.surface_green(s.Haml, s.Hamlv, s.Hamr, s.Hamrv, E, op, want_plus_minus=True)
            o = O.rgf(s.Hd, s.Hdv, s.Hld, s.Hrd, s.Haml, s.Hamr, g["gl_adv"].conj().T, g["gr_ret"], -g["gml_adv"],
                      g["gmr_ret"], g["sl_adv"].conj().T, g["sr_ret"], E, op, Hdv_back=s.Hdv_back,
                      Hrd_fwd=s.Hrd_fwd)
            m = M.vel_mat_tdb(E, g["gl_adv"], g["gr_ret"], g["gl_plus"], g["gr_minus"], g["gml_adv"], g["gmr_ret"],
                              o["G_D_RL"], o["G_D_LR"], o["G_L"], o["G_R"], s.Haml, s.Hamr, s.Hamlv, s.Hamrv,
                              op.eta, op.eta0, 1e-3, a, a, False, True, 2)
            kzs = [m["k_z_" + key] for key in ("L_adv", "L_ret", "R_ret", "R_adv")]
            if _degenerate(kzs):
                continue
            U = {key: m["U_" + key] for key in ("L_adv", "L_ret", "R_ret", "R_adv")}
            lam = {key: m["lam_" + key] for key in ("L_adv", "L_ret", "R_ret", "R_adv")}
            _, _, spl = UF.unfold_point(m, U, lam, k, recivec, reci_uc, reci_uc, 2, 1, pos, setl, setr, ndims=10)
            for n in names:
                assert abs(r[n][ie, ik] - spl[n]) <= 2e-6 * max(1.0, abs(spl[n])), (ik, E, n, r[n][ie, ik], spl[n])
            if defect is None:
                for n in ("transl_ns", "transr_ns", "refll_ns", "reflr_ns"):
                    assert abs(r[n][ie, ik]) <= 1e-6, (ik, E, n, r[n][ie, ik])
            checked += 1
    assert checked >= 2


def _rows(kx, other):
    """(k_x, k_y, k_z, other...) of the propagating modes, sorted by k."""
    sel = kx[:, 2] < 99.9
    a = np.concatenate([kx[sel], other[sel]], axis=1)
    return a[np.lexsort((np.round(a[:, 1], 6), np.round(a[:, 0], 6), np.round(a[:, 2], 6)))]


def test_run_tb_primitive_cell_unfolding(engine, Tm, tmp_path):
    """With the leads' primitive-cell inputs the run unfolds every
    propagating mode to the primitive cells (find_k_pc_l / _r,
    tb/hamiltonian.f90:666-916) and writes the per-mode files of the loop.
    At a sample of points the unfolded k, the primitive k and the velocities
    equal the oracle chain's (oracle.unfold for k, oracle.kpc_tb on the
    primitive cells read by oracle.kpc_tb.read_pc_fdf, a velocity compared
    where one candidate matches and its band is not degenerate); the files
    hold one record
    per propagating incident mode with these columns, md2md records with
    verbosity = 1."""
    from oracle import kpc_tb as KT
    from oracle import modes as M
    from oracle import unfold as UF
    a = 5.5
    cfg, lat, pos, sp = R.write_run(str(tmp_path), cells=(2, 1, 8), nk=(2, 1), ne=4, n_bloch=(2, 1), pc=True,
                                    verbosity=1)
    print("run_tb with primitive-cell unfolding", flush=True)
    r = driver.run_tb(cfg, engine=engine, dos=False)
    assert (r["status"] == 0).all()
    mp = r["mode_points"]
    P = tb.read_param(R.PARAM)
    cell = np.array([[a / 2, -a / 2, 0.0], [a / 2, a / 2, 0.0], [0.0, 0.0, a]])
    reci_uc = 2 * np.pi * np.linalg.inv(cell).T
    recivec = 2 * np.pi * np.linalg.inv(lat).T
    pcl = KT.read_pc_fdf(str(tmp_path / "pc_l.fdf"))
    pcr = KT.read_pc_fdf(str(tmp_path / "pc_r.fdf"))
    opc = {"L": KT.TbPrimitiveCell(P, pcl["positions_pc"], pcl["latvec_pc"], 2.6),
           "R": KT.TbPrimitiveCell(P, pcr["positions_pc"], pcr["latvec_pc"], 2.6)}
    rpc = {"L": pcl["recivec_pc"], "R": pcr["recivec_pc"]}
    nl, per = r["nl"], 8
    setl = UF.transverse_uc_setup(pos, cell, per, per, 2, 1, 1e-3, ndims=10)
    setr = UF.transverse_uc_setup(pos, cell, len(pos) - 2 * per, per, 2, 1, 1e-5, ndims=10)
    p = Tm.params("tb", eta=1e-5, convergence=1e-10)
    op = O.OrParams(p.eta, p.eta0, p.convergence, p.device_eta_factor, p.eta_mode, p.max_iter)
    sets = ("L_adv", "L_ret", "R_ret", "R_adv")
    checked = nvel = 0
    for ik, k in enumerate(r["k"]):
        H = TB.gen_ham(P, pos, sp, lat, k, 2.6)
        s = D.slices(H, r["layer_list"], nl, nl, 0, 0, variant="tb")
        for ie in (0, 3):
            E = r["E"][ie]
            q = ik * len(r["E"]) + ie
            g = O.surface_green(s.Haml, s.Hamlv, s.Hamr, s.Hamrv, E, op, want_plus_minus=True)
            o = O.rgf(s.Hd, s.Hdv, s.Hld, s.Hrd, s.Haml, s.Hamr, g["gl_adv"].conj().T, g["gr_ret"], -g["gml_adv"],
                      g["gmr_ret"], g["sl_adv"].conj().T, g["sr_ret"], E, op, Hdv_back=s.Hdv_back,
                      Hrd_fwd=s.Hrd_fwd)
            m = M.vel_mat_tdb(E, g["gl_adv"], g["gr_ret"], g["gl_plus"], g["gr_minus"], g["gml_adv"], g["gmr_ret"],
                              o["G_D_RL"], o["G_D_LR"], o["G_L"], o["G_R"], s.Haml, s.Hamr, s.Hamlv, s.Hamrv,
                              op.eta, op.eta0, 1e-3, a, a, False, True, 2)
            if _degenerate([m["k_z_" + key] for key in sets]):
                continue
            U = {key: m["U_" + key] for key in sets}
            lam = {key: m["lam_" + key] for key in sets}
            _, kx, _ = UF.unfold_point(m, U, lam, k, recivec, reci_uc, reci_uc, 2, 1, pos, setl, setr, ndims=10)
            for key in sets:
                side = key[0]
                ref, vok = [], []
                for i in np.nonzero(kx[key][:, 2] < 99.9)[0]:
                    kp, v = KT.mode_pc(E, kx[key][i], reci_uc, rpc[side], 2, 1, opc[side])
                    kpcs, _ = KT.find_k_pc(E, kx[key][i], reci_uc, rpc[side], 2, 1, opc[side])
                    e = np.linalg.eigvalsh(opc[side].ham(kp))
                    # a velocity is fixed by the mode when one candidate matches
                    # and one band alone lies near E there
                    vok.append(float(len(kpcs) == 1 and (np.abs(e - E) < 1e-4).sum() == 1))
                    ref.append(np.concatenate([kx[key][i], kp, v]))
                ref = np.array(ref).reshape(-1, 9)
                got = _rows(mp["k_xyz_" + key][q], np.concatenate([mp["kp_" + key][q], mp["v_" + key][q]], axis=1))
                ref = np.concatenate([ref, np.array(vok)[:, None]], axis=1) if len(ref) else ref
                ref = ref[np.lexsort((np.round(ref[:, 1], 6), np.round(ref[:, 0], 6), np.round(ref[:, 2], 6)))] \
                    if len(ref) else ref
                assert got.shape[0] == ref.shape[0], (ik, E, key)
                if not len(ref):
                    continue
                assert np.abs(got[:, :6] - ref[:, :6]).max() < 1e-8, (ik, E, key)
                sel = ref[:, 9] > 0
                if sel.any():
                    assert np.abs(got[sel, 6:9] - ref[sel, 6:9]).max() < 1e-6 * max(1.0, np.abs(ref[sel, 6:9]).max())
                    nvel += int(sel.sum())
            checked += 1
    assert checked >= 2 and nvel > 0
    out = os.path.join(str(tmp_path), "out")
    for ik in range(len(r["k"])):
        rows = open(os.path.join(out, "trl_specular_param_%06d.dat" % (ik + 1))).read().splitlines()
        nin = sum(int((mp["k_xyz_L_adv"][ik * 4 + ie][:, 2] < 99.9).sum()) for ie in range(4))
        assert len(rows) == nin and all(len(x) == 13 * 20 for x in rows)
        md = open(os.path.join(out, "mdl_specular_param_%06d.dat" % (ik + 1))).read().splitlines()
        assert len(md) == nin and all(len(x) == 5 + 14 * 20 for x in md)      # 14 values of '(1I5,15e20.10)'
        m2 = open(os.path.join(out, "tlr_md2md_%06d.dat" % (ik + 1))).read().splitlines()
        assert all(len(x) == 10 + 21 * 20 + 10 for x in m2) and len(m2) > 0
    # the first record of the first point: E, k, kp, v of its first propagating incident mode
    rows = open(os.path.join(out, "trl_specular_param_000001.dat")).read().splitlines()
    q0 = next(ie for ie in range(4) if (mp["k_xyz_L_adv"][ie][:, 2] < 99.9).any())
    i0 = int(np.nonzero(mp["k_xyz_L_adv"][q0][:, 2] < 99.9)[0][0])
    vals = [float(rows[0][20 * c:20 * (c + 1)]) for c in range(13)]
    want = [r["E"][q0], *mp["k_xyz_L_adv"][q0][i0], *mp["kp_L_adv"][q0][i0], *mp["v_L_adv"][q0][i0]]
    assert np.allclose(vals[:10], want, rtol=1e-9, atol=1e-10)


def test_run_tb_pristine_si_integer_channels(engine, tmp_path):
    cfg, *_ = R.write_run(str(tmp_path), split=False, nk=(1, 1), emin=-1.2, emax=2.4, ne=4, eta="1.0d-7")
    print("run_tb pristine Si", flush=True)
    r = driver.run_tb(cfg, engine=engine)
    t = r["transl"][:, 0]
    assert np.abs(t - np.round(t)).max() < 2e-3, t
    assert np.abs(r["transl"] - r["transr"]).max() < 1e-6
    assert t[0] >= 1 and t[-1] >= 1


def test_run_tb_kpath_mode(engine, Tm, tmp_path):
    """path_mode = 3: the k-points of kpath_file (third component set to 0,
    then moved into the Wigner-Seitz cell) against the oracle chain"""
    kp = [[0.1, 0.2, 0.7], [0.45, -0.3, 0.0], [0.8, 0.6, 0.1]]
    cfg, lat, pos, sp = R.write_run(str(tmp_path), path_mode=3, kpath=kp, ne=3)
    print("run_tb kpath", flush=True)
    r = driver.run_tb(cfg, engine=engine)
    assert r["k"].shape == (3, 3) and np.all(r["k"][:, 2] == 0.0)
    assert np.allclose(r["k"][2, :2], [-0.2, -0.4])         # (0.8, 0.6) moved by (1, 1)
    P = tb.read_param(R.PARAM)
    for ik, k in enumerate(r["k"]):
        for ie, E in enumerate(r["E"]):
            tl, tr = oracle_T(Tm, P, lat, pos, sp, k, E, 1e-5, r["layer_list"], 40, 40)
            assert abs(r["transl"][ie, ik] - tl) <= 1e-8 * max(1.0, tl), (ik, E)
            assert abs(r["transr"][ie, ik] - tr) <= 1e-8 * max(1.0, tr), (ik, E)
