"""Primitive-cell unfolding on the GPU (nt_unfold_pc_dev, find_k_pc) against
the reference's committed per-mode tables (columns kp_b and v_b of trl and
mdr) and against the oracle (oracle/kpc.py, itself pinned to those columns).
Points with degenerate propagating modes are skipped: their unfolded labels
follow the eigenvector basis (see tests/test_gpu_unfold.py)."""
import math
import os
import sys

import numpy as np
import pytest

from oracle import kpc as K

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import si_si as S  # noqa: E402

pytestmark = pytest.mark.gpu
SETS = ("L_adv", "L_ret", "R_ret", "R_adv")
KPTS = (10, 24)


@pytest.fixture(scope="module")
def run(engine):
    from phinfortran_amd import transport as T
    from phinfortran_amd.unfold import PrimitiveCell, UnfoldGeometry
    cfg = S.read_config()
    qe = S.read_qe_fc(cfg)
    asm = S.TdbAssembler(qe, cfg)
    ks, E = S.k_grid(cfg), S.e_grid(cfg)
    az = qe["cell"][2, 2]
    pc = PrimitiveCell.from_qe(qe, cfg, os.path.join(S.FIXTURE, "pc1.fdf"), S.SPECIES_MASS)
    d = engine.upload_blocks([asm.blocks(ks[ik]) for ik in KPTS])
    ne = len(E)
    kidx = np.repeat(np.arange(len(KPTS), dtype=np.int32), ne)
    mask = np.tile((np.arange(ne) == 0).astype(np.int32), len(KPTS))
    p = T.params("phonon", eta=cfg["eta"], eta0=cfg["eta0"], convergence=cfg["convergence"], max_iter=500,
                 lambda_eta=5e-3)
    print("kpc si-si", flush=True)
    r = engine.solve_modes(d, np.tile(E, len(KPTS)), p, az, az, kidx=kidx, first_point_mask=mask, want=(),
                           unfold=UnfoldGeometry.from_qe(qe, cfg), ksc=ks[list(KPTS)], pc=pc)
    assert (r["modes_status"].cpu().numpy() == 0).all()
    host = {k: v.cpu().numpy() for k, v in r.items() if hasattr(v, "cpu")}
    return dict(cfg=cfg, qe=qe, E=E, ne=ne, r=host)


def degenerate(r, q, tol=1e-6):
    for key in SETS:
        kz = np.sort(r["k_z_" + key][q][r["k_z_" + key][q] < 99.9])
        if len(kz) > 1 and (np.diff(kz) < tol).any():
            return True
    return False


def rows_key(a):
    return a[np.lexsort((np.round(a[:, 1], 6), np.round(a[:, 0], 6), np.round(a[:, 2], 6)))]


def test_primitive_cell_columns_against_reference(run):
    r, E, ne = run["r"], run["E"], run["ne"]
    nrow = nv = 0
    for j, ik in enumerate(KPTS):
        gold = {"trl": S.golden_rows("trl", ik), "mdr": S.golden_rows("mdr", ik)}
        for ie in range(ne):
            q = j * ne + ie
            w = round(math.sqrt(E[ie]), 8)
            if w not in gold["trl"] or degenerate(r, q):
                continue
            for name, key in (("trl", "L_adv"), ("mdr", "R_adv")):
                kx = r["k_xyz_" + key][q]
                inc = kx[:, 2] < 99.9
                a = rows_key(np.concatenate([kx[inc], r["kp_" + key][q][inc], r["v_" + key][q][inc]], axis=1))
                g = rows_key(gold[name][w][:, 1:10])
                assert a.shape == g.shape, (ik, ie, name)
                assert np.abs(a[:, :6] - g[:, :6]).max() < 1e-7, (ik, ie, name, a[:, 3:6], g[:, 3:6])
                assert np.abs(a[:, 6:] - g[:, 6:]).max() < 1e-6 * max(1.0, np.abs(g[:, 6:]).max()), (ik, ie, name)
                nrow += len(a)
                nv += int((np.abs(g[:, 6:]).sum(axis=1) > 0).sum())
    print(f"primitive-cell columns: {nrow} rows, {nv} with a velocity")
    assert nrow > 300 and nv > 50


def test_against_oracle(run):
    """kp, v and the candidate count of sampled modes, the oracle fed the GPU's
    unfolded k"""
    r, E, ne, qe = run["r"], run["E"], run["ne"], run["qe"]
    pcf = K.read_pc_fdf(os.path.join(S.FIXTURE, "pc1.fdf"), S.SPECIES_MASS)
    dyn = K.PcDynamics(qe, pcf["natoms_pc"], pcf["mass_uc"])
    reci_uc = 2 * np.pi * np.linalg.inv(qe["cell"]).T
    checked = 0
    for j in range(len(KPTS)):
        for ie in range(2, ne, 7):
            q = j * ne + ie
            for key in SETS:
                kx = r["k_xyz_" + key][q]
                for i in np.where(kx[:, 2] < 99.9)[0][:3]:
                    kpcs, vs = K.find_k_pc(E[ie], kx[i], reci_uc, pcf["recivec_pc"], 2, 2, dyn)
                    kp, v = K.mode_pc(E[ie], kx[i], reci_uc, pcf["recivec_pc"], 2, 2, dyn)
                    assert r["ncand_" + key][q][i] == len(kpcs), (j, ie, key, i)
                    assert np.abs(r["kp_" + key][q][i] - kp).max() < 1e-12, (j, ie, key, i)
                    assert np.abs(r["v_" + key][q][i] - v).max() < 1e-6 * max(1.0, np.abs(v).max()), (j, ie, key, i)
                    checked += 1
            for key in SETS:
                ev = r["k_xyz_" + key][q][:, 2] >= 99.9
                assert (r["kp_" + key][q][ev] == 100.0).all() and (r["ncand_" + key][q][ev] == 0).all()
    print(f"oracle: {checked} modes")
    assert checked > 50


def test_tight_binding_primitive_cell_against_oracle(engine, tmp_path):
    """nt_unfold_pc_tb_dev (find_k_pc_l, tb/hamiltonian.f90:666-916) on the Si
    primitive cell of the tight-binding example's parameters, seen from the
    four-atom tetragonal lead cell: energies on the primitive bands at random
    k, the mode's k folded into the lead cell's zone, some on its boundary;
    kp, v and the candidate count against oracle/kpc_tb.py (v where one
    candidate matches)."""
    import torch
    from oracle import kpc_tb as KT
    from phinfortran_amd import tb
    from phinfortran_amd.unfold import TbPrimitiveCell
    from tests import tb_run_inputs as R
    a = 5.5
    path = R.write_pc(str(tmp_path), "pc.fdf", 1, a)
    P = tb.read_param(R.PARAM)
    cell = np.array([[a / 2, -a / 2, 0.0], [a / 2, a / 2, 0.0], [0.0, 0.0, a]])
    reci_uc = 2 * np.pi * np.linalg.inv(cell).T
    cfg = dict(n_bloch_x=1, n_bloch_y=1)
    pc = TbPrimitiveCell.from_fdf(P, path, reci_uc, cfg, 2.6)
    o = KT.read_pc_fdf(path)
    opc = KT.TbPrimitiveCell(P, o["positions_pc"], o["latvec_pc"], 2.6)
    assert pc.n_pc == opc.n == 20
    rng = np.random.default_rng(11)
    Es, ks = [], []
    for _ in range(24):
        kpc = rng.uniform(-0.5, 0.5, 3) @ reci_uc
        e = np.linalg.eigvalsh(opc.ham(kpc))
        Es.append(float(e[rng.integers(4, 12)]))
        # the mode's k: kpc moved into the lead cell's zone
        f = np.linalg.solve(reci_uc.T, kpc)
        ks.append((f - np.round(f)) @ reci_uc)
    ks[3] = 0.5 * reci_uc[2]                                 # on the zone boundary: its own k
    Es[3] = float(np.linalg.eigvalsh(opc.ham(ks[3]))[6])
    ks.append(np.array([0.0, 0.0, 100.0]))                   # an evanescent mode
    Es.append(0.3)
    E = np.array(Es)
    kx = np.array(ks)
    print("kpc tight binding", flush=True)
    u = engine.unfold_pc(pc, torch.as_tensor(E), torch.as_tensor(kx))
    kp, v, nc = u["kp"].cpu().numpy(), u["v"].cpu().numpy(), u["ncand"].cpu().numpy()
    assert np.all(kp[-1] == 100.0) and nc[-1] == 0
    nv = 0
    for i in range(len(E) - 1):
        okp, ov = KT.mode_pc(E[i], kx[i], reci_uc, o["recivec_pc"], 1, 1, opc)
        kpcs, _ = KT.find_k_pc(E[i], kx[i], reci_uc, o["recivec_pc"], 1, 1, opc)
        assert nc[i] == len(kpcs), i
        assert np.abs(kp[i] - okp).max() < 1e-10, (i, kp[i], okp)
        e = np.linalg.eigvalsh(opc.ham(okp))
        if len(kpcs) == 1 and (np.abs(e - E[i]) < 1e-4).sum() == 1:
            assert np.abs(v[i] - ov).max() < 1e-8 * max(1.0, np.abs(ov).max()), (i, v[i], ov)
            nv += 1
    assert nv >= 5
