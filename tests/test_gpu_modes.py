"""Mode-resolved transmission on the GPU (nt_modes_tdb_batched_dev) against
the oracle (oracle/modes.py, itself pinned to the reference's per-mode
tables) and against those tables directly.

Mode order inside the propagating set follows the eigensolver's deflation
order, which is not LAPACK's; the Gram-Schmidt step then mixes
non-orthogonal modes differently, so per-mode values are compared where they
are basis independent (pristine leads, single-mode cases) and the sums over
modes everywhere."""
import math
import os
import sys

import numpy as np
import pytest

import oracle as O
from oracle import modes as M
from phinfortran_amd import lattice as lt

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import si_si as S  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from phinfortran_amd import transport
    return transport


def oparams(p):
    return O.OrParams(p.eta, p.eta0, p.convergence, p.device_eta_factor, p.eta_mode, p.max_iter)


def test_chain_single_mode(engine, T):
    blk = lt.chain_blocks(L=4)
    p = T.params("phonon", eta=1e-9, eta0=1e-10, convergence=1e-14, max_iter=2000)
    E = np.array([0.3, 1.7, 3.5])
    d = engine.upload_blocks(blk)
    print("modes chain", flush=True)
    r = engine.solve_modes(d, E, p, 1.0, 1.0)
    assert (r["modes_status"].cpu().numpy() == 0).all()
    npr = r["n_prop"].cpu().numpy()
    assert (npr == 1).all()
    kz = r["k_z_L_adv"].cpu().numpy()[:, 0]
    assert np.abs(np.abs(kz) - np.arccos(1 - E / 2)).max() < 1e-9
    assert np.abs(r["t_mode_L"].cpu().numpy()[:, 0] - 1).max() < 1e-6


@pytest.mark.parametrize("n,L,nk", [(6, 3, 1), (8, 4, 2)])
def test_random_system_vs_oracle(engine, T, n, L, nk):
    blocks = [lt.random_blocks(n, n, L, seed=70 + i, scale=0.6) for i in range(nk)]
    for b in blocks:      # Hermitian leads with Hamrv = Hamlv^+ mirrored, device as in random_blocks
        b.Hamr[...] = b.Haml
        b.Hamrv[...] = b.Hamlv.conj().T
    p = T.params("phonon", eta=1e-7, eta0=1e-8, convergence=1e-13, max_iter=3000)
    op = oparams(p)
    rng = np.random.default_rng(2)
    npts = 8
    E = rng.uniform(0.05, 0.6, npts)
    kidx = rng.integers(0, nk, npts).astype(np.int32)
    az = 1.3
    d = engine.upload_blocks(blocks)
    print(f"modes random n={n} L={L} nk={nk}", flush=True)
    r = engine.solve_modes(d, E, p, az, az, kidx=kidx if nk > 1 else None)
    assert (r["modes_status"].cpu().numpy() == 0).all()
    sums = r["sums"].cpu().numpy()
    npr = r["n_prop"].cpu().numpy()
    kz = {k: r[k].cpu().numpy() for k in ("k_z_L_adv", "k_z_L_ret", "k_z_R_ret", "k_z_R_adv")}
    for i in range(npts):
        m = M.solve_point_modes(blocks[kidx[i] if nk > 1 else 0], E[i], op, p.lambda_eta, az, az)
        assert list(npr[i]) == [m["n_prop"][k] for k in ("L_adv", "L_ret", "R_ret", "R_adv")]
        for k in kz:
            a = np.sort(kz[k][i][kz[k][i] < 99])
            b = np.sort(m[k][m[k] < 99])
            assert np.abs(a - b).max(initial=0) < 1e-8, (i, k)
        inc, out = m["k_z_L_adv"] < 99, m["k_z_R_ret"] < 99
        ref = [(np.abs(m["trl"][np.ix_(out, inc)]) ** 2).sum(),
               (np.abs(m["tlr"][np.ix_(m["k_z_L_ret"] < 99, m["k_z_R_adv"] < 99)]) ** 2).sum(),
               (np.abs(m["rrr"][np.ix_(out, m["k_z_R_adv"] < 99)]) ** 2).sum(),
               (np.abs(m["rll"][np.ix_(m["k_z_L_ret"] < 99, inc)]) ** 2).sum()]
        assert np.abs(sums[i] - ref).max() < 1e-6 * max(1.0, max(ref)), (i, sums[i], ref)
        # mode sums against Caroli
        assert abs(sums[i][0] - r["T_RL"][i].item()) < 1e-4 * max(1.0, sums[i][0])


def test_si_si_per_mode_tables_gpu(engine, T):
    cfg = S.read_config()
    qe = S.read_qe_fc(cfg)
    asm = S.TdbAssembler(qe, cfg)
    ks, E = S.k_grid(cfg), S.e_grid(cfg)
    az = qe["cell"][2, 2]
    kp = list(S.MODE_KPOINTS)
    d = engine.upload_blocks([asm.blocks(ks[ik]) for ik in kp])
    ne = len(E)
    kidx = np.repeat(np.arange(len(kp), dtype=np.int32), ne)
    mask = np.tile((np.arange(ne) == 0).astype(np.int32), len(kp))
    p = T.params("phonon", eta=cfg["eta"], eta0=cfg["eta0"], convergence=cfg["convergence"], max_iter=500,
                 lambda_eta=5e-3)
    print("modes si-si", flush=True)
    r = engine.solve_modes(d, np.tile(E, len(kp)), p, az, az, kidx=kidx, first_point_mask=mask, want=())
    assert (r["modes_status"].cpu().numpy() == 0).all()
    kz = r["k_z_L_adv"].cpu().numpy()
    tm = r["t_mode_L"].cpu().numpy()
    dk = dt = 0.0
    nmodes = 0
    for j, ik in enumerate(kp):
        gold = S.golden_modes(ik)
        for ie in range(ne):
            q = j * ne + ie
            inc = np.where(kz[q] < 99.9)[0]
            rows = gold.get(round(math.sqrt(E[ie]), 8))
            if rows is None:
                assert len(inc) == 0
                continue
            assert len(inc) == len(rows), (ik, ie)
            o = np.argsort(kz[q][inc])
            g = rows[np.argsort(rows[:, 0])]
            dk = max(dk, np.abs(kz[q][inc][o] - g[:, 0]).max())
            dt = max(dt, np.abs(tm[q][inc][o] - g[:, 1]).max())
            nmodes += len(inc)
    print(f"si-si per-mode: {nmodes} modes, max |dk_z| {dk:.2e}, max |dT_n| {dt:.2e}")
    assert dk < 1e-7 and dt < 2e-6


def test_tb_ribbon_vs_oracle(engine, T):
    """tight-binding conventions (tb=1) on a disordered armchair ribbon"""
    blk = lt.agnr_blocks(n_dimer=4, L=3, seed=2)
    p = T.params("tb", eta=1e-7, convergence=1e-13, max_iter=3000)
    op = oparams(p)
    E = np.array([-1.3, -0.4, 0.45, 1.1])
    d = engine.upload_blocks(blk)
    print("modes tb ribbon", flush=True)
    r = engine.solve_modes(d, E, p, 1.0, 1.0, tb=True)
    assert (r["modes_status"].cpu().numpy() == 0).all()
    sums = r["sums"].cpu().numpy()
    npr = r["n_prop"].cpu().numpy()
    for i in range(len(E)):
        m = M.solve_point_modes(blk, E[i], op, p.lambda_eta, 1.0, 1.0, tb=True)
        assert list(npr[i]) == [m["n_prop"][k] for k in ("L_adv", "L_ret", "R_ret", "R_adv")]
        inc, out = m["k_z_L_adv"] < 99, m["k_z_R_ret"] < 99
        ref = (np.abs(m["trl"][np.ix_(out, inc)]) ** 2).sum()
        assert abs(sums[i][0] - ref) < 1e-6 * max(1.0, ref), (i, sums[i][0], ref)
        assert abs(sums[i][0] - r["T_RL"][i].item()) < 1e-4 * max(1.0, ref)


def _set_sums(kz, tm, tol=1e-4):
    """(k_z, summed per-mode transmission) over sets of incident modes whose
    k_z agree to tol (a degenerate set: eigenvalues within 1e-4 -- or a
    single mode), sorted by k_z"""
    inc = np.where(kz < 99.9)[0]
    order = inc[np.argsort(kz[inc])]
    out, cur = [], None
    for i in order:
        if cur is not None and abs(kz[i] - cur[0]) < tol:
            cur[1] += tm[i]
            cur[2] += 1
        else:
            if cur is not None:
                out.append(cur)
            cur = [kz[i], tm[i], 1]
    if cur is not None:
        out.append(cur)
    return out


def test_degenerate_rotation_vs_oracle(engine, T):
    """path_mode = 1 (degenerate-subspace rotation) at the si-si Gamma point,
    whose leads carry degenerate transverse modes: k_z, the transmission and
    reflection totals and the per-mode transmission summed over every set of
    incident modes sharing one k_z against the oracle's path_mode = 1 (the
    basis inside a degenerate set follows the eigensolver, here and in the
    reference), and the same against the GPU's own path_mode = 2."""
    cfg = S.read_config()
    qe = S.read_qe_fc(cfg)
    asm = S.TdbAssembler(qe, cfg)
    ks, E = S.k_grid(cfg), S.e_grid(cfg)
    az = qe["cell"][2, 2]
    blk = asm.blocks(ks[24])
    d = engine.upload_blocks(blk)
    ies = np.arange(1, len(E), 3)
    p1 = T.params("phonon", eta=cfg["eta"], eta0=cfg["eta0"], convergence=cfg["convergence"], max_iter=500,
                  lambda_eta=5e-3, path_mode=1)
    p2 = T.params("phonon", eta=cfg["eta"], eta0=cfg["eta0"], convergence=cfg["convergence"], max_iter=500,
                  lambda_eta=5e-3)
    op = oparams(p1)
    print("modes path_mode 1", flush=True)
    r1 = engine.solve_modes(d, E[ies], p1, az, az, want=())
    r2 = engine.solve_modes(d, E[ies], p2, az, az, want=())
    assert (r1["modes_status"].cpu().numpy() == 0).all()
    kz1, tm1, s1 = (r1[k].cpu().numpy() for k in ("k_z_L_adv", "t_mode_L", "sums"))
    kz2, tm2, s2 = (r2[k].cpu().numpy() for k in ("k_z_L_adv", "t_mode_L", "sums"))
    ndeg = 0
    for q, ie in enumerate(ies):
        m = M.solve_point_modes(blk, E[ie], op, 5e-3, az, az, path_mode=1)
        ref = [m["T_modes_RL"], m["T_modes_LR"], m["R_modes_R"], m["R_modes_L"]]
        assert np.abs(s1[q] - ref).max() < 1e-6 * max(1.0, max(ref)), (ie, s1[q], ref)
        assert np.abs(s1[q] - s2[q]).max() < 1e-8 * max(1.0, s2[q].max()), ie
        a = _set_sums(kz1[q], tm1[q])
        b = _set_sums(m["k_z_L_adv"], m["t_mode_L"])
        c = _set_sums(kz2[q], tm2[q])
        assert [x[2] for x in a] == [x[2] for x in b] == [x[2] for x in c], ie
        for x, y, z in zip(a, b, c):
            assert abs(x[0] - y[0]) < 1e-7 and abs(x[1] - y[1]) < 2e-6 and abs(x[1] - z[1]) < 1e-8, (ie, x, y, z)
        ndeg += sum(1 for x in a if x[2] > 1)
    print(f"path_mode 1: {ndeg} degenerate incident sets over {len(ies)} energies")
    assert ndeg >= 1
