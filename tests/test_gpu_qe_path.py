"""The QE force-constant path end to end on the GPU: H(k) assembled by the
GPU kernel equals the host assembly (which the golden tables pin), and the
driver writes the si-si example's T tables in the reference's format with
the golden tables' values (same band-edge tolerance as the batched test)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import si_si as S  # noqa: E402

pytestmark = pytest.mark.gpu


def test_gpu_hk_equals_host_assembly(engine):
    cfg = S.read_config()
    qe = S.read_qe_fc(cfg)
    asm = S.TdbAssembler(qe, cfg)
    ks = S.k_grid(cfg)[[0, 7, 24, 48]]
    print("hk assembly", flush=True)
    d = asm.upload(engine, ks)
    for j, kg in enumerate(ks):
        ref = asm.blocks(kg)
        for name in ("Haml", "Hamlv", "Hamr", "Hamrv", "Hamld", "Hamrd", "Hamd", "Hamdv"):
            g = d[name][j].cpu().numpy()
            g = np.swapaxes(g, -1, -2)                        # ABI storage -> row-major
            r = ref[name]
            if name == "Hamdv" and r.shape[0] == 0:
                continue
            assert np.abs(g - r).max() <= 1e-13 * max(1.0, np.abs(r).max()), name


def test_driver_writes_reference_tables(engine, tmp_path):
    from phinfortran_amd import driver
    out = driver.run_qe_tdb(os.path.join(S.FIXTURE, "config"), os.path.join(S.FIXTURE, "Si444.fc"),
                            str(tmp_path), engine=engine)
    assert (out["status"] == 0).all()
    for which in ("transl", "transr"):
        gold = np.loadtxt(os.path.join(S.FIXTURE, f"{which}_tot.dat"))
        mine = np.loadtxt(os.path.join(tmp_path, f"{which}_tot.dat"))
        assert mine.shape == gold.shape
        assert np.abs(mine[:, 0] - gold[:, 0]).max() == 0.0
        d = np.abs(mine[:, 1:] - gold[:, 1:])
        print(which, "max", d.max(), "beyond 1e-8:", int((d > 1e-8).sum()))
        assert (d > 1e-8).mean() < 0.02 and d.max() < 1e-6


def test_driver_mode_resolved_run(engine, tmp_path):
    """The mode-resolved run writes the reference's ten tables and six per-mode
    files per k-point: the split tables match the committed ones (s + ns at
    every point, s and ns separately away from Gamma, where degenerate modes of
    different unfolded k make the split basis-dependent), and every per-mode
    file has the committed file's row count at every frequency."""
    from phinfortran_amd import driver
    print("driver mode-resolved", flush=True)
    out = driver.run_qe_tdb(os.path.join(S.FIXTURE, "config"), os.path.join(S.FIXTURE, "Si444.fc"),
                            str(tmp_path), engine=engine, modes=True, pc_fdf=os.path.join(S.FIXTURE, "pc1.fdf"))
    assert (out["status"] == 0).all() and (out["modes_status"] == 0).all()
    Eg, split = S.golden_split()
    for base in ("transl", "transr", "refll", "reflr"):
        s = np.loadtxt(os.path.join(tmp_path, base + "_s.dat"))
        ns = np.loadtxt(os.path.join(tmp_path, base + "_ns.dat"))
        assert s.shape == (61, 50) and np.abs(s[:, 0] - Eg).max() == 0.0
        for j, ik in enumerate(S.MODE_KPOINTS):
            gs, gns = split[base + "_s"][:, j], split[base + "_ns"][:, j]
            assert np.abs((s[:, 1 + ik] + ns[:, 1 + ik]) - (gs + gns)).max() < 2e-6 * max(1.0, gs.max())
            if ik != 24:
                assert np.abs(s[:, 1 + ik] - gs).max() < 2e-6 * max(1.0, gs.max()), (base, ik)
    for ik in S.MODE_KPOINTS:
        gold = S.golden_rows("trl", ik)
        path = os.path.join(tmp_path, "trl_specular_param_%06d.dat" % (ik + 1))
        rows = np.atleast_2d(np.loadtxt(path))
        counts = {}
        for r in rows:
            counts[round(r[1], 8)] = counts.get(round(r[1], 8), 0) + 1
        assert counts == {w: len(v) for w, v in gold.items()}, ik
        # the primitive-cell columns are filled (find_k_pc): rows with a
        # primitive-cell k, as many as in the committed table
        g = np.concatenate(list(gold.values()))
        n_gold = int((np.abs(g[:, 4:7]).sum(axis=1) > 0).sum())
        n_mine = int((np.abs(rows[:, 5:8]).sum(axis=1) > 0).sum())
        assert abs(n_mine - n_gold) <= max(2, 0.05 * n_gold), (ik, n_mine, n_gold)
        for name in ("tlr", "right", "left", "mdl", "mdr"):
            assert os.path.getsize(os.path.join(tmp_path, "%s_specular_param_%06d.dat" % (name, ik + 1))) > 0


def test_driver_buttiker_run(engine, tmp_path):
    """The si-si example with buttiker = 1: two iterations of the probe loop,
    its unformatted files, and transmission tables equal to a direct probe
    solve at the temperatures of the last iteration (and different from the
    probe-free tables)."""
    from phinfortran_amd import buttiker as PB
    from phinfortran_amd import driver
    from phinfortran_amd import transport as T
    txt = open(os.path.join(S.FIXTURE, "config")).read()
    txt = txt.replace("buttiker = 0", "buttiker = 1").replace("lifetime = 0", "lifetime = 1")
    txt = txt.replace("t_left = 300.0d0", "t_left = 310.0d0").replace("t_right = 300.0d0", "t_right = 290.0d0")
    cfgp = os.path.join(tmp_path, "config")
    open(cfgp, "w").write(txt)
    print("driver buttiker", flush=True)
    out = driver.run_qe_tdb(cfgp, os.path.join(S.FIXTURE, "Si444.fc"), str(tmp_path), engine=engine, max_outer=2)
    loop = out["buttiker"]
    assert loop["iterations"] == 2 and len(loop["tdiff"]) == 2
    nk, ne = out["transl"].shape[1], out["transl"].shape[0]
    L = out["blocks"]["L"]
    flux = PB.read_fortran_records(os.path.join(tmp_path, "flux.dat"))[0]
    assert flux.size == nk * ne * L
    assert np.array_equal(PB.read_fortran_records(os.path.join(tmp_path, "temperature_i.dat"))[0],
                          loop["temperature_i"])
    cfg = S.read_config()
    p = T.params("phonon", eta=cfg["eta"], eta0=cfg["eta0"], convergence=cfg["convergence"], max_iter=500)
    kidx = np.repeat(np.arange(nk, dtype=np.int32), ne)
    mask = np.tile((np.arange(ne) == 0).astype(np.int32), nk)
    r = engine.solve_buttiker(out["blocks"], np.tile(out["E"], nk), p, loop["temperature_used"], 310.0, 290.0,
                              lifetime=1, kidx=kidx, first_point_mask=mask)
    tl = r["T_LR"].cpu().numpy().reshape(nk, ne).T
    assert np.array_equal(tl, out["transl"])
    mine = np.loadtxt(os.path.join(tmp_path, "transl_tot.dat"))
    gold = np.loadtxt(os.path.join(S.FIXTURE, "transl_tot.dat"))
    assert mine.shape == gold.shape and np.isfinite(mine).all()
    assert np.abs(mine[:, 1:] - gold[:, 1:]).max() > 1e-6       # the probes scatter


def test_driver_eigenvalue_mode(engine, tmp_path):
    """path_mode = 1 on the si-si inputs along a two-point k path: the per-mode
    files of both k-points and no transmission tables (phinfortran.f90:1524);
    the mode-resolved totals equal the path_mode = 2 run's at the same k."""
    from phinfortran_amd import driver
    txt = open(os.path.join(S.FIXTURE, "config")).read().replace("path_mode = 2", "path_mode = 1")
    txt = txt.replace('kpath_file = "kpoint.dat_LXKnew"', 'kpath_file = "kp.dat"')
    cfgp = os.path.join(tmp_path, "config")
    open(cfgp, "w").write(txt)
    (tmp_path / "kp.dat").write_text("2\n0.0 0.0 0.0\n0.142857142857 0.0 0.0\n")
    print("driver eigenvalue mode", flush=True)
    out = driver.run_qe_tdb(cfgp, os.path.join(S.FIXTURE, "Si444.fc"), str(tmp_path), engine=engine)
    assert (out["modes_status"] == 0).all() and out["k"].shape == (2, 3)
    assert not os.path.exists(os.path.join(tmp_path, "transl_tot.dat"))
    assert os.path.exists(os.path.join(tmp_path, "trl_specular_param_000001.dat"))
    assert os.path.exists(os.path.join(tmp_path, "trl_specular_param_000002.dat"))
    # the total transmission of the Gamma point equals the mesh run's (k index 24)
    ref = driver.run_qe_tdb(os.path.join(S.FIXTURE, "config"), os.path.join(S.FIXTURE, "Si444.fc"), engine=engine)
    assert np.abs(out["transl"][:, 0] - ref["transl"][:, 24]).max() < 1e-12
