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
