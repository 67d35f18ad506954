"""GPU parity at the real shapes of the five benchmark workloads.

Each test builds the workload's seeded blocks at full cross-section and full
device length and checks a sample of its (omega,k) points against the C
oracle (oracle/nt_oracle.c, the restatement of surface_green_tdb +
get_G_full_tdb, negf/surface.f90:385-605, negf/lesser.f90:475-809):

  wire      Si/Ge alloy wire 2x2 cells, b = 96, L = 20, 512 frequencies in
            (0, 16] THz, nk = 1: every 16th frequency, plus mode-resolved
            sums against Caroli
  ribbon    armchair graphene nanoribbon, 40 dimers (b = 80), L = 60,
            2048 energies in [-3, 3] eV, eta = 1e-8 eV: every 32nd energy,
            plus the pristine ribbon's integer channel counts against the
            lead's propagating-mode count
  uneven    Si/Ge superlattice films of the uneven-batching workload, one per
            block size 24/96/216/384 at its drawn period and length
  film      Si|Ge interface film 4x2 cells, b = 192, L = 40, 8x8 k: two
            k-points x every 32nd of 1024 frequencies
  large     Si film 8x4 cells, b = 768, L = 100, rough 3-layer interface:
            two points against committed oracle values (minutes per point on
            the CPU), a short device of the same cross-section live, and
            full-size invariants

FP64 tolerance, per point: |T_gpu - T_oracle| <= max(1e-10, 256 eps kappa)
max(1, T), kappa the largest 1-norm condition number of the blocks the
oracle's solve inverted.  Both are FP64, but the engine inverts by blocked
Gauss-Jordan with MFMA sums and the oracle by LU with sequential sums, so
they agree to eps times the conditioning of what is inverted: 1e-10 where
that is small, and the rounding scale where it is not (the AGNR at
eta = 1e-8 eV has kappa up to ~1e7).  The printed lines give the achieved
agreement and how many points the rounding scale relaxed.
"""
import json
import os
import time

import numpy as np
import pytest

import oracle as O
from phinfortran_amd import lattice as lt

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-9          # committed b = 768 fixture (no per-point kappa stored)
EPS = 2.220446049250313e-16
NTHREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def T():
    from phinfortran_amd import transport
    return transport


def oparams(p):
    return O.OrParams(p.eta, p.eta0, p.convergence, p.device_eta_factor, p.eta_mode, p.max_iter)


def omega_grid(n, wmax=16.0):
    return wmax * np.arange(1, n + 1) / n


def check_vs_oracle(name, blk, E, r, op):
    t0 = time.time()
    otl, otr, ost, kappa = O.solve_tdb_batch_kappa(blk.as_dict(), E, op, nthreads=NTHREADS)
    assert (ost == 0).all()
    tol = np.maximum(1e-10, 256 * EPS * kappa)
    tl, tr = r["T_LR"].cpu().numpy(), r["T_RL"].cpu().numpy()
    dl = np.abs(tl - otl) / np.maximum(1.0, np.abs(otl))
    dr = np.abs(tr - otr) / np.maximum(1.0, np.abs(otr))
    print(f"{name}: {E.size} points, max |dT|/max(1,T) LR {dl.max():.2e} RL {dr.max():.2e}, "
          f"max |dT|/tol {max((dl / tol).max(), (dr / tol).max()):.2f}, "
          f"{int((tol > 1e-10).sum())} point(s) at the rounding scale (kappa up to {kappa.max():.1e}), "
          f"T in [{otl.min():.3g}, {otl.max():.3g}], oracle {time.time() - t0:.1f}s", flush=True)
    assert (dl <= tol).all(), (name, int((dl / tol).argmax()), tl, otl, kappa)
    assert (dr <= tol).all(), (name, int((dr / tol).argmax()), tr, otr, kappa)
    return otl, otr


def test_wire_b96_L20(engine, T):
    """Si/Ge alloy wire at its full shape; eta(E) = 1e-6 E + 1e-6 omega_max
    (eta0 = 1.6e-5 THz^2)."""
    blk = lt.si_ge_wire(nx=2, ny=2, L=20, seed=0)
    assert (blk.b, blk.L) == (96, 20)
    p = T.params("phonon", eta=1e-6, eta0=1e-6 * 16.0)
    E = omega_grid(512)[7::16] ** 2
    print("wire b=96 L=20: solve", flush=True)
    d = engine.upload_blocks(blk)
    r = engine.solve_tdb(d, E, p)
    assert (r["status"].cpu().numpy() == 0).all()
    otl, _ = check_vs_oracle("wire", blk, E, r, oparams(p))
    # mode-resolved: sums over incident modes equal Caroli (up to the
    # evanescent tails the finite eta leaves in Caroli)
    print("wire b=96 L=20: modes", flush=True)
    m = engine.solve_modes(d, E, p, lt.A_SI, lt.A_SI)
    assert (m["modes_status"].cpu().numpy() == 0).all()
    sums = m["sums"].cpu().numpy()
    tl, tr = m["T_LR"].cpu().numpy(), m["T_RL"].cpu().numpy()
    err = np.abs(sums[:, 0] - tr) / np.maximum(1.0, tr)
    err2 = np.abs(sums[:, 1] - tl) / np.maximum(1.0, tl)
    print(f"wire modes: max |sum T_n - T|/max(1,T) {err.max():.2e} / {err2.max():.2e}", flush=True)
    assert err.max() < 1e-4 and err2.max() < 1e-4


def test_ribbon_b80_L60(engine, T):
    blk = lt.agnr_blocks(n_dimer=40, L=60, seed=0)
    assert (blk.b, blk.L) == (80, 60)
    p = T.params("tb", eta=1e-8)
    E = np.linspace(-3.0, 3.0, 2048)[5::32]
    print("ribbon b=80 L=60: solve", flush=True)
    r = engine.solve_tdb(engine.upload_blocks(blk), E, p)
    assert (r["status"].cpu().numpy() == 0).all()
    check_vs_oracle("ribbon", blk, E, r, oparams(p))


def test_pristine_ribbon_channels_b80_L60(engine, T):
    """Pristine ribbon (device = lead): T is the number of right-moving
    propagating lead modes, an integer up to the eta damping of slow modes
    over the 60 layers (|T - round(T)| ~ 1e-5 at eta = 1e-8 eV), at every
    energy; the same T as the oracle."""
    blk = lt.agnr_blocks(n_dimer=40, L=60, hop_noise=0.0, onsite_w=0.0, seed=0)
    p = T.params("tb", eta=1e-8)
    E = np.linspace(-3.0, 3.0, 2048)[5::32]
    d = engine.upload_blocks(blk)
    print("pristine ribbon b=80 L=60: modes", flush=True)
    m = engine.solve_modes(d, E, p, 1.0, 1.0, tb=True)
    assert (m["status"].cpu().numpy() == 0).all()
    tl, tr = m["T_LR"].cpu().numpy(), m["T_RL"].cpu().numpy()
    nprop = m["n_prop"].cpu().numpy()
    frac = np.abs(tl - np.round(tl))
    print(f"pristine ribbon: max |T - round(T)| {frac.max():.2e}, |T_LR - T_RL| {np.abs(tl - tr).max():.2e}, "
          f"channels {sorted(set(np.round(tl).astype(int)))}", flush=True)
    assert frac.max() < 1e-3
    assert np.abs(tl - tr).max() < 1e-6
    assert (np.round(tl).astype(int) == nprop[:, 0]).all()
    assert (np.round(tl).astype(int) == nprop[:, 2]).all()
    assert len(set(np.round(tl).astype(int))) >= 5
    check_vs_oracle("pristine ribbon", blk, E, m, oparams(p))


UNEVEN = lt.superlattice_set(256, seed=0)


def _first_of_size(n):
    return next(d for d in UNEVEN if d["n"] == n)


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_uneven_device(engine, T, n):
    spec = _first_of_size(n)
    ks = lt.mp_grid(4, 4)
    p = T.params("phonon")
    op = oparams(p)
    npts = 8 if n < 4 else 4
    for kk in (0, 9):
        blk = lt.si_ge_superlattice_film(spec["n"], spec["period"], spec["L"], ks[kk], seed=spec["seed"])
        E = omega_grid(256)[np.linspace(3, 250, npts).astype(int)] ** 2
        print(f"uneven b={spec['b']} L={spec['L']} period={spec['period']} k={kk}", flush=True)
        r = engine.solve_tdb(engine.upload_blocks(blk), E, p)
        assert (r["status"].cpu().numpy() == 0).all()
        check_vs_oracle(f"uneven b={spec['b']}", blk, E, r, op)


@pytest.mark.parametrize("kk", [9, 46])
def test_film_b192_L40(engine, T, kk):
    ks = lt.mp_grid(8, 8)
    blk = lt.si_ge_interface_film(nx=4, ny=2, L=40, kfrac=ks[kk])
    assert (blk.b, blk.L) == (192, 40)
    p = T.params("phonon")
    E = omega_grid(1024)[11::32] ** 2
    print(f"film b=192 L=40 k={kk}", flush=True)
    r = engine.solve_tdb(engine.upload_blocks(blk), E, p)
    assert (r["status"].cpu().numpy() == 0).all()
    check_vs_oracle("film", blk, E, r, oparams(p))


def test_large_b768_vs_committed_oracle(engine, T):
    """Two points of the b = 768, L = 100 workload against the oracle values
    committed by tests/golden/make_c5_oracle.py."""
    ref = json.load(open(os.path.join(HERE, "golden", "c5_oracle.json")))
    p = T.params("phonon")
    # max_iter differs (oracle default 200, engine 300) but is not reached:
    # the fixture's decimations converge in about 24 iterations
    for key in ("eta", "eta0", "convergence", "device_eta_factor"):
        assert getattr(p, key) == ref["params"][key]
    assert max(pt["iters"] for pt in ref["points"]) < min(p.max_iter, ref["params"]["max_iter"])
    for pt in ref["points"]:
        blk = lt.rough_interface_film(nx=8, ny=4, L=100, rough_layers=3, seed=0, kfrac=tuple(pt["kfrac"]))
        assert (blk.b, blk.L) == (768, 100)
        print(f"large b=768 L=100 E={pt['E']}", flush=True)
        r = engine.solve_tdb(engine.upload_blocks(blk), np.array([pt["E"]]), p)
        assert r["status"].item() == 0
        tl, tr = r["T_LR"].item(), r["T_RL"].item()
        print(f"large: T_LR {tl!r} vs {pt['T_LR']!r}, T_RL {tr!r} vs {pt['T_RL']!r}", flush=True)
        assert abs(tl - pt["T_LR"]) < TOL * max(1.0, pt["T_LR"])
        assert abs(tr - pt["T_RL"]) < TOL * max(1.0, pt["T_RL"])


def test_large_b768_short_device_live(engine, T):
    """Full b = 768 cross-section, 6 layers around the rough interface,
    checked live against the oracle."""
    masses = lt.rough_interface_masses(256, L=6, rough_layers=3, seed=1)
    blk = lt.rough_interface_film(nx=8, ny=4, L=6, kfrac=(0.1875, 0.0625), masses=masses)
    p = T.params("phonon")
    E = np.array([3.1]) ** 2
    print("large b=768 L=6 live", flush=True)
    r = engine.solve_tdb(engine.upload_blocks(blk), E, p)
    assert (r["status"].cpu().numpy() == 0).all()
    check_vs_oracle("large short", blk, E, r, oparams(p))


def test_large_b768_L100_invariants(engine, T):
    """Full-size invariants at b = 768, L = 100: T_LR = T_RL up to the eta
    asymmetry of the two sweeps, mode sums equal Caroli, and T bounded by
    the number of propagating modes of either lead."""
    blk = lt.rough_interface_film(nx=8, ny=4, L=100, rough_layers=3, seed=0, kfrac=(0.3125, 0.4375))
    p = T.params("phonon")
    E = np.array([1.3, 3.7]) ** 2
    d = engine.upload_blocks(blk)
    print("large b=768 L=100 modes", flush=True)
    t0 = time.time()
    m = engine.solve_modes(d, E, p, lt.A_SI, lt.A_SI)
    print(f"large modes: {time.time() - t0:.1f}s", flush=True)
    assert (m["status"].cpu().numpy() == 0).all()
    assert (m["modes_status"].cpu().numpy() == 0).all()
    tl, tr = m["T_LR"].cpu().numpy(), m["T_RL"].cpu().numpy()
    sums, nprop = m["sums"].cpu().numpy(), m["n_prop"].cpu().numpy()
    print(f"large: T_LR {tl}, T_RL {tr}, sums {sums[:, :2]}, n_prop {nprop}", flush=True)
    assert (tl > 0).all()
    assert (np.abs(tl - tr) / tl < 1e-3).all()
    assert (np.abs(sums[:, 0] - tr) / np.maximum(1.0, tr) < 1e-4).all()
    assert (np.abs(sums[:, 1] - tl) / np.maximum(1.0, tl) < 1e-4).all()
    assert (tl <= np.minimum(nprop[:, 0], nprop[:, 2]) + 1e-6).all()
