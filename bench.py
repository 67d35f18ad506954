#!/usr/bin/env python3
"""Throughput of the batched (omega,k)-solve at block 192.

Workload (one GPU): Si|Ge interface film, 4x2 conventional-cell cross-section
(192 dof per principal layer), 40 device layers, transverse k on an 8x8 grid,
1024 frequencies in (0, 16] THz.  One step = the full hot path (four lead
decimations' worth of surface GFs needed for T, both RGF sweeps and Caroli T)
for all 1024 frequencies of one k-point, inputs resident in HBM.

Prints one JSON line (see the repository's measurement contract).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "mixed": 157.3}     # MI355X dense MFMA TFLOP/s (FP64, FP32)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_per_launch(kernel, prec, b, points):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 --pmc summary of the same workload (FETCH_SIZE and WRITE_SIZE
    per dispatch, each scaled by the counter's reading for a copy of known
    size); None when no summary covers this kernel and shape."""
    import re
    path = os.path.join(ROOT, "profiles", f"{kernel}_b{b}_{prec}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+-]+)", txt)
    p = re.search(r"points_per_launch:\s*(\d+)", txt)
    if not m or not p or int(p.group(1)) != points:
        return None
    return float(m.group(1))


def kgrid(n):
    """Monkhorst-Pack style grid centred like negf/config.f90:156-175."""
    ks = []
    for i in range(n):
        for j in range(n):
            ks.append((i / n - 0.5 * (n - 1) / n, j / n - 0.5 * (n - 1) / n))
    return ks


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=1024, help="frequencies per step (one k-point)")
    ap.add_argument("--layers", type=int, default=40)
    ap.add_argument("--precision", choices=["auto", "fp64", "mixed"], default="auto",
                    help="path reported as value: auto = mixed when it meets dT/T <= 1e-6 and is faster")
    ap.add_argument("--cpu-sample", type=int, default=16, help="solves in the CPU baseline sample (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-adjoint", dest="adjoint", action="store_false", help="skip the adjoint measurement")
    ap.add_argument("--no-mixed", dest="mixed", action="store_false", help="skip the mixed-precision run")
    ap.add_argument("--adam-steps", type=int, default=50,
                    help="Adam steps maximising G over the 2048 bonds on a sample of the workload (0 = skip)")
    ap.add_argument("--adam-freqs", type=int, default=32, help="frequencies of the Adam sample (per k-point)")
    ap.add_argument("--mass-points", type=int, default=16,
                    help="frequencies of the interface-mass gradient at b = 768, L = 100 (0 = skip)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="before timing, run one device-to-device copy of this many MiB (HBM-counter calibration "
                         "under rocprofv3 --pmc)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false")
    ap.add_argument("--point-split", type=int, default=2,
                    help="parts a solve of many points runs as, side by side on worker streams (1 = one stream)")
    ap.add_argument("--serial-sweeps", action="store_true",
                    help="every launch on one stream in every step: the two RGF sweeps (with two sweeps) and the "
                         "parts of a split solve (for per-kernel profiles)")
    ap.add_argument("--c3-devices", type=int, default=16,
                    help="uneven-batching sample: devices of the seeded 256-device Si/Ge superlattice draw (the "
                         "first of each block size 24/96/216/384 in draw order, in equal numbers; 256 = the whole "
                         "draw), solved in one multi-device call (0 = skip)")
    ap.add_argument("--c3-freqs", type=int, default=64,
                    help="frequencies per device of the uneven-batching sample (of its 256 in (0, 16] THz)")
    ap.add_argument("--c3-k", type=int, default=2,
                    help="transverse k-grid side of the uneven-batching sample (2: 2x2 k-points per device)")
    ap.add_argument("--no-modes", dest="modes", action="store_false",
                    help="skip the mode-resolved sample (the si-si example's full mode-resolved run: solve, "
                         "modes, Brillouin-zone and primitive-cell unfolding of 49 k x 61 omega)")
    ap.add_argument("--no-buttiker", dest="buttiker", action="store_false",
                    help="skip the Buttiker-probe measurement")
    ap.add_argument("--buttiker-points", type=int, default=128, help="frequencies of the probe solve")
    ap.add_argument("--buttiker-cpu-sample", type=int, default=1, help="points of the probe CPU sample (0 = skip)")
    ap.add_argument("--no-configs", dest="configs", action="store_false",
                    help="skip the full-size wire (b=96, L=20, 512 omega, with modes) and ribbon (b=80, L=60, "
                         "2048 E) legs")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="skip the graph-replay leg (a small device solved eagerly and as a captured HIP graph)")
    ap.add_argument("--no-tb-hk", dest="tb_hk", action="store_false",
                    help="skip the tight-binding H(k) assembly measurement")
    ap.add_argument("--tb-hk-only", action="store_true",
                    help="measure only the tight-binding H(k) assembly (for kernel profiles)")
    ap.add_argument("--no-tb-run", dest="tb_run", action="store_false",
                    help="skip the tight-binding transport run measurement")
    ap.add_argument("--tb-run-only", action="store_true",
                    help="measure only the tight-binding transport run (for kernel profiles)")
    ap.add_argument("--tb-hk-k", type=int, default=64, help="k-points per tight-binding H(k) launch")
    ap.add_argument("--modes-cpu-sample", type=int, default=6,
                    help="points of the mode-resolved CPU baseline sample (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the (omega,k) points shard across GPUs with no "
              "inter-GPU traffic but a final gather, and multi-GPU runs are not part of this build; "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from phinfortran_amd import lattice as lt
    from phinfortran_amd import transport as T
    from phinfortran_amd import _native as N

    dev = torch.device("cuda", 0)
    L = args.layers
    ks = kgrid(8)
    nsteps = args.warmup + args.steps
    t0 = time.time()
    blocks = [lt.si_ge_interface_film(nx=4, ny=2, L=L, kfrac=ks[i % len(ks)]) for i in range(nsteps)]
    log(f"built {nsteps} k-point block sets in {time.time() - t0:.1f}s (b={blocks[0].b}, L={L})")
    omega = 16.0 * np.arange(1, args.points + 1) / args.points
    E = torch.as_tensor(omega ** 2, dtype=torch.float64, device=dev)

    eng = T.Engine(0, batch_limit=args.points)
    eng.set_point_split(1 if args.serial_sweeps else args.point_split)
    if args.serial_sweeps:
        eng.set_overlap(False)
    dblocks = [eng.upload_blocks(b) for b in blocks]
    torch.cuda.synchronize()
    if args.copy_probe_mib:
        src = torch.empty(args.copy_probe_mib * (1 << 20) // 8, dtype=torch.float64, device=dev).fill_(1.0)
        dst = src.clone()
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written")
        del src, dst
    def run_tb_hk():
        """gen_ham of the reference's tight-binding example (Si/Ge superlattice
        2 x 2 x 8 four-atom cells: 128 atoms, sp3d5s*, 1280 orbitals) for
        tb_hk_k k-points per launch; the oracle's gen_ham on one k-point
        beside it"""
        from phinfortran_amd import tb
        P = tb.read_param(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "tb", "param"))
        st = tb.diamond_superlattice(5.5, 2, 2, 8)
        A = tb.TbAssembler(P, st["latvec"], st["positions"], st["species"], 2.6)
        n = int(round(np.sqrt(args.tb_hk_k)))
        kk = np.array([[(i + 0.5) / n - 0.5, (j + 0.5) / n - 0.5, 0.0] for i in range(n) for j in range(n)])
        H = A.assemble(eng, kk)                                           # warm-up, allocates the output
        eng.profile_reset()
        eng.profile(True)
        torch.cuda.synchronize()
        nrep = 10
        t0 = time.perf_counter()
        for s in range(nrep):
            A.assemble(eng, kk, out=H)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        eng.profile(False)
        e = eng.profile_get(N.NT_K_ELEM)
        avg_ms = e["ms"] / max(1, e["launches"])
        nk, N_ = len(kk), A.norb
        alg = 16.0 * N_ * N_ * nk                                          # H(k) written once
        import re
        prof = os.path.join(ROOT, "profiles", f"tb_hk_sige_N{N_}_k{nk}_kernel_trace.md")
        traffic = None
        if os.path.exists(prof):
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+-]+)", open(prof).read())
            traffic = float(m.group(1)) if m else None
        res = {"k_points_per_s": nrep * nk / el, "unit": "H(k)/s", "norb": N_, "k_per_launch": nk,
               "segments": A.nseg, "entries": A.nent, "avg_launch_ms": avg_ms, "launches": e["launches"],
               "roofline": {"bound": "hbm", "achieved": alg / (avg_ms * 1e-3) / 1e9, "peak": 8000.0,
                            "unit": "GB/s", "frac": alg / (avg_ms * 1e-3) / 1e9 / 8000.0,
                            "algorithmic_bytes": alg, "traffic": traffic}}
        # write-only reference: torch's fill of the same buffer, timed on torch's stream
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        H.zero_()
        ev0.record()
        for s in range(nrep):
            H.zero_()
        ev1.record()
        torch.cuda.synchronize()
        fill_ms = ev0.elapsed_time(ev1) / nrep
        res["write_reference"] = {"what": "torch zero_() of the same output buffer (write only)",
                                  "ms": fill_ms, "GB_per_s": alg / (fill_ms * 1e-3) / 1e9}
        log(f"tb hk: {nk} k of N={N_} in {avg_ms:.3f} ms per launch ({res['roofline']['frac']:.2f} of HBM); "
            f"fill of the same buffer {fill_ms:.3f} ms")
        if args.cpu_sample > 0:
            from oracle import tb_ham as OT
            t0 = time.perf_counter()
            OT.gen_ham(P, st["positions"], st["species"], st["latvec"], kk[1], 2.6)
            dt = time.perf_counter() - t0
            res["cpu_baseline"] = {"value": 1.0 / dt, "unit": "H(k)/s", "cores": 1, "kind": "port",
                                   "sample": f"one k-point of the same structure, the oracle's gen_ham (pure "
                                             f"Python loops), {dt:.1f} s"}
        return res

    def run_tb_transport():
        """The tight-binding driver (driver.run_tb) on the reference's example
        structure: Si/Ge superlattice 2 x 2 x 8 cells (1280 orbitals), one
        16-atom cell layer per lead period, 4 x 4 k-points x 64 energies; the
        oracle chain (surface GFs + RGF on the same slices) on a sample"""
        import tempfile
        from phinfortran_amd import driver as DRV
        from phinfortran_amd import tb
        from phinfortran_amd import dense as D
        here = os.path.dirname(os.path.abspath(__file__))
        with tempfile.TemporaryDirectory() as tmp:
            st = tb.diamond_superlattice(5.5, 2, 2, 8)
            with open(os.path.join(tmp, "input.fdf"), "w") as f:
                f.write(f"NumberOfAtoms {len(st['positions'])}\nNumberOfSpecies 2\nLatticeConstant 1.0 Ang\n"
                        "%block LatticeVectors\n")
                for r in st["latvec"]:
                    f.write(" ".join(f"{x:.13f}" for x in r) + "\n")
                f.write("%endblock LatticeVectors\n%block AtomicCoordinatesAndAtomicSpecies\n")
                for x, sp in zip(st["positions"], st["species"]):
                    f.write(" ".join(f"{v:.13f}" for v in x) + f" {sp}\n")
                f.write("%endblock AtomicCoordinatesAndAtomicSpecies\n")
            with open(os.path.join(tmp, "confige"), "w") as f:
                f.write("&configlist\n k_trans_mesh = 4 4 1, emin = -1.5, emax = 2.2, ne = 64\n"
                        " r_cutoff = 2.6, filename_input = 'input.fdf'\n/\n"
                        "&setup\n period_left = 16, period_right = 16, buffer_left = 0, buffer_right = 0\n/\n"
                        "&convergelist\n eta = 1.0d-5, convergence = 1.0d-10\n/\n"
                        "&unfolding\n n_bloch_x = 1, n_bloch_y = 1, circular_flag = 0\n/\n"
                        "&kpath\n path_mode = 2\n/\n&saving\n output_dir = '', verbosity = 0\n/\n")
            prm = os.path.join(here, "tests", "golden", "tb", "param")
            cfgp = os.path.join(tmp, "confige")
            DRV.run_tb(cfgp, engine=eng, param_path=prm, dos=False, modes=False)        # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = DRV.run_tb(cfgp, engine=eng, param_path=prm, dos=False, modes=False)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            DRV.run_tb(cfgp, engine=eng, param_path=prm, modes=False)      # warm-up, with the device DOS
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rd = DRV.run_tb(cfgp, engine=eng, param_path=prm, modes=False)
            torch.cuda.synchronize()
            eld = time.perf_counter() - t0
            # the reference's whole loop: transmission, device DOS and vel_mat of every point
            DRV.run_tb(cfgp, engine=eng, param_path=prm)                   # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rm = DRV.run_tb(cfgp, engine=eng, param_path=prm)
            torch.cuda.synchronize()
            elm = time.perf_counter() - t0
        npt = r["transl"].size
        res = {"points_per_s": npt / el, "unit": "(E,k)-points/s", "points": npt, "seconds": el,
               "with_device_dos_points_per_s": npt / eld,
               "with_dos_and_modes_points_per_s": npt / elm,
               "modes_max_unitarity_dev": float(np.abs(rm["mode_sums"][..., 0] + rm["mode_sums"][..., 3]
                                                       - rm["n_prop"][..., 0]).max()),
               "norb": r["norb"], "nl": r["nl"], "layer_list": r["layer_list"],
               "status_ok": bool((r["status"] == 0).all() and (rd["status"] == 0).all()
                                 and (rm["status"] == 0).all()),
               "what": "whole run: table, H(k) of 16 k-points on the GPU, blockwise_setup (host), device "
                       "slicing, one batched solve (layers padded to the widest), FP64; with the device DOS "
                       "(get_ldos of every point) beside it, and with the device DOS and the mode-resolved "
                       "path (vel_mat of every point: lead modes, t/r, unfolding) as the reference's loop"}
        log(f"tb run: {npt} points in {el:.2f}s ({npt / el:.1f} points/s), layers {r['layer_list']}; "
            f"with the device DOS {eld:.2f}s ({npt / eld:.1f} points/s); with DOS and modes {elm:.2f}s "
            f"({npt / elm:.1f} points/s)")
        if args.cpu_sample > 0:
            import oracle as O
            from oracle import tb_ham as OT
            P = tb.read_param(prm)
            p = T.params("tb", eta=1e-5, convergence=1e-10)
            op = O.OrParams(p.eta, p.eta0, p.convergence, p.device_eta_factor, p.eta_mode, p.max_iter)
            t0 = time.perf_counter()
            Hk = OT.gen_ham(P, st["positions"], st["species"], st["latvec"], r["k"][5], 2.6)
            sl = D.slices(Hk, r["layer_list"], r["nl"], r["nr"], 0, 0, variant="tb")
            for e in r["E"][::16]:
                g = O.surface_green(sl.Haml, sl.Hamlv, sl.Hamr, sl.Hamrv, e, op, want_plus_minus=False)
                O.rgf(sl.Hd, sl.Hdv, sl.Hld, sl.Hrd, sl.Haml, sl.Hamr, g["gl_adv"].conj().T, g["gr_ret"],
                      -g["gml_adv"], g["gmr_ret"], g["sl_adv"].conj().T, g["sr_ret"], e, op,
                      Hdv_back=sl.Hdv_back, Hrd_fwd=sl.Hrd_fwd)
            dt = time.perf_counter() - t0
            res["cpu_baseline"] = {"value": 4 / dt, "unit": "(E,k)-points/s", "cores": 1, "kind": "port",
                                   "sample": f"4 energies of one k-point: the oracle's gen_ham (Python) and the C "
                                             f"oracle's decimation + RGF on the same layers, {dt:.1f} s"}
        return res

    if args.tb_run_only:
        print(json.dumps({"tb_transport_run": run_tb_transport()}))
        eng.close()
        return
    if args.tb_hk_only:
        print(json.dumps({"tb_hk_assembly": run_tb_hk()}))
        eng.close()
        return
    tuned, gtuned = {}, {}
    if args.autotune:
        for prec in ("fp64", "mixed"):
            tuned[prec] = eng.autotune_inverse(blocks[0].b, args.points, prec)
            gtuned[prec] = eng.autotune_gemm(blocks[0].b, args.points, prec)
            log(f"inverse panel width for b={blocks[0].b} {prec}: {tuned[prec][0]} ({tuned[prec][1]:.2f} ms); "
                f"gemm tile {gtuned[prec][0]} ({gtuned[prec][1]:.3f} ms)")

    mixed_counts = {}

    def run(prec):
        """warmup + timed steps for one precision; returns (s, per-step T, kernel stats, mean iters)"""
        p = T.params("phonon", precision=prec)
        for s in range(args.warmup):
            r = eng.solve_tdb(dblocks[s], E, p)
            log(f"{prec} warmup {s}: ok={int((r['status'] == 0).sum())}/{args.points}, "
                f"mean iters {eng.last_mean_iters():.1f}")
        eng.profile_reset()
        eng.profile(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outs, iters = [], []
        counts = None
        for s in range(args.steps):
            r = eng.solve_tdb(dblocks[args.warmup + s], E, p, stream_sync=False)
            outs.append(r)
            iters.append(eng.last_mean_iters())
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        eng.profile(False)
        kern = {"gemm": eng.profile_get(N.NT_K_GEMM), "inverse": eng.profile_get(N.NT_K_INV),
                "elementwise": eng.profile_get(N.NT_K_ELEM)}
        log(f"{prec}: {args.steps * args.points / el:.1f} solves/s; " + ", ".join(
            f"{k}: {v['launches']} launches {v['ms']:.1f} ms {v['flops'] / max(1e-9, v['ms'] * 1e-3) / 1e12:.2f} TF/s"
            for k, v in kern.items()))
        ok = all(bool((r["status"] == 0).all()) for r in outs)
        if prec == "mixed":
            counts = eng.mixed_counts()
            log(f"mixed: points {counts[0]}, second refinement steps {counts[1]}, FP64 recomputes {counts[2]}")
            mixed_counts.update(points=counts[0], second_steps=counts[1], fp64_recomputes=counts[2])
        return el, [r["T_LR"].cpu().numpy() for r in outs], kern, float(np.mean(iters)), ok

    el64, T64, k64, nbar, ok64 = run("fp64")

    def isolated_kernels(prec):
        """one more step with every launch on one stream: per-launch kernel
        times that include no time shared with another stream's kernels"""
        p = T.params("phonon", precision=prec)
        eng.set_overlap(False)
        eng.set_point_split(1)
        eng.profile_reset()
        eng.profile(True)
        eng.solve_tdb(dblocks[args.warmup], E, p)
        eng.profile(False)
        eng.set_overlap(not args.serial_sweeps)
        eng.set_point_split(1 if args.serial_sweeps else args.point_split)
        return {"gemm": eng.profile_get(N.NT_K_GEMM), "inverse": eng.profile_get(N.NT_K_INV)}

    iso64 = isolated_kernels("fp64")
    elmx, Tmx, kmx, _, okmx = run("mixed") if args.mixed else (el64, T64, k64, nbar, ok64)

    def run_adjoint():
        """forward + adjoint of the conductance weights over the window of the
        2048 trainable bonds (the bonds owned by the 16 layers around the
        interface: device layers 13..28, 1-based, and their couplings), against
        the full forward (both sweeps) and a single left-to-right sweep (the
        adjoint entry with a one-layer window at the last layer)"""
        from phinfortran_amd.adjoint import conductance_weights
        p = T.params("phonon")
        w = torch.as_tensor(conductance_weights(np.sqrt(omega ** 2), 300.0, (4 * 2) * (lt.A_SI * 1e-10) ** 2),
                            dtype=torch.float64, device=dev)
        lo, hi = L // 2 - 8, L // 2 + 8

        def timed(f):
            for s in range(args.warmup):
                f(dblocks[s])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for s in range(args.steps):
                r = f(dblocks[args.warmup + s])
            torch.cuda.synchronize()
            return time.perf_counter() - t0, bool((r["status"] == 0).all())
        el, ok = timed(lambda d: eng.solve_tdb_adjoint(d, E, p, w, lo, hi, stream_sync=False))
        el1, _ = timed(lambda d: eng.solve_tdb_adjoint(d, E, p, w, L - 1, L - 1, stream_sync=False))
        log(f"adjoint (window {lo}..{hi}): {args.steps * args.points / el:.1f} (omega,k)-gradients/s; "
            f"single-sweep forward {args.steps * args.points / el1:.1f} solves/s")
        res = {"grads_per_s": args.steps * args.points / el, "unit": "(omega,k)-gradients/s",
               "what": f"T_LR and dW/dH of device layers {lo}..{hi} (0-based), the blocks the 2048 trainable bonds "
                       f"of the 16 interfacial layers enter; W = conductance-weighted sum of T_LR, FP64",
               "window_layers": [lo, hi], "bytes_per_point": eng.adjoint_bytes_per_point(dblocks[0], lo, hi),
               "single_sweep_forward_solves_per_s": args.steps * args.points / el1,
               "cost_vs_single_sweep_forward": el / el1, "status_ok": ok}
        if args.adam_steps > 0:
            from phinfortran_amd.adjoint import InterfaceFilm, maximize_conductance
            ks = kgrid(8)
            f = 16.0 * np.arange(1, args.points + 1) / args.points
            f = f[:: max(1, args.points // args.adam_freqs)]
            t0 = time.perf_counter()
            film = InterfaceFilm(eng, 4, 2, L, [ks[9], ks[36]], f, n_iface=16, params=p)
            tb = time.perf_counter() - t0
            rng = np.random.default_rng(0)
            beta = torch.as_tensor(13.8 + 0.5 * rng.standard_normal(film.nparam), device=film.dev)
            t0 = time.perf_counter()
            _, hist = maximize_conductance(film, beta, steps=args.adam_steps, lr=0.2)
            ta = time.perf_counter() - t0
            log(f"adam: {args.adam_steps} steps, G {hist[0]:.4e} -> {hist[-1]:.4e} in {ta:.1f}s")
            res["adam"] = {"steps": args.adam_steps, "parameters": film.nparam, "G_first": hist[0],
                           "G_last": hist[-1], "seconds": ta, "setup_seconds": tb,
                           "sample": f"2 k-points of the 8x8 grid x {len(f)} of the {args.points} frequencies",
                           "unit": "W m^-2 K^-1"}
        return el, ok, res

    def run_mass_adjoint():
        """gradient of G(300 K) with respect to the 768 interface masses of the
        Si film with a rough Si/Ge interface (8x4 cells, b = 768, L = 100),
        with checkpointed recomputation of the layer Green's functions"""
        from phinfortran_amd.adjoint import MassFilm
        f = 16.0 * np.arange(1, args.mass_points + 1) / args.mass_points
        t0 = time.perf_counter()
        mf = MassFilm(eng, 8, 4, 100, [kgrid(8)[27]], f, rough_layers=3, seed=0, params=T.params("phonon"))
        tb = time.perf_counter() - t0
        m = mf.initial()
        out = {}
        for every in (0, -1):
            eng.set_adjoint_checkpoint(every)
            nbytes = eng.adjoint_bytes_per_point(mf.dblocks, mf.w_lo, mf.w_hi)
            mf.conductance(m)                      # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            G, g, _ = mf.conductance(m)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            out["checkpointed" if every else "full_storage"] = {
                "grads_per_s": len(f) / el, "adjoint_bytes_per_point": nbytes, "G": float(G),
                "max_abs_dG_dm": float(g.abs().max())}
            log(f"mass adjoint b=768 L=100 ({'checkpointed' if every else 'full storage'}): "
                f"{len(f) / el:.1f} gradients/s, {nbytes / 2 ** 20:.0f} MiB per point")
        eng.set_adjoint_checkpoint(0)
        # the forward solve of the same points in FP64 and in mixed precision
        fwd = {}
        for prec in ("fp64", "mixed"):
            pp = T.params("phonon", precision=prec)
            eng.solve_tdb(mf.dblocks, mf.E, pp, kidx=mf.kidx)                # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = eng.solve_tdb(mf.dblocks, mf.E, pp, kidx=mf.kidx)
            el = time.perf_counter() - t0
            fwd[prec] = (len(f) / el, r["T_LR"].cpu().numpy(), bool((r["status"] == 0).all()))
        t64, tmx = fwd["fp64"][1], fwd["mixed"][1]
        sig = t64 > 1e-3 * t64.max()
        mx_err = float(np.max(np.abs(tmx[sig] - t64[sig]) / t64[sig])) if sig.any() else 0.0
        out["forward"] = {"fp64_solves_per_s": fwd["fp64"][0], "mixed_solves_per_s": fwd["mixed"][0],
                          "mixed_max_rel_dT": mx_err, "status_ok": fwd["fp64"][2] and fwd["mixed"][2]}
        log(f"b=768 L=100 forward: fp64 {fwd['fp64'][0]:.1f}, mixed {fwd['mixed'][0]:.1f} solves/s, "
            f"mixed max rel dT {mx_err:.2e}")
        out.update(unit="(omega,k)-gradients/s", points=len(f), setup_seconds=tb,
                   what="Si film 8x4 cells (b = 768), L = 100, 3-layer rough Si/Ge interface: T_LR and "
                        "dG/dm of the 768 interface masses, one k-point, FP64")
        return out

    eladj, okadj, adjres = run_adjoint() if args.adjoint else (None, True, None)
    massadj = run_mass_adjoint() if args.mass_points > 0 else None

    def run_c3():
        """uneven devices in one call, FP64 and mixed; per-device mixed error;
        devices of the same block size share their (pristine Si) leads, whose
        Green's functions the engine then computes once; FP64 solves/s per
        block size from one call per size"""
        draw = lt.superlattice_set(256, seed=3)
        if args.c3_devices >= len(draw):
            specs = draw
        else:
            per = max(1, args.c3_devices // 4)
            specs = [sp for n in (1, 2, 3, 4) for sp in [d for d in draw if d["n"] == n][:per]]
        ks = lt.mp_grid(args.c3_k, args.c3_k)
        # every device: an evenly spread sample of its 256 frequencies at every k-point of the grid
        f = (16.0 * np.arange(1, 257) / 256)[np.linspace(0, 255, args.c3_freqs).astype(int)]
        devs, npts = [], 0
        t0 = time.time()
        for sp in specs:
            blk = [lt.si_ge_superlattice_film(sp["n"], sp["period"], sp["L"], kf, seed=sp["seed"]) for kf in ks]
            E = np.tile(f ** 2, len(ks))
            kidx = np.repeat(np.arange(len(ks), dtype=np.int32), len(f))
            devs.append(dict(blocks=eng.upload_blocks(blk), E=E, kidx=kidx))
            npts += E.size
        log(f"c3: {len(devs)} devices built in {time.time() - t0:.1f}s")
        tune = {}
        if args.autotune:
            # per block size: the inverse's panel width and the GEMM's output
            # tile, timed on batches of one device's points
            for bs in sorted({sp["b"] for sp in specs}):
                for prec in ("fp64", "mixed"):
                    nb, ims = eng.autotune_inverse(bs, 256, prec)
                    tb, gms = eng.autotune_gemm(bs, 256, prec)
                    tune[f"{bs}/{prec}"] = dict(inverse_panel=nb, inverse_ms=ims, gemm_tile=tb, gemm_ms=gms)
                    log(f"c3 tuning b={bs} {prec}: inverse panel {nb} ({ims:.3f} ms), gemm tile {tb} ({gms:.3f} ms)")
        res = {}
        for prec in ("fp64", "mixed") if args.mixed else ("fp64",):
            p = T.params("phonon", precision=prec)
            eng.solve_tdb_multi(devs, p)                      # warm-up
            torch.cuda.synchronize()
            log(f"c3 {prec}: warm-up call done")
            t0 = time.perf_counter()
            outs = eng.solve_tdb_multi(devs, p)
            el = time.perf_counter() - t0
            res[prec] = (npts / el, [o["T_LR"].cpu().numpy() for o in outs],
                         all(bool((o["status"] == 0).all()) for o in outs))
            log(f"c3 {prec}: {npts / el:.1f} solves/s")
        # the same FP64 call with the groups one after the other on the context
        p = T.params("phonon")
        eng.set_multi_workers(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.solve_tdb_multi(devs, p)
        one_worker = npts / (time.perf_counter() - t0)
        eng.set_multi_workers(4)
        log(f"c3 fp64, groups one after the other: {one_worker:.1f} solves/s")
        errs = []
        for a, b in zip(res["fp64"][1], res["mixed"][1]) if "mixed" in res else ():
            sig = a > 1e-3 * max(1e-300, a.max())
            errs.append(float(np.max(np.abs(b[sig] - a[sig]) / a[sig])) if sig.any() else 0.0)
        per_b = {}
        p = T.params("phonon")
        for bs in sorted({sp["b"] for sp in specs}):
            sub = [d for d, sp in zip(devs, specs) if sp["b"] == bs]
            n = sum(len(d["E"]) for d in sub)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            eng.solve_tdb_multi(sub, p)
            per_b[bs] = n / (time.perf_counter() - t0)
            log(f"c3 fp64 b={bs}: {per_b[bs]:.1f} solves/s ({len(sub)} devices)")
        return {"devices": [dict(b=sp["b"], L=sp["L"], period=sp["period"]) for sp in specs],
                "points": npts, "fp64_solves_per_s": res["fp64"][0],
                "mixed_solves_per_s": res["mixed"][0] if "mixed" in res else None,
                "fp64_solves_per_s_per_block": per_b, "workers": 4,
                "fp64_solves_per_s_one_worker": one_worker,
                "mixed_max_rel_dT_per_device": errs,
                "status_ok": res["fp64"][2] and (res["mixed"][2] if "mixed" in res else True),
                "tuning": tune,
                "what": f"{len(specs)} devices of the seeded 256-device Si/Ge superlattice draw (block size, "
                        f"period and length from the draw), {f.size} of the 256 frequencies x {len(ks)} k-points "
                        "each, all in one multi-device call; devices with the same leads and points share their "
                        "lead Green's functions"}

    c3 = run_c3() if args.c3_devices > 0 else None

    def run_modes():
        """the si-si example's mode-resolved run (its committed inputs) on the GPU;
        the CPU restatement of the same chain on a few points beside it"""
        from phinfortran_amd import driver
        fix = os.path.join(ROOT, "tests", "golden", "si_si")
        cfgp, fcp, pcp = (os.path.join(fix, f) for f in ("config", "Si444.fc", "pc1.fdf"))
        driver.run_qe_tdb(cfgp, fcp, engine=eng, modes=True, pc_fdf=pcp)            # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        o = driver.run_qe_tdb(cfgp, fcp, engine=eng, modes=True, pc_fdf=pcp)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        npts = o["transl"].size
        ok = bool((o["status"] == 0).all() and (o["modes_status"] == 0).all())
        log(f"mode-resolved si-si: {npts} points in {el:.2f}s")
        res = {"points": npts, "points_per_s": npts / el, "seconds": el, "status_ok": ok,
               "what": "si-si example (b=48, L=2, 49 k x 61 omega): H(k) assembly, decimation, RGF, four "
                       "mode sets, t/r, Brillouin-zone unfolding and find_k_pc of every propagating mode, FP64"}
        if args.modes_cpu_sample > 0:
            import oracle as O
            from oracle import kpc as K
            from oracle import modes as OM
            from oracle import unfold as OU
            from phinfortran_amd import qe as Q
            cfg = Q.read_config(cfgp)
            qe = Q.read_qe_fc(cfg, fcp)
            asm = Q.TdbAssembler(qe, cfg)
            ks, E = Q.k_grid(cfg), Q.e_grid(cfg)
            pos = qe["pos_all"]
            N, per, buf = len(pos), cfg["period"], cfg["buffer"]
            nbx, nby = cfg["n_bloch_x"], cfg["n_bloch_y"]
            setl = OU.transverse_uc_setup(pos, qe["cell"], buf + per, per, nbx, nby, 1e-3)
            setr = OU.transverse_uc_setup(pos, qe["cell"], N - buf - 2 * per, per, nbx, nby, 1e-5)
            recivec = 2 * np.pi * np.linalg.inv(qe["cell_all"]).T
            reci_uc = 2 * np.pi * np.linalg.inv(qe["cell"]).T
            pcf = K.read_pc_fdf(pcp, Q.SPECIES_MASS)
            dyn = K.PcDynamics(qe, pcf["natoms_pc"], pcf["mass_uc"])
            op = O.OrParams(cfg["eta"], cfg["eta0"], cfg["convergence"], 10.0, O.ETA_RELATIVE, 500)
            az = qe["cell"][2, 2]
            ik = 24
            blk = asm.blocks(ks[ik])
            pts = np.linspace(1, len(E) - 1, args.modes_cpu_sample).astype(int)
            t0 = time.perf_counter()
            for ie in pts:
                m = OM.solve_point_modes(blk, E[ie], op, cfg["lambda_eta"], az, az)
                U = {k: m["U_" + k] for k in ("L_adv", "L_ret", "R_ret", "R_adv")}
                lam = {k: m["lam_" + k] for k in ("L_adv", "L_ret", "R_ret", "R_adv")}
                bz, kx, sp = OU.unfold_point(m, U, lam, ks[ik], recivec, reci_uc, reci_uc, nbx, nby, pos,
                                             setl, setr)
                for key in ("L_adv", "L_ret", "R_ret", "R_adv"):
                    for i in np.where(kx[key][:, 2] < 99.9)[0]:
                        K.mode_pc(E[ie], kx[key][i], reci_uc, pcf["recivec_pc"], nbx, nby, dyn)
            dt = time.perf_counter() - t0
            res["cpu_baseline"] = {"value": len(pts) / dt, "unit": "(omega,k)-points/s", "cores": 1,
                                   "kind": "port",
                                   "sample": f"{len(pts)} points of the Gamma k-point, the same chain in numpy "
                                             f"(one core), {dt:.1f} s"}
            log(f"mode-resolved cpu sample: {len(pts)} points in {dt:.1f}s")
        return res

    modes = run_modes() if args.modes else None

    def run_configs():
        """the Si/Ge alloy wire (2x2 cells, b = 96, L = 20, 512 frequencies,
        Caroli and mode-resolved T) and the armchair ribbon (40 dimers, b = 80,
        L = 60, 2048 energies, eta = 1e-8 eV) at their full sizes, FP64, one
        batched call each after a warm-up; the GEMM tile and inverse panel of
        each block size from the autotuners"""
        out = {}
        wire = lt.si_ge_wire(nx=2, ny=2, L=20, seed=0)
        rib = lt.agnr_blocks(n_dimer=40, L=60, seed=0)
        cases = [("wire_b96_L20", wire, (16.0 * np.arange(1, 513) / 512) ** 2,
                  T.params("phonon", eta=1e-6, eta0=1e-6 * 16.0)),
                 ("ribbon_b80_L60", rib, np.linspace(-3.0, 3.0, 2048), T.params("tb", eta=1e-8))]
        for name, blk, E, p in cases:
            tune = None
            if args.autotune:
                nb, _ = eng.autotune_inverse(blk.b, len(E), "fp64")
                tb, _ = eng.autotune_gemm(blk.b, len(E), "fp64")
                tune = dict(inverse_panel=nb, gemm_tile=tb)
            d = eng.upload_blocks(blk)
            Ed = torch.as_tensor(E, dtype=torch.float64, device="cuda")
            eng.solve_tdb(d, Ed, p)                              # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = eng.solve_tdb(d, Ed, p)
            el = time.perf_counter() - t0
            ok = bool((r["status"] == 0).all())
            res = {"b": blk.b, "L": blk.L, "points": len(E), "solves_per_s": len(E) / el, "seconds": el,
                   "mean_decimation_iters": eng.last_mean_iters(), "status_ok": ok, "tuning": tune}
            log(f"{name}: {len(E)} points in {el * 1e3:.1f} ms ({len(E) / el:.1f} solves/s), tuning {tune}")
            if name.startswith("wire"):
                eng.solve_modes(d, Ed, p, lt.A_SI, lt.A_SI)     # warm-up
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                m = eng.solve_modes(d, Ed, p, lt.A_SI, lt.A_SI)
                elm = time.perf_counter() - t0
                res["modes_points_per_s"] = len(E) / elm
                res["modes_status_ok"] = bool((m["modes_status"] == 0).all())
                log(f"{name} mode-resolved: {len(E)} points in {elm:.2f} s ({len(E) / elm:.1f} points/s)")
            out[name] = res
        return out

    configs = run_configs() if args.configs else None

    def run_buttiker():
        """probe solve (both sweeps with a probe in every layer, every block of
        the full G, probe currents and Jacobian) on the headline k-point blocks;
        the oracle's get_G_full_tdb + get_G_lesser_tdb on a sample beside it"""
        npb = args.buttiker_points
        Eb = E[:: max(1, args.points // npb)][:npb].contiguous()
        p = T.params("phonon")
        temp = np.linspace(310.0, 290.0, L)
        eng.solve_buttiker(dblocks[0], Eb, p, temp, 320.0, 280.0)          # warm-up
        eng.profile_reset()
        eng.profile(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nrep = max(1, args.steps // 2)
        for s in range(nrep):
            r = eng.solve_buttiker(dblocks[args.warmup + s], Eb, p, temp, 320.0, 280.0, stream_sync=False)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        eng.profile(False)
        g = eng.profile_get(N.NT_K_GEMM)
        npt = nrep * Eb.numel()
        res = {"points_per_s": npt / el, "unit": "(omega,k)-probe-solves/s", "points": npt,
               "status_ok": bool((r["status"] == 0).all()),
               "what": f"b={blocks[0].b}, L={L}, {Eb.numel()} frequencies per k-point: decimation, both sweeps "
                       f"with a probe self-energy in every layer, all L^2 blocks of the full G, probe currents, "
                       f"Jacobian, contact currents, FP64",
               "gemm": {"launches": g["launches"], "ms": g["ms"],
                        "tflops": g["flops"] / max(1e-9, g["ms"] * 1e-3) / 1e12,
                        "share_of_wall": g["ms"] * 1e-3 / el}}
        log(f"buttiker: {npt / el:.1f} probe solves/s, gemm {res['gemm']['tflops']:.1f} TF/s")
        if args.buttiker_cpu_sample > 0:
            import oracle as O
            from oracle import buttiker as OB
            blk = blocks[args.warmup].as_dict()
            op = O.OrParams(p.eta, p.eta0, p.convergence, p.device_eta_factor, p.eta_mode, p.max_iter)
            Es = Eb.cpu().numpy()[:: max(1, npb // args.buttiker_cpu_sample)][: args.buttiker_cpu_sample]
            t0 = time.perf_counter()
            for e in Es:
                sg = O.surface_green(blk["Haml"], blk["Hamlv"], blk["Hamr"], blk["Hamrv"], e, op,
                                     want_plus_minus=False)
                gg = dict(gl=sg["gl_adv"].conj().T, gr=sg["gr_ret"], gml=-sg["gml_adv"], gmr=sg["gmr_ret"],
                          sl=sg["sl_adv"].conj().T, sr=sg["sr_ret"])
                sig, dsig = OB.sigma_bp(np.array([e]), temp, T.B_PROBE, 430.0)
                f = OB.full_tdb(blk, e, op.eta, op.eta0, sig=sig[:, 0], **gg)
                OB.lesser_tdb(blk, e, f, gg["gl"], gg["gr"], sig[:, 0], dsig[:, 0], temp, 320.0, 280.0)
            dt = time.perf_counter() - t0
            nth = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
            res["cpu_baseline"] = {"value": len(Es) / dt, "unit": "(omega,k)-probe-solves/s", "cores": nth,
                                   "kind": "port",
                                   "sample": f"{len(Es)} point(s) at b={blocks[0].b}, L={L}: the reference's "
                                             f"recursion (dGn for every layer pair) in numpy, threaded BLAS, "
                                             f"{dt:.1f} s"}
            log(f"buttiker cpu sample: {len(Es)} points in {dt:.1f}s")
        return res

    probes = run_buttiker() if args.buttiker else None

    def run_graph():
        """the allocation-free FP64 solve of a small device (b = 24, L = 40,
        64 points: the launch-bound end of the uneven-batching workload)
        eagerly and as a captured HIP graph (decimation check off, max_iter
        64); replays are checked against the eager result"""
        blk = lt.si_ge_superlattice_film(1, 4, 40, (0.125, 0.25), seed=7)
        d = eng.upload_blocks(blk)
        p = T.params("phonon", max_iter=64)
        Ed = torch.as_tensor((16.0 * np.arange(1, 65) / 64) ** 2, dtype=torch.float64, device="cuda")
        reps = 20
        eng.set_decimation_check(4)
        ref = eng.solve_tdb(d, Ed, p, noalloc=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            eng.solve_tdb(d, Ed, p, noalloc=True, stream_sync=False)
        torch.cuda.synchronize()
        t_eager = (time.perf_counter() - t0) / reps
        eng.set_decimation_check(0)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            eng.solve_tdb(d, Ed, p, noalloc=True, stream_sync=False)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            out = eng.solve_tdb(d, Ed, p, noalloc=True, stream_sync=False)
        eng.set_decimation_check(4)
        g.replay()
        torch.cuda.synchronize()
        same = bool(torch.equal(out["T_LR"], ref["T_LR"]) and torch.equal(out["T_RL"], ref["T_RL"]))
        t0 = time.perf_counter()
        for _ in range(reps):
            g.replay()
        torch.cuda.synchronize()
        t_graph = (time.perf_counter() - t0) / reps
        log(f"graph: b=24 L=40 64 points, eager {t_eager * 1e3:.2f} ms, graph replay {t_graph * 1e3:.2f} ms, "
            f"equal {same}")
        return {"what": "b=24, L=40 superlattice device, 64 frequencies, allocation-free FP64 solve",
                "eager_ms": t_eager * 1e3, "graph_replay_ms": t_graph * 1e3,
                "eager_solves_per_s": 64 / t_eager, "graph_solves_per_s": 64 / t_graph,
                "replay_equals_eager": same, "graph_decimation_iterations": 64}


    graph = run_graph() if args.graph else None
    tbhk = run_tb_hk() if args.tb_hk else None

    tbrun = run_tb_transport() if args.tb_run else None
    t64 = np.concatenate(T64)
    tmx = np.concatenate(Tmx)
    sig = t64 > 1e-3 * max(1e-300, t64.max())
    rel_err = float(np.max(np.abs(tmx[sig] - t64[sig]) / t64[sig])) if sig.any() else 0.0
    solves = args.steps * args.points
    v64, vmx = solves / el64, solves / elmx
    use_mixed = args.precision == "mixed" or (rel_err <= 1e-6 and vmx > v64)
    value = vmx if use_mixed else v64
    kern = kmx if use_mixed else k64
    t_el = elmx if use_mixed else el64
    prec_used = "mixed" if use_mixed else "fp64"
    b = blocks[0].b
    flop_solve = b ** 3 * (112 * nbar + 40 * L + 104)
    dom_name, dom = max(((k, v) for k, v in kern.items() if k != "elementwise"), key=lambda kv: kv[1]["ms"])
    avg_ms_conc = dom["ms"] / max(1, dom["launches"])
    avg_flops = dom["flops"] / max(1, dom["launches"])
    achieved_conc = avg_flops / (avg_ms_conc * 1e-3) / 1e12
    # the roofline's per-launch time comes from the serialised step (same
    # kernels and shapes, no second stream sharing the GPU)
    iso = iso64[dom_name] if prec_used == "fp64" else None
    if iso and iso["launches"]:
        avg_ms = iso["ms"] / iso["launches"]
        achieved = iso["flops"] / iso["launches"] / (avg_ms * 1e-3) / 1e12
    else:
        avg_ms, achieved = avg_ms_conc, achieved_conc
    peak = PEAK[prec_used] if dom_name == "gemm" else PEAK["fp64"]
    if prec_used == "mixed" and dom_name == "inverse":
        peak = PEAK["mixed"]

    cpu = None
    parity = None
    if args.cpu_sample > 0:
        import oracle as O
        nth = args.cpu_threads or min(16, os.cpu_count() or 1)
        ns = args.cpu_sample
        # the frequencies of the first timed step's k-point, so the oracle's
        # T doubles as a parity check of the timed results
        sel = np.arange(len(omega))[:: max(1, len(omega) // ns)][:ns]
        sample_E = (omega ** 2)[sel]
        p = T.params("phonon")
        op = O.OrParams(p.eta, p.eta0, p.convergence, p.device_eta_factor, p.eta_mode, p.max_iter)
        t0 = time.perf_counter()
        tl, tr, st = O.solve_tdb_batch(blocks[args.warmup].as_dict(), sample_E, op, nthreads=nth)
        dt = time.perf_counter() - t0
        cpu = {"value": ns / dt, "unit": "(omega,k)-solves/s", "cores": nth, "kind": "port",
               "sample": f"{ns} solves at b={b}, L={L} (the first timed step's k-point, {ns} of its "
                         f"frequencies), FP64 C oracle, OpenMP over points, {dt:.1f} s"}
        log(f"cpu baseline: {ns} solves in {dt:.1f}s on {nth} threads")
        d64 = np.abs(T64[0][sel] - tl) / np.maximum(1.0, np.abs(tl))
        dmx = np.abs(Tmx[0][sel] - tl) / np.maximum(1.0, np.abs(tl))
        parity = {"points": ns, "status_ok": bool((st == 0).all()),
                  "fp64_max_abs_dT_over_max1T": float(d64.max()),
                  "mixed_max_abs_dT_over_max1T": float(dmx.max()) if args.mixed else None,
                  "what": "timed T_LR of the first timed step against the C oracle on the same inputs"}
        log(f"timed-step parity vs oracle: fp64 {d64.max():.2e}, mixed {dmx.max():.2e}")

    out = {
        "metric": "RGF (omega,k)-solves/s at block=192",
        "value": value,
        "unit": "(omega,k)-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if prec_used == "fp64" else "c64 sweep + c128 decimation",
        "data": "synthetic",
        "config": {"workload": "Si|Ge interface film 4x2 cells, b=192, L=%d, 8x8 k-grid, "
                               "%d frequencies in (0,16] THz; one k-point per step" % (L, args.points),
                   "block": b, "layers": L, "points_per_step": args.points, "precision": prec_used,
                   "parallelism": "single GPU, batched over frequencies"},
        "roofline": {"bound": "mfma", "kernel": dom_name, "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": traffic_per_launch(dom_name, prec_used, b, args.points),
                     "launches": (iso or dom)["launches"], "avg_launch_ms": avg_ms,
                     "timing": "HIP events around every launch on the launching stream, in one extra step with "
                               "every launch on one stream (no kernel shares the GPU with another)",
                     "concurrent": {"avg_launch_ms": avg_ms_conc, "achieved": achieved_conc,
                                    "what": f"the same kernel in the timed steps, where the step's points run as "
                                            f"{args.point_split} part(s) side by side on worker streams "
                                            f"(nt_set_point_split) and launches share the GPU"},
                     "inverse_isolated": {"avg_launch_ms": iso64["inverse"]["ms"] / max(1, iso64["inverse"]["launches"]),
                                          "tflops": iso64["inverse"]["flops"] / max(1e-9, iso64["inverse"]["ms"] * 1e-3) / 1e12}},
        "precision_compare": {"fp64_solves_per_s": v64, "mixed_solves_per_s": vmx,
                              "mixed_speedup": vmx / v64, "mixed_max_rel_dT": rel_err,
                              "criterion": "max |T_mixed - T_fp64| / T_fp64 over T > 1e-3 max T, all timed steps",
                              "refinement": mixed_counts,
                              "status_ok": bool(ok64 and okmx)} if args.mixed else None,
        "tuning": {"inverse_panel": {k: v[0] for k, v in tuned.items()},
                   "gemm_tile": {k: v[0] for k, v in gtuned.items()}, "sweeps": 1,
                   "decimation_check_every": 4},
        "solve_model": {"mean_decimation_iters": nbar, "flop_per_solve": flop_solve,
                        "solve_tflops": flop_solve * value / 1e12},
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
    }
    if c3 is not None:
        out["uneven_batching"] = c3
    if modes is not None:
        out["mode_resolved"] = modes
    if configs is not None:
        out["workloads"] = configs
    if graph is not None:
        out["graph_replay"] = graph
    if probes is not None:
        out["buttiker"] = probes
    if tbhk is not None:
        out["tb_hk_assembly"] = tbhk
    if tbrun is not None:
        out["tb_transport_run"] = tbrun
    if eladj is not None:
        adjres["cost_vs_fp64_forward"] = eladj / el64
        out["adjoint"] = adjres
    if massadj is not None:
        out["mass_adjoint"] = massadj
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
