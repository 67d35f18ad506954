"""Self-consistent Buttiker-probe temperatures on the batched GPU solve: the
driver's outer loop for buttiker = 1 (negf/phinfortran.f90:1150-1330).

Each iteration solves every (omega, k) point with the probes of the current
temperatures (nt_buttiker_tdb_batched_dev), integrates the probe currents and
their Jacobian over frequency on the GPU, then takes the reference's host
steps: phonon temperatures (get_ph_temperature), one Newton step of the
probe temperatures (update_temperature) and the unformatted files of
write_flux / write_temperature.  It stops when the rms temperature change
falls below 1e-4 K or after 100 iterations, as the reference.

    initial_temperatures   negf/phinfortran.f90:187-199
    simpson_weights        simpson (negf/util.f90:1015-1043) as a weight vector
    integrate_jacobian     negf/lesser.f90:1506-1529
    update_temperature     negf/lesser.f90:1531-1557
    get_ph_temperature     negf/lesser.f90:1588-1623
    write_flux             negf/lesser.f90:1559-1586
    write_temperature      negf/lesser.f90:1625-1640
"""
from __future__ import annotations

import os

import numpy as np

from .transport import B_PROBE

PI = 3.141592653589793238
KB = 1.380649e-23
HBAR = 1.0545718e-34
MAX_OUTER = 100          # niit of the driver (phinfortran.f90:1151-1155)
TDIFF_STOP = 1.0e-4      # phinfortran.f90:1330
PH_TOL = 1.0e-4          # get_ph_temperature's stopping test (lesser.f90:1608)
PH_MAX_STEPS = 200       # bound on that Newton loop (the reference's is unbounded)


def bose(freq, T):
    """bose_einstein (negf/util.f90:987-998); freq in THz."""
    x = np.exp(HBAR * np.asarray(freq) * 1.0e12 * 2 * PI / KB / T)
    return 1.0 / (x - 1.0)


def dbose(freq, T):
    """dbosedT (negf/util.f90:1000-1013)."""
    x = np.exp(HBAR * np.asarray(freq) * 1.0e12 * 2 * PI / KB / T)
    return x / (x - 1.0) ** 2 * HBAR * np.asarray(freq) * 1.0e12 * 2 * PI / KB / T ** 2


def initial_temperatures(L: int, t_left: float, t_right: float) -> np.ndarray:
    """Probe temperatures before the first iteration: from a quarter of the
    way to three quarters of the way from t_left to t_right."""
    if L == 1:
        return np.array([float(t_left)])
    i = np.arange(L, dtype=float)
    return t_left + 0.25 * (t_right - t_left) + 0.5 * i * (t_right - t_left) / (L - 1)


def simpson_weights(x) -> np.ndarray:
    """w with simpson(x, y) = w . y: dx/3 (1, 4, 2, 4, ..., 2, 4, 1), dx the
    last spacing; the number of points must be odd."""
    x = np.asarray(x, float)
    n = len(x)
    if n < 3 or (n - 1) % 2:
        raise ValueError("ne should be an odd number")
    w = np.full(n, 2.0)
    w[1::2] = 4.0
    w[0] = w[-1] = 1.0
    return w * (x[n - 1] - x[n - 2]) / 3.0


def update_temperature(temp, jac, flux):
    """One Newton step T <- T - Re(J^-1 flux); every temperature becomes 1 K
    when one turns negative.  Returns (new temperatures, rms change)."""
    t_old = np.asarray(temp, float)
    t_new = t_old - np.real(np.linalg.solve(np.asarray(jac), np.asarray(flux)))
    if t_new.min() < 0.0:
        t_new[:] = 1.0
    return t_new, float(np.sqrt(np.dot(t_new - t_old, t_new - t_old) / len(t_new)))


def get_ph_temperature(a_sum, gn_sum, E, t_ph):
    """Per layer, the temperature whose Bose occupation reproduces the layer's
    energy density: Newton on int (Gn - f_T A) E dw = 0.  a_sum, gn_sum: (ne, L)
    (k-averaged traces); returns the updated temperatures."""
    E = np.asarray(E, float)
    w = np.sqrt(E)
    sw = simpson_weights(w)
    t_ph = np.array(t_ph, float)
    for nn in range(a_sum.shape[1]):
        t_old = -100.0
        for _ in range(PH_MAX_STEPS):
            if not abs(t_old - t_ph[nn]) > PH_TOL:
                break
            t_old = t_ph[nn]
            jdiff = float(np.real(sw @ (-dbose(w, t_ph[nn]) * E * a_sum[:, nn])))
            diff = float(np.real(sw @ (gn_sum[:, nn] * E - bose(w, t_ph[nn]) * E * a_sum[:, nn])))
            t_ph[nn] = t_old - diff / jdiff
    return t_ph


def integrate_jacobian(jac_pts, flux_pts, E, nk: int):
    """Frequency integral of the k-summed Jacobian and probe currents:
    J_ij = simpson(w, sum_k jac(k, :, i, j) w / nk).  jac_pts (nk*ne, L, L) and
    flux_pts (nk*ne, L), point = k * ne + ie, torch tensors on the GPU;
    returns numpy (L x L, L)."""
    import torch
    E = np.asarray(E, float)
    w = np.sqrt(E)
    wt = np.tile(simpson_weights(w) * w / nk, nk)
    wt = torch.as_tensor(wt, dtype=torch.float64, device=jac_pts.device)
    L = flux_pts.shape[1]
    jac = (wt @ jac_pts.reshape(-1, L * L)).reshape(L, L)
    flux = wt @ flux_pts
    return jac.cpu().numpy(), flux.cpu().numpy()


def _fortran_record(f, arr):
    """one sequential unformatted record: int32 length, data, int32 length"""
    data = np.asarray(arr, dtype="<f8").ravel(order="F").tobytes()
    n = np.array([len(data)], dtype="<i4").tobytes()
    f.write(n + data + n)


def write_flux(out_dir, flux_all, gn_all, a_all, flux_left, flux_right):
    """flux.dat, gn.dat, A.dat (nk, ne, L) and flux_left.dat, flux_right.dat
    (nk, ne): real parts, Fortran order, one unformatted record each.  The
    reference opens flux_left.dat on a second unit for flux_right
    (negf/lesser.f90:1573-1574), which a Fortran runtime refuses while the
    first unit holds the file; flux_right goes to its own file here."""
    os.makedirs(out_dir, exist_ok=True)
    for name, a in (("flux.dat", flux_all), ("gn.dat", gn_all), ("A.dat", a_all), ("flux_left.dat", flux_left),
                    ("flux_right.dat", flux_right)):
        with open(os.path.join(out_dir, name), "wb") as f:
            _fortran_record(f, np.real(a))


def write_temperature(out_dir, temperature_i, temperature_ph):
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "temperature_i.dat"), "wb") as f:
        _fortran_record(f, temperature_i)
    with open(os.path.join(out_dir, "temperature_ph.dat"), "wb") as f:
        _fortran_record(f, temperature_ph)


def read_fortran_records(path):
    """the float64 records of a sequential unformatted file"""
    out = []
    raw = open(path, "rb").read()
    pos = 0
    while pos < len(raw):
        n = int(np.frombuffer(raw, "<i4", 1, pos)[0])
        out.append(np.frombuffer(raw, "<f8", n // 8, pos + 4).copy())
        if int(np.frombuffer(raw, "<i4", 1, pos + 4 + n)[0]) != n:
            raise ValueError(f"{path}: record markers differ")
        pos += n + 8
    return out


def run(engine, dblocks, E, p, t_left: float, t_right: float, B_probe: float = B_PROBE, T_probe: float = 430.0,
        lifetime: int = 1, max_outer: int = MAX_OUTER, out_dir=None, temperature=None, log=None):
    """The self-consistent loop over every (omega, k) point of the k-points in
    dblocks (Engine.upload_blocks) and the energy grid E (ne odd; point ie = 0
    is the reference's ie = 1 and gets its zero-frequency treatment).

    Returns dict(temperature_i, temperature_ph, tdiff (per iteration),
    temperature_used (the probe temperatures of the last iteration's solve),
    T_LR, T_RL (nk, ne) of the last iteration, flux, gn, A (nk, ne, L),
    flux_left, flux_right (nk, ne), iterations)."""
    import torch
    E = np.asarray(E, float)
    ne, nk, L = len(E), dblocks["nk"], dblocks["L"]
    simpson_weights(np.sqrt(E))            # ne odd, before any GPU work
    temp = initial_temperatures(L, t_left, t_right) if temperature is None else np.array(temperature, float)
    t_ph = temp.copy()
    Ep = np.tile(E, nk)
    kidx = np.repeat(np.arange(nk, dtype=np.int32), ne)
    mask = np.zeros(nk * ne, np.int32)
    mask[::ne] = 1
    hist = []
    res = None
    used = temp
    for it in range(max_outer):
        used = temp
        r = engine.solve_buttiker(dblocks, Ep, p, temp, t_left, t_right, B_probe=B_probe, T_probe=T_probe,
                                  lifetime=lifetime, kidx=kidx if nk > 1 else None, first_point_mask=mask)
        bad = int(torch.count_nonzero(r["status"]).item())
        if bad:
            raise RuntimeError(f"{bad} points failed (first status {int(r['status'][r['status'] != 0][0])})")
        jac, flux = integrate_jacobian(r["jacobian"], r["flux"], E, nk)
        res = {k: r[k].cpu().numpy() for k in ("T_LR", "T_RL", "flux", "gn_sum", "a_sum", "flux_left", "flux_right")}
        shape3 = (nk, ne, L)
        fl, gn, A = (res[k].reshape(shape3) for k in ("flux", "gn_sum", "a_sum"))
        if out_dir:
            write_flux(out_dir, fl, gn, A, res["flux_left"].reshape(nk, ne), res["flux_right"].reshape(nk, ne))
        t_ph = get_ph_temperature(A.sum(axis=0) / nk, gn.sum(axis=0) / nk, E, t_ph)
        temp, tdiff = update_temperature(temp, jac, flux)
        hist.append(tdiff)
        if out_dir:
            write_temperature(out_dir, temp, t_ph)
        if log:
            log(f"probe iteration {it + 1}: |T_old - T_new| = {tdiff:.3e}")
        if tdiff < TDIFF_STOP:
            break
    return dict(temperature_i=temp, temperature_ph=t_ph, tdiff=hist, iterations=len(hist), temperature_used=used,
                T_LR=res["T_LR"].reshape(nk, ne), T_RL=res["T_RL"].reshape(nk, ne),
                flux=res["flux"].reshape(nk, ne, L), gn=res["gn_sum"].reshape(nk, ne, L),
                A=res["a_sum"].reshape(nk, ne, L), flux_left=res["flux_left"].reshape(nk, ne),
                flux_right=res["flux_right"].reshape(nk, ne))
