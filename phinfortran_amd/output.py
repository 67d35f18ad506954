"""The reference's transmission output files (negf/write_data.f90:340-497).

write_trans writes ten tables, one row per energy: the first column is
E(i, 1) (the energy grid of the first k-point), then one value per k-point,
Fortran edit descriptor 1e24.16; values below 1e-60 are written as zero.
The *_s/_ns/refl files end each row with a blank (" "), the *_tot files
with nothing (the reference writes "(A)" of " " and of "" respectively).
"""
from __future__ import annotations

import math
import os
from typing import Optional

import numpy as np


def fortran_e(x: float, w: int = 24, d: int = 16) -> str:
    """Fortran Ew.d editing: 0.DDDD...E+XX right-justified in w columns
    (three-digit exponents drop the E as Fortran does)."""
    x = float(x)
    if math.isnan(x):                   # gfortran writes NaN / Infinity right-justified
        return "NaN".rjust(w)
    if math.isinf(x):
        body = "Infinity" if x > 0 else "-Infinity"
        return body.rjust(w) if len(body) <= w else "*" * w
    if x == 0.0:
        body = "0." + "0" * d + "E+00"
    else:
        sgn = "-" if x < 0 else ""
        m, e = ("%.*e" % (d - 1, abs(x))).split("e")
        digits = m.replace(".", "")
        ex = int(e) + 1
        if abs(ex) <= 99:
            body = "%s0.%sE%s%02d" % (sgn, digits, "+" if ex >= 0 else "-", abs(ex))
        else:
            body = "%s0.%s%s%03d" % (sgn, digits, "+" if ex >= 0 else "-", abs(ex))
    if len(body) > w:
        return "*" * w
    return body.rjust(w)


def _table(path: str, E1, T, tail: str, w: int = 24, d: int = 16, clamp: bool = True):
    T = np.asarray(T, float)
    with open(path, "w") as f:
        for i in range(T.shape[0]):
            row = [fortran_e(E1[i], w, d)] + [fortran_e(v if (v >= 1e-60 or not clamp or math.isnan(v)) else 0.0, w, d) for v in T[i]]
            f.write("".join(row) + tail + "\n")


def write_trans(output_dir: str, E, transl, transr, transl_s: Optional[np.ndarray] = None,
                transl_ns=None, transr_s=None, transr_ns=None, refll_s=None, refll_ns=None, reflr_s=None,
                reflr_ns=None, tb: bool = False):
    """Write the tables of write_trans.  E: (ne, nk) energy grid (or the
    first column); the T arrays are (ne, nk).  Tables passed as None are
    not written (the Caroli-only run produces only *_tot).  tb=True: the
    tight-binding driver's writer (tb/write_data.f90:362-470), edit
    descriptor 1e20.10 and values written as they are."""
    fmt = dict(w=20, d=10, clamp=False) if tb else {}
    os.makedirs(output_dir, exist_ok=True)
    E = np.asarray(E, float)
    E1 = E[:, 0] if E.ndim == 2 else E
    for name, T in (("transl_s", transl_s), ("transl_ns", transl_ns), ("refll_s", refll_s),
                    ("refll_ns", refll_ns)):
        if T is not None:
            _table(os.path.join(output_dir, name + ".dat"), E1, T, " ", **fmt)
    for name, T in (("transr_s", transr_s), ("transr_ns", transr_ns), ("reflr_s", reflr_s),
                    ("reflr_ns", reflr_ns)):
        if T is not None:
            _table(os.path.join(output_dir, name + ".dat"), E1, T, "", **fmt)
    _table(os.path.join(output_dir, "transl_tot.dat"), E1, transl, "", **fmt)
    _table(os.path.join(output_dir, "transr_tot.dat"), E1, transr, "", **fmt)


# ------------------------------------------------------------ per-mode tables
def _mode_row(count: int, vals) -> str:
    """One record of '(1I5,13e20.10)' / '(1I5,15e20.10)'."""
    return "%5d" % count + "".join(fortran_e(v, 20, 10) for v in vals)


def _clamp(x: float) -> float:
    return 1.0e-60 if 0.0 < x < 1.0e-60 else x


def specular_rows(omega, k_b, total, spec, kp_b=None, v_b=None):
    """write_specular's records for one (omega, k) point
    (negf/write_data.f90:36-73): one per incident propagating mode (k_b(i,3)
    < 99.9, numbered from 1): omega, k_b(i,1:3), kp_b(i,1:3), v_b(i,1:3),
    r_s, r_t, r_s / r_t (0 when r_t = 0); r_s and r_t below 1e-60 become
    1e-60.  total / spec: the (total, specular) sums of the mode
    (nt_unfold_out spec_*).  kp_b / v_b (the primitive-cell k and velocity of
    find_k_pc) default to zero."""
    k_b = np.asarray(k_b, float)
    n = len(k_b)
    kp_b = np.zeros((n, 3)) if kp_b is None else np.asarray(kp_b, float)
    v_b = np.zeros((n, 3)) if v_b is None else np.asarray(v_b, float)
    out, count = [], 0
    for i in range(n):
        if not k_b[i, 2] < 99.9:
            continue
        count += 1
        rt, rs = _clamp(float(total[i])), _clamp(float(spec[i]))
        out.append(_mode_row(count, [omega, *k_b[i], *kp_b[i], *v_b[i], rs, rt, rs / rt if rt != 0.0 else 0.0]))
    return out


def md_rows(omega, k_b, r_total, r_spec, t_total, t_spec, kp_b=None, v_b=None):
    """write_md's records (negf/write_data.f90:113-171): omega, k_b(1:3),
    kp_b(1:3), v_b(1:3), r_s, r_t, t_s, t_t per incident propagating mode."""
    k_b = np.asarray(k_b, float)
    n = len(k_b)
    kp_b = np.zeros((n, 3)) if kp_b is None else np.asarray(kp_b, float)
    v_b = np.zeros((n, 3)) if v_b is None else np.asarray(v_b, float)
    out, count = [], 0
    for i in range(n):
        if not k_b[i, 2] < 99.9:
            continue
        count += 1
        vals = [_clamp(float(x)) for x in (r_spec[i], r_total[i], t_spec[i], t_total[i])]
        out.append(_mode_row(count, [omega, *k_b[i], *kp_b[i], *v_b[i], *vals]))
    return out


def write_mode_tables(output_dir: str, nn: int, omega: float, point: dict):
    """Append one (omega, k) point to the six per-mode files of vel_mat_tdb
    (negf/velocity_mat.f90:3101-3130): {tlr,trl,right,left,mdl,mdr}_specular_param_NNNNNN.dat,
    NNNNNN the 1-based k index nn.  point: host arrays of one point from
    Engine.solve_modes(..., unfold=...): k_xyz_L_adv / k_xyz_R_adv (n x 3)
    and spec_trl, spec_tlr, spec_rrr, spec_rll (n x 2: total, specular);
    optional kp_L_adv, v_L_adv, kp_R_adv, v_R_adv (n x 3, solve_modes with
    pc=...) fill the primitive-cell columns, zero otherwise."""
    os.makedirs(output_dir, exist_ok=True)
    kl, kr = point["k_xyz_L_adv"], point["k_xyz_R_adv"]
    pl = dict(kp_b=point.get("kp_L_adv"), v_b=point.get("v_L_adv"))
    pr = dict(kp_b=point.get("kp_R_adv"), v_b=point.get("v_R_adv"))
    s = {k: np.asarray(point["spec_" + k]) for k in ("trl", "tlr", "rrr", "rll")}
    tables = {
        "tlr": specular_rows(omega, kr, s["tlr"][:, 0], s["tlr"][:, 1], **pr),
        "trl": specular_rows(omega, kl, s["trl"][:, 0], s["trl"][:, 1], **pl),
        "right": specular_rows(omega, kr, s["rrr"][:, 0], s["rrr"][:, 1], **pr),
        "left": specular_rows(omega, kl, s["rll"][:, 0], s["rll"][:, 1], **pl),
        "mdl": md_rows(omega, kl, s["rll"][:, 0], s["rll"][:, 1], s["trl"][:, 0], s["trl"][:, 1], **pl),
        "mdr": md_rows(omega, kr, s["rrr"][:, 0], s["rrr"][:, 1], s["tlr"][:, 0], s["tlr"][:, 1], **pr),
    }
    for name, rows in tables.items():
        with open(os.path.join(output_dir, "%s_specular_param_%06d.dat" % (name, nn)), "a") as f:
            for r in rows:
                f.write(r + "\n")
