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


# ------------------------------------------------ tight-binding per-mode tables
def specular_rows_tb(E, k_f, total, spec, kp_f=None, v_f=None):
    """The tight-binding write_specular's records for one (E, k) point
    (tb/write_data.f90:3-67): one per incident propagating mode (k_f(i,3) <
    99.9), '(13e20.10)': E, k_f(i,1:3), kp_f(i,1:3), v_f(i,1:3), r_s, r_t,
    r_s / r_t (0 when r_t = 0); no mode number and no 1e-60 floor."""
    k_f = np.asarray(k_f, float)
    n = len(k_f)
    kp_f = np.zeros((n, 3)) if kp_f is None else np.asarray(kp_f, float)
    v_f = np.zeros((n, 3)) if v_f is None else np.asarray(v_f, float)
    out = []
    for i in range(n):
        if not k_f[i, 2] < 99.9:
            continue
        rt, rs = float(total[i]), float(spec[i])
        vals = [E, *k_f[i], *kp_f[i], *v_f[i], rs, rt, rs / rt if rt != 0.0 else 0.0]
        out.append("".join(fortran_e(v, 20, 10) for v in vals))
    return out


def md2md_rows_tb(E, M, k_f, k_b, kp_f=None, kp_b=None, v_f=None, v_b=None):
    """The tight-binding write_md2md's records (tb/write_data.f90:166-239):
    for every propagating outgoing mode i (k_f, numbered count1 from 1) and
    propagating incident mode j (k_b, count2 running over all pairs), '(2I5,
    21e20.10,2I5)': count1, count2, E, k_f(i), k_b(j), kp_f(i), kp_b(j),
    v_f(i), v_b(j), ts, tns, i, j (1-based), |M(i, j)|^2 in ts when the
    in-plane k agree (|dk|^2 < 1e-5), else in tns.  M: outgoing x incident."""
    k_f, k_b = np.asarray(k_f, float), np.asarray(k_b, float)
    z = lambda a, n: np.zeros((n, 3)) if a is None else np.asarray(a, float)
    kp_f, kp_b, v_f, v_b = z(kp_f, len(k_f)), z(kp_b, len(k_b)), z(v_f, len(k_f)), z(v_b, len(k_b))
    out, c1, c2 = [], 0, 0
    for i in range(len(k_f)):
        if not k_f[i, 2] < 99.9:
            continue
        c1 += 1
        for j in range(len(k_b)):
            if not k_b[j, 2] < 99.9:
                continue
            c2 += 1
            d = k_f[i] - k_b[j]
            d[2] = 0.0
            a = abs(M[i, j]) ** 2
            ts, tns = (a, 0.0) if d @ d < 1.0e-5 else (0.0, a)
            vals = [E, *k_f[i], *k_b[j], *kp_f[i], *kp_b[j], *v_f[i], *v_b[j], ts, tns]
            out.append("%5d%5d" % (c1, c2) + "".join(fortran_e(v, 20, 10) for v in vals) + "%5d%5d" % (i + 1, j + 1))
    return out


def write_mode_tables_tb(output_dir: str, nn: int, E: float, point: dict, md2md: bool = False):
    """Append one (E, k) point to the per-mode files of the tight-binding
    vel_mat (tb/velocity_mat.f90:1511-1544): {left,right,trl,tlr}_specular_
    param_NNNNNN.dat (write_specular of r_LL, r_RR, t_RL, t_LR per incident
    mode), {mdl,mdr}_specular_param_NNNNNN.dat (write_md, the negf format)
    and, with md2md (verbosity = 1), {tlr,trl,rrr,rll}_md2md_NNNNNN.dat.
    point: host arrays of one point from Engine.solve_modes(..., unfold=...,
    pc=...): k_xyz_* and kp_* / v_* (n x 3) of the four mode sets, spec_*
    (n x 2: total, specular), and for md2md the t/r matrices (outgoing x
    incident)."""
    os.makedirs(output_dir, exist_ok=True)
    k = {s: point["k_xyz_" + s] for s in ("L_adv", "L_ret", "R_ret", "R_adv")}
    kp = {s: point.get("kp_" + s) for s in k}
    v = {s: point.get("v_" + s) for s in k}
    sp = {m: np.asarray(point["spec_" + m]) for m in ("trl", "tlr", "rrr", "rll")}
    tables = {
        "left_specular_param": specular_rows_tb(E, k["L_adv"], sp["rll"][:, 0], sp["rll"][:, 1], kp["L_adv"], v["L_adv"]),
        "right_specular_param": specular_rows_tb(E, k["R_adv"], sp["rrr"][:, 0], sp["rrr"][:, 1], kp["R_adv"],
                                                 v["R_adv"]),
        "trl_specular_param": specular_rows_tb(E, k["L_adv"], sp["trl"][:, 0], sp["trl"][:, 1], kp["L_adv"], v["L_adv"]),
        "tlr_specular_param": specular_rows_tb(E, k["R_adv"], sp["tlr"][:, 0], sp["tlr"][:, 1], kp["R_adv"], v["R_adv"]),
        "mdl_specular_param": md_rows(E, k["L_adv"], sp["rll"][:, 0], sp["rll"][:, 1], sp["trl"][:, 0],
                                      sp["trl"][:, 1], kp_b=kp["L_adv"], v_b=v["L_adv"]),
        "mdr_specular_param": md_rows(E, k["R_adv"], sp["rrr"][:, 0], sp["rrr"][:, 1], sp["tlr"][:, 0],
                                      sp["tlr"][:, 1], kp_b=kp["R_adv"], v_b=v["R_adv"]),
    }
    if md2md:
        tables["tlr_md2md"] = md2md_rows_tb(E, point["tlr"], k["L_ret"], k["R_adv"], kp["L_ret"], kp["R_adv"],
                                            v["L_ret"], v["R_adv"])
        tables["trl_md2md"] = md2md_rows_tb(E, point["trl"], k["R_ret"], k["L_adv"], kp["R_ret"], kp["L_adv"],
                                            v["R_ret"], v["L_adv"])
        tables["rrr_md2md"] = md2md_rows_tb(E, point["rrr"], k["R_ret"], k["R_adv"], kp["R_ret"], kp["R_adv"],
                                            v["R_ret"], v["R_adv"])
        tables["rll_md2md"] = md2md_rows_tb(E, point["rll"], k["L_ret"], k["L_adv"], kp["L_ret"], kp["L_adv"],
                                            v["L_ret"], v["L_adv"])
    for name, rows in tables.items():
        with open(os.path.join(output_dir, "%s_%06d.dat" % (name, nn)), "a") as f:
            for r in rows:
                f.write(r + "\n")
