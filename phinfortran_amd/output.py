"""The reference's transmission output files (negf/write_data.f90:340-497).

write_trans writes ten tables, one row per energy: the first column is
E(i, 1) (the energy grid of the first k-point), then one value per k-point,
Fortran edit descriptor 1e24.16; values below 1e-60 are written as zero.
The *_s/_ns/refl files end each row with a blank (" "), the *_tot files
with nothing (the reference writes "(A)" of " " and of "" respectively).
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np


def fortran_e(x: float, w: int = 24, d: int = 16) -> str:
    """Fortran Ew.d editing: 0.DDDD...E+XX right-justified in w columns
    (three-digit exponents drop the E as Fortran does)."""
    x = float(x)
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


def _table(path: str, E1, T, tail: str):
    T = np.asarray(T, float)
    with open(path, "w") as f:
        for i in range(T.shape[0]):
            row = [fortran_e(E1[i])] + [fortran_e(v if v >= 1e-60 else 0.0) for v in T[i]]
            f.write("".join(row) + tail + "\n")


def write_trans(output_dir: str, E, transl, transr, transl_s: Optional[np.ndarray] = None,
                transl_ns=None, transr_s=None, transr_ns=None, refll_s=None, refll_ns=None, reflr_s=None,
                reflr_ns=None):
    """Write the tables of write_trans.  E: (ne, nk) energy grid (or the
    first column); the T arrays are (ne, nk).  Tables passed as None are
    not written (the Caroli-only run produces only *_tot)."""
    os.makedirs(output_dir, exist_ok=True)
    E = np.asarray(E, float)
    E1 = E[:, 0] if E.ndim == 2 else E
    for name, T in (("transl_s", transl_s), ("transl_ns", transl_ns), ("refll_s", refll_s),
                    ("refll_ns", refll_ns)):
        if T is not None:
            _table(os.path.join(output_dir, name + ".dat"), E1, T, " ")
    for name, T in (("transr_s", transr_s), ("transr_ns", transr_ns), ("reflr_s", reflr_s),
                    ("reflr_ns", reflr_ns)):
        if T is not None:
            _table(os.path.join(output_dir, name + ".dat"), E1, T, "")
    _table(os.path.join(output_dir, "transl_tot.dat"), E1, transl, "")
    _table(os.path.join(output_dir, "transr_tot.dat"), E1, transr, "")
