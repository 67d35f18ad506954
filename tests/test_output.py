"""write_trans output (negf/write_data.f90:340-497): the committed tables of
the si-si example are rewritten byte for byte from their own values."""
import os

import numpy as np

from phinfortran_amd import output as OUT

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "si_si")


def test_fortran_e_editing():
    assert OUT.fortran_e(1e-6) == "  0.1000000000000000E-05"
    assert OUT.fortran_e(0.0) == "  0.0000000000000000E+00"
    assert OUT.fortran_e(3.999984712361426) == "  0.3999984712361426E+01"
    assert OUT.fortran_e(-2.5e-30) == " -0.2500000000000000E-29"
    assert OUT.fortran_e(1e-100) == "  0.1000000000000000E-99"
    assert OUT.fortran_e(1e-101) == "  0.1000000000000000-100"
    assert len(OUT.fortran_e(123.456)) == 24


def test_write_trans_reproduces_committed_tables(tmp_path):
    tabs = {}
    for name in ("transl_tot", "transr_tot"):
        path = os.path.join(FIX, name + ".dat")
        tabs[name] = np.loadtxt(path)
    E = tabs["transl_tot"][:, 0]
    OUT.write_trans(str(tmp_path), E, tabs["transl_tot"][:, 1:], tabs["transr_tot"][:, 1:])
    for name in ("transl_tot", "transr_tot"):
        assert open(os.path.join(tmp_path, name + ".dat"), "rb").read() == \
            open(os.path.join(FIX, name + ".dat"), "rb").read(), name
