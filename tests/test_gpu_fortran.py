"""The Fortran caller (ISO_C_BINDING module) drives the GPU engine."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fortran_caller():
    exe = os.path.join(ROOT, "tests", "fortran", "fortran_abi_check.x")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "FORTRAN_ABI_OK" in r.stdout
