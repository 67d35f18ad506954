"""CPU checks of the drop-in boundary: the built library loads and exports every
symbol include/negf_transport.h declares, the Fortran interface module binds
exactly those names, and argument validation rejects bad input without a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "negf_transport.h")
FMOD = os.path.join(ROOT, "include", "negf_transport.f90")
LIB = os.path.join(ROOT, "phinfortran_amd", "libnegf_transport.so")


def header_functions():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(nt_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_entry_points():
    names = header_functions()
    for must in ("nt_create", "nt_surface_green_tdb", "nt_get_G_full_tdb", "nt_G_block",
                 "nt_solve_tdb_batched", "nt_solve_tdb_batched_dev", "nt_zgemm_batched_dev", "nt_zinv_batched_dev"):
        assert must in names


def test_library_exports_every_header_symbol():
    assert os.path.exists(LIB), "build the library first (python -m phinfortran_amd.build)"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from phinfortran_amd import _native
    assert set(header_functions()) <= set(_native.EXPORTS) | {"nt_zgemm_batched_dev", "nt_zinv_batched_dev"}


def test_fortran_module_binds_header_names():
    src = open(FMOD).read()
    bound = set(re.findall(r'bind\(C,\s*name="(nt_\w+)"\)', src))
    assert bound, "no bind(C) interfaces"
    assert bound <= set(header_functions())


def test_fortran_module_and_caller_compile(tmp_path):
    from phinfortran_amd import build as B
    exe = B.build_fortran()
    assert os.path.exists(exe)


def test_argument_validation_without_gpu():
    """Entry points reject bad arguments before touching the device."""
    from phinfortran_amd import _native as N
    L = N.lib()
    p = N.NtParams()
    L.nt_params_default(ctypes.byref(p))
    assert p.eta == 1e-6 and p.eta0 == 1e-6 and p.device_eta_factor == 10.0 and p.max_iter > 0
    # null context
    assert L.nt_surface_green_tdb(None, ctypes.byref(p), 4, 4, 0.1, *([None] * 13)) == N.NT_ERR_ARG
    assert L.nt_sync(None) == N.NT_ERR_ARG
    assert L.nt_status_string(N.NT_ERR_NOT_CONVERGED).decode().startswith("surface decimation")
    assert L.nt_status_string(3).decode() == "Matrix is numerically singular!"


def test_circular_sizes_validated_without_gpu():
    """circular decimation needs n divisible by n_bloch_x * n_bloch_y"""
    from phinfortran_amd import _native as N
    L = N.lib()
    p = N.NtParams()
    L.nt_params_default(ctypes.byref(p))
    assert p.circular == 0 and p.n_bloch_x == 1 and p.n_bloch_y == 1
    assert L.nt_params_validate(ctypes.byref(p), 12, 12) == N.NT_OK
    p.circular, p.n_bloch_x, p.n_bloch_y = 1, 3, 2
    assert L.nt_params_validate(ctypes.byref(p), 12, 12) == N.NT_OK
    assert L.nt_params_validate(ctypes.byref(p), 12, 8) == N.NT_ERR_ARG
    assert L.nt_params_validate(ctypes.byref(p), 9, 9) == N.NT_ERR_ARG
    p.n_bloch_x = 0
    assert L.nt_params_validate(ctypes.byref(p), 12, 12) == N.NT_ERR_ARG
    p.n_bloch_x, p.precision = 1, 7
    assert L.nt_params_validate(ctypes.byref(p), 12, 12) == N.NT_ERR_ARG


def header_param_counts():
    """name -> number of parameters of every function declared in the header"""
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(nt_\w+)\s*\(([^;]*?)\)\s*;", src, flags=re.M | re.S):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_python_argtypes_match_header():
    """every ctypes binding declares exactly the header's number of parameters"""
    from phinfortran_amd import _native
    L = _native.lib()
    counts = header_param_counts()
    assert counts, "no declarations parsed"
    for name, k in counts.items():
        f = getattr(L, name)
        if f.argtypes is not None:
            assert len(f.argtypes) == k, (name, len(f.argtypes), k)


def test_fortran_interfaces_match_header():
    """every bind(C) interface of the Fortran module has the header's parameter count"""
    src = open(FMOD).read()
    counts = header_param_counts()
    found = list(re.finditer(r"(?:function|subroutine)\s+\w+\s*\(((?:[^()]|&\s*\n)*?)\)\s*(?:&\s*\n\s*)?"
                             r"bind\(C,\s*name=\"(nt_\w+)\"\)", src, flags=re.S))
    assert len(found) == len(re.findall(r'bind\(C,\s*name="nt_', src))
    for m in found:
        args = [a for a in re.sub(r"&\s*\n", " ", m.group(1)).replace("&", " ").split(",") if a.strip()]
        assert len(args) == counts[m.group(2)], (m.group(2), len(args), counts[m.group(2)])


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """without the built HIP library the product path raises instead of
    falling back to anything on the CPU"""
    from phinfortran_amd import _native as N
    from phinfortran_amd import transport as T
    monkeypatch.setattr(N, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(N, "_lib", None)
    with pytest.raises(ImportError, match="no CPU fallback"):
        N.lib()
    with pytest.raises(ImportError):
        T.params("tb")


def test_dos_mode_default_and_validation_without_gpu():
    """The surface-DOS convention is an explicit parameter: the default follows
    eta_mode, out-of-range values are rejected, and the Python parameter
    builder names the convention of its driver family."""
    from phinfortran_amd import _native as N
    from phinfortran_amd import transport as T
    L = N.lib()
    p = N.NtParams()
    L.nt_params_default(ctypes.byref(p))
    assert p.dos_mode == N.NT_DOS_FROM_ETA
    for m in (N.NT_DOS_FROM_ETA, N.NT_DOS_NEGF, N.NT_DOS_TB):
        p.dos_mode = m
        assert L.nt_params_validate(ctypes.byref(p), 4, 4) == N.NT_OK
    for m in (-1, 3):
        p.dos_mode = m
        assert L.nt_params_validate(ctypes.byref(p), 4, 4) == N.NT_ERR_ARG
    assert T.params("phonon").dos_mode == N.NT_DOS_NEGF
    assert T.params("tb").dos_mode == N.NT_DOS_TB
    # layout of the struct seen from Python equals the C one
    assert ctypes.sizeof(N.NtParams) == 5 * 8 + 8 * 4
