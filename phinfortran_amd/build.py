"""Builds the HIP transport library in-tree: phinfortran_amd/libnegf_transport.so.

Every .hip file under csrc/ is compiled for gfx950 with packed-FP32 code
generation disabled, then linked into one shared library that exports the C ABI
of include/negf_transport.h.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnegf_transport.so")
BUILD_DIR = os.path.join(HERE, "_build")
ARCH = os.environ.get("NT_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(h):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build phinfortran_amd")
    return h


def _compile(src):
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    cmd = [_hipcc()] + HIPCC_FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link the library.  Always rebuilds."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB




def build_fortran(verbose: bool = False) -> str:
    """Compile the Fortran interface module and the Fortran caller check
    (tests/fortran/fortran_abi_check.x) against the built library."""
    root = os.path.dirname(HERE)
    fc = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    if not os.path.exists(fc):
        raise RuntimeError("flang not found")
    mod_dir = os.path.join(BUILD_DIR, "fmod")
    os.makedirs(mod_dir, exist_ok=True)
    exe = os.path.join(root, "tests", "fortran", "fortran_abi_check.x")
    cmd = [fc, "-O2", "-module-dir", mod_dir, os.path.join(root, "include", "negf_transport.f90"),
           os.path.join(root, "tests", "fortran", "fortran_abi_check.f90"), "-o", exe,
           "-L" + HERE, "-lnegf_transport", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../../phinfortran_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=BUILD_DIR)
    if r.returncode != 0:
        raise RuntimeError(f"flang failed:\n{r.stderr}")
    if verbose:
        print(f"built {exe}", file=sys.stderr)
    return exe


if __name__ == "__main__":
    build(verbose=True)
    build_fortran(verbose=True)
