"""MI355X-native engine for phinfortran's per-(omega,k) transport solve.

    lattice     seeded block-Hamiltonian builders for the transport workloads
    transport   host mirror of the reference interface over the HIP C ABI
    build       in-tree build of libnegf_transport.so (gfx950)
"""
import importlib

from . import lattice  # noqa: F401

__all__ = ["lattice", "transport", "build"]


def __getattr__(name):
    if name in ("transport", "build", "_native"):
        return importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
