"""MI355X-native engine for phinfortran's per-(omega,k) transport solve.

    lattice     seeded block-Hamiltonian builders for the transport workloads
    transport   host mirror of the reference interface over the HIP C ABI
    build       in-tree build of libnegf_transport.so (gfx950)
"""
from . import lattice  # noqa: F401

__all__ = ["lattice", "transport", "build"]


def __getattr__(name):
    if name == "transport":
        from . import transport
        return transport
    if name == "build":
        from . import build
        return build
    raise AttributeError(name)
