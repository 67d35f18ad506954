"""ctypes binding of the CPU oracle (oracle/nt_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP engine and the CPU baseline
of bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module; the product path (phinfortran_amd) never does.

All arrays are numpy complex128 in Fortran (column-major) order, matching the
reference's complex(kind=8) arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

ETA_RELATIVE = 0
ETA_ABSOLUTE = 1


class CDouble(C.Structure):
    """double _Complex passed by value (two SSE registers, like a C struct of two doubles)."""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class OrParams(C.Structure):
    _fields_ = [
        ("eta", C.c_double),
        ("eta0", C.c_double),
        ("convergence", C.c_double),
        ("device_eta_factor", C.c_double),
        ("eta_mode", C.c_int),
        ("max_iter", C.c_int),
    ]


def phonon_params(eta=1e-6, eta0=1e-6, convergence=1e-12, max_iter=200):
    """negf defaults (config.f90:88-89): eta = eta0 = 1e-6, device eta x10."""
    return OrParams(eta, eta0, convergence, 10.0, ETA_RELATIVE, max_iter)


def tb_params(eta=1e-8, convergence=1e-12, max_iter=200):
    """tb conventions: constant eta in leads and device (tb/surface.f90:218, tb/device.f90:354)."""
    return OrParams(eta, 0.0, convergence, 1.0, ETA_ABSOLUTE, max_iter)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        _lib.or_inv.argtypes = [C.c_int, P, P]
        _lib.or_decimate.argtypes = [C.c_int, P, P, P, CDouble, C.c_double, C.c_int, P, C.POINTER(C.c_int)]
        _lib.or_surface_green.argtypes = [C.c_int, C.c_int, P, P, P, P, C.c_double, C.POINTER(OrParams), C.c_int,
                                          P, P, P, P, P, P, P, P, P]
        _lib.or_rgf.argtypes = [C.c_int, P, P, P, P, C.c_int, C.c_int, P, P, P, P, P,
                                P, P, P, P, P, P, C.c_double, C.POINTER(OrParams), C.c_int,
                                P, P, P, P, P, P]
        _lib.or_solve_tdb.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, P, P, P, P, P,
                                      C.c_double, C.POINTER(OrParams), C.c_int, P, P, P]
        _lib.or_solve_tdb_batch.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, P, P, P, P, P,
                                            C.c_int, P, C.POINTER(OrParams), C.c_int, P, P, P]
        _lib.or_solve_tdb_batch_kappa.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, P, P, P, P, P,
                                                  C.c_int, P, C.POINTER(OrParams), C.c_int, P, P, P, P]
        _lib.or_circular_g.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, CDouble, C.c_double, C.c_int,
                                       P, C.POINTER(C.c_int)]
        _lib.or_get_ldos.argtypes = [C.c_int, P, P, P, P, C.c_int, C.c_int, P, P, P, P, P, C.c_double,
                                     C.POINTER(OrParams), C.c_int, P, P]
    return _lib


def _f(a):
    """complex128 Fortran-ordered copy (the oracle's layout)."""
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _z(z):
    return CDouble(float(np.real(z)), float(np.imag(z)))


class OracleError(RuntimeError):
    def __init__(self, status, what):
        super().__init__(f"{what}: oracle status {status}")
        self.status = status


def inv(A):
    A = _f(A)
    n = A.shape[0]
    out = np.zeros((n, n), np.complex128, order="F")
    st = lib().or_inv(n, _p(A), _p(out))
    if st:
        raise OracleError(st, "or_inv")
    return out


def decimate(H00, alpha0, beta0, z, conv, max_iter=500):
    H00, alpha0, beta0 = _f(H00), _f(alpha0), _f(beta0)
    n = H00.shape[0]
    g = np.zeros((n, n), np.complex128, order="F")
    it = C.c_int(0)
    st = lib().or_decimate(n, _p(H00), _p(alpha0), _p(beta0), _z(z), conv, max_iter, _p(g), C.byref(it))
    if st:
        raise OracleError(st, "or_decimate")
    return g, it.value


def circular_g(H00, alpha0, nby, nbx, z, conv, max_iter=500):
    H00, alpha0 = _f(H00), _f(alpha0)
    n = H00.shape[0]
    g = np.zeros((n, n), np.complex128, order="F")
    it = C.c_int(0)
    st = lib().or_circular_g(n, nby, nbx, _p(H00), _p(alpha0), None, _z(z), conv, max_iter, _p(g), C.byref(it))
    if st:
        raise OracleError(st, "or_circular_g")
    return g, it.value


def surface_green(Haml, Hamlv, Hamr, Hamrv, E, params, want_plus_minus=True):
    """surface_green_tdb: returns dict with gl_adv, gr_ret, gl_plus, gr_minus,
    gml_adv, gmr_ret, sl_adv, sr_ret and the four iteration counts."""
    Haml, Hamlv, Hamr, Hamrv = map(_f, (Haml, Hamlv, Hamr, Hamrv))
    nl, nr = Haml.shape[0], Hamr.shape[0]
    names_l = ["gl_adv", "gl_plus", "gml_adv", "sl_adv"]
    names_r = ["gr_ret", "gr_minus", "gmr_ret", "sr_ret"]
    out = {k: np.zeros((nl, nl), np.complex128, order="F") for k in names_l}
    out.update({k: np.zeros((nr, nr), np.complex128, order="F") for k in names_r})
    iters = (C.c_int * 4)()
    st = lib().or_surface_green(nl, nr, _p(Haml), _p(Hamlv), _p(Hamr), _p(Hamrv), float(E), C.byref(params),
                                int(want_plus_minus),
                                _p(out["gl_adv"]), _p(out["gr_ret"]), _p(out["gl_plus"]), _p(out["gr_minus"]),
                                _p(out["gml_adv"]), _p(out["gmr_ret"]), _p(out["sl_adv"]), _p(out["sr_ret"]),
                                C.cast(iters, C.c_void_p))
    if st:
        raise OracleError(st, "or_surface_green")
    out["iters"] = list(iters)
    return out


def rgf(Hd, Hdv, Hld, Hrd, Haml, Hamr, gl, gr, gml, gmr, sl, sr, E, params,
        Hdv_back=None, Hrd_fwd=None, zero_transmission=False):
    """RGF sweep + Caroli on (possibly variable-size) blocks.  gl/gml/sl are the
    retarded left quantities (the reference passes gl_adv^+, -gml_adv, sl_adv^+)."""
    L = len(Hd)
    Hd = [_f(h) for h in Hd]
    Hdv = [_f(h) for h in Hdv]
    bsz = (C.c_int * L)(*[h.shape[0] for h in Hd])
    pHd = (C.c_void_p * L)(*[h.ctypes.data for h in Hd])
    pHdv = (C.c_void_p * max(L - 1, 1))(*[h.ctypes.data for h in Hdv])
    pback = None
    if Hdv_back is not None:
        Hdv_back = [_f(h) for h in Hdv_back]
        pback = (C.c_void_p * max(L - 1, 1))(*[h.ctypes.data for h in Hdv_back])
    Hld, Hrd, Haml, Hamr, gl, gr, gml, gmr, sl, sr = map(_f, (Hld, Hrd, Haml, Hamr, gl, gr, gml, gmr, sl, sr))
    Hrd_fwd = None if Hrd_fwd is None else _f(Hrd_fwd)
    nl, nr = Haml.shape[0], Hamr.shape[0]
    G_D_LR = np.zeros((nl, nr), np.complex128, order="F")
    G_D_RL = np.zeros((nr, nl), np.complex128, order="F")
    G_L = np.zeros((nl, nl), np.complex128, order="F")
    G_R = np.zeros((nr, nr), np.complex128, order="F")
    tl, tr = C.c_double(0), C.c_double(0)
    st = lib().or_rgf(L, C.cast(bsz, C.c_void_p), C.cast(pHd, C.c_void_p), C.cast(pHdv, C.c_void_p),
                      None if pback is None else C.cast(pback, C.c_void_p), nl, nr,
                      _p(Hld), _p(Hrd), _p(Hrd_fwd), _p(Haml), _p(Hamr), _p(gl), _p(gr), _p(gml), _p(gmr),
                      _p(sl), _p(sr), float(E), C.byref(params), int(zero_transmission),
                      _p(G_D_LR), _p(G_D_RL), _p(G_L), _p(G_R), C.cast(C.byref(tl), C.c_void_p),
                      C.cast(C.byref(tr), C.c_void_p))
    if st:
        raise OracleError(st, "or_rgf")
    return dict(G_D_LR=G_D_LR, G_D_RL=G_D_RL, G_L=G_L, G_R=G_R, T_LR=tl.value, T_RL=tr.value)


def solve_tdb(blocks, E, params, zero_transmission=False):
    """Full (omega,k)-solve on a tdb block set (dict with Haml, Hamlv, Hamr, Hamrv,
    Hamd (L,b,b), Hamdv (L-1,b,b), Hamld (nl,b), Hamrd (nr,b)).  Returns
    (T_LR, T_RL, mean decimation iterations)."""
    nl, b, L, nr = _dims(blocks)
    arrs = _pack(blocks)
    tl, tr, it = C.c_double(0), C.c_double(0), C.c_double(0)
    st = lib().or_solve_tdb(nl, nr, b, L, *[_p(a) for a in arrs], float(E), C.byref(params),
                            int(zero_transmission), C.cast(C.byref(tl), C.c_void_p),
                            C.cast(C.byref(tr), C.c_void_p), C.cast(C.byref(it), C.c_void_p))
    if st:
        raise OracleError(st, "or_solve_tdb")
    return tl.value, tr.value, it.value


def solve_tdb_batch(blocks, E, params, nthreads=0):
    nl, b, L, nr = _dims(blocks)
    arrs = _pack(blocks)
    E = np.ascontiguousarray(E, dtype=np.float64)
    tl = np.zeros(E.size)
    tr = np.zeros(E.size)
    status = np.zeros(E.size, np.int32)
    lib().or_solve_tdb_batch(nl, nr, b, L, *[_p(a) for a in arrs], E.size, _p(E), C.byref(params),
                             int(nthreads), _p(tl), _p(tr), _p(status))
    return tl, tr, status


def solve_tdb_batch_kappa(blocks, E, params, nthreads=0):
    """solve_tdb_batch plus, per point, the largest 1-norm condition number of
    the blocks the solve inverted: two correct FP64 algorithms may differ by
    about eps times it."""
    nl, b, L, nr = _dims(blocks)
    arrs = _pack(blocks)
    E = np.ascontiguousarray(E, dtype=np.float64)
    tl, tr, kappa = np.zeros(E.size), np.zeros(E.size), np.zeros(E.size)
    status = np.zeros(E.size, np.int32)
    lib().or_solve_tdb_batch_kappa(nl, nr, b, L, *[_p(a) for a in arrs], E.size, _p(E), C.byref(params),
                                   int(nthreads), _p(tl), _p(tr), _p(status), _p(kappa))
    return tl, tr, status, kappa


def _dims(blocks):
    Hamd = blocks["Hamd"]
    return blocks["Haml"].shape[0], Hamd.shape[1], Hamd.shape[0], blocks["Hamr"].shape[0]


def _pack(blocks):
    """Layer-major Fortran layout: Hamd as b x b x L with each layer contiguous."""
    Hamd = np.asarray(blocks["Hamd"], np.complex128)
    L, b = Hamd.shape[0], Hamd.shape[1]
    Hamdv = np.asarray(blocks["Hamdv"], np.complex128)
    hd = np.asfortranarray(np.transpose(Hamd, (1, 2, 0)))
    hdv = np.asfortranarray(np.transpose(Hamdv[: max(L - 1, 0)], (1, 2, 0))) if L > 1 else np.zeros((b, b, 1), np.complex128, order="F")
    return [_f(blocks["Haml"]), _f(blocks["Hamlv"]), _f(blocks["Hamr"]), _f(blocks["Hamrv"]),
            hd, hdv, _f(blocks["Hamld"]), _f(blocks["Hamrd"])]


def get_ldos(Hd, Hdv, Hld, Hrd, gl, gr, E, params, Hdv_back=None, Hrd_fwd=None, maxlen=None):
    """get_ldos (tb/device.f90:610-855): (ldos (L, maxlen), ddos)."""
    L = len(Hd)
    Hd = [_f(h) for h in Hd]
    Hdv = [_f(h) for h in Hdv]
    bsz = (C.c_int * L)(*[h.shape[0] for h in Hd])
    maxlen = maxlen or max(h.shape[0] for h in Hd)
    pHd = (C.c_void_p * L)(*[h.ctypes.data for h in Hd])
    pHdv = (C.c_void_p * max(L - 1, 1))(*[h.ctypes.data for h in Hdv])
    pback = None
    if Hdv_back is not None:
        Hdv_back = [_f(h) for h in Hdv_back]
        pback = (C.c_void_p * max(L - 1, 1))(*[h.ctypes.data for h in Hdv_back])
    Hld, Hrd, gl, gr = map(_f, (Hld, Hrd, gl, gr))
    Hrd_fwd = None if Hrd_fwd is None else _f(Hrd_fwd)
    nl, nr = gl.shape[0], gr.shape[0]
    ldos = np.zeros((L, maxlen), order="F")
    ddos = C.c_double(0)
    st = lib().or_get_ldos(L, C.cast(bsz, C.c_void_p), C.cast(pHd, C.c_void_p), C.cast(pHdv, C.c_void_p),
                           None if pback is None else C.cast(pback, C.c_void_p), nl, nr,
                           _p(Hld), _p(Hrd), _p(Hrd_fwd), _p(gl), _p(gr), float(E), C.byref(params),
                           int(maxlen), ldos.ctypes.data_as(C.c_void_p), C.cast(C.byref(ddos), C.c_void_p))
    if st:
        raise OracleError(st, "or_get_ldos")
    return np.ascontiguousarray(ldos), ddos.value
