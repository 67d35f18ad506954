"""ctypes binding of libnegf_transport.so (the C ABI in include/negf_transport.h).

The library is required: importing the solver API without it raises, there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnegf_transport.so")

NT_OK = 0
NT_ERR_NAN = -1
NT_ERR_NOT_CONVERGED = -2
NT_ERR_ARG = -3
NT_ERR_HIP = -4
NT_ERR_NOMEM = -5

NT_ETA_RELATIVE = 0
NT_ETA_ABSOLUTE = 1
NT_PREC_FP64 = 0
NT_PREC_MIXED = 1
NT_DOS_FROM_ETA = 0
NT_DOS_NEGF = 1
NT_DOS_TB = 2

NT_K_GEMM = 0
NT_K_INV = 1
NT_K_ELEM = 2
NT_K_EIG = 3

# every symbol the header declares
EXPORTS = [
    "nt_params_default", "nt_params_validate", "nt_create", "nt_destroy", "nt_sync", "nt_stream", "nt_set_stream",
    "nt_set_batch_limit", "nt_last_error", "nt_status_string", "nt_surface_green_tdb",
    "nt_get_G_full_tdb", "nt_G_block", "nt_solve_tdb_batched_dev", "nt_solve_tdb_batched",
    "nt_solve_tdb_batched_noalloc_dev",
    "nt_solve_tdb_adjoint_batched_dev", "nt_buttiker_tdb_batched_dev", "nt_get_ldos", "nt_modes_tdb_batched_dev", "nt_modes_unfold_tdb_batched_dev", "nt_unfold_pc_dev", "nt_unfold_pc_tb_dev", "nt_solve_tdb_multi_dev", "nt_hk_assemble_dev", "nt_hk_assemble_orb_dev", "nt_autotune_inverse", "nt_set_inverse_panel", "nt_autotune_gemm", "nt_set_gemm_tile",
    "nt_set_overlap", "nt_set_gemm_3m", "nt_set_workspace_poison", "nt_set_point_split",
    "nt_set_adjoint_checkpoint", "nt_adjoint_bytes_per_point", "nt_set_sweeps", "nt_set_decimation_check", "nt_set_multi_workers", "nt_profile_enable", "nt_profile_get", "nt_profile_reset", "nt_last_mean_iters", "nt_mixed_counts",
    "nt_zgemm_batched_dev", "nt_zinv_batched_dev", "nt_zgeev_batched_dev", "nt_zsqrtm_batched_dev",
]


class NtC128(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class NtModesOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("trl", "tlr", "rrr", "rll", "k_z_L_adv", "k_z_L_ret", "k_z_R_ret",
                                          "k_z_R_adv", "n_prop", "t_mode_L", "t_mode_R", "sums")]


class NtUnfoldDesc(C.Structure):
    _fields_ = [("n_bloch_x", C.c_int), ("n_bloch_y", C.c_int), ("recivec", C.c_double * 9),
                ("reci_uc_l", C.c_double * 9), ("reci_uc_r", C.c_double * 9)] + \
               [(n, C.c_void_p) for n in ("ksc", "pos_l", "pos_uc_l", "pos_r", "pos_uc_r")]


UNFOLD_OUT_FIELDS = ("bz_L_adv", "bz_L_ret", "bz_R_ret", "bz_R_adv", "k_xyz_L_adv", "k_xyz_L_ret", "k_xyz_R_ret",
                     "k_xyz_R_adv", "split", "spec_trl", "spec_tlr", "spec_rrr", "spec_rll")


class NtUnfoldOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in UNFOLD_OUT_FIELDS]


class NtProbeDesc(C.Structure):
    _fields_ = [("temperature", C.c_void_p), ("t_left", C.c_double), ("t_right", C.c_double),
                ("B_probe", C.c_double), ("T_probe", C.c_double), ("lifetime", C.c_int)]


PROBE_OUT_FIELDS = ("flux", "jacobian", "flux_left", "flux_right", "a_sum", "gn_sum")


class NtProbeOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in PROBE_OUT_FIELDS]


class NtDeviceDesc(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nl", "nr", "b", "L", "nk", "npts")] + \
               [(n, C.c_void_p) for n in ("Haml", "Hamlv", "Hamr", "Hamrv", "Hamd", "Hamdv", "Hamld", "Hamrd", "E",
                                          "kidx", "first_point_mask", "T_LR", "T_RL", "status")]


class NtParams(C.Structure):
    _fields_ = [
        ("eta", C.c_double),
        ("eta0", C.c_double),
        ("convergence", C.c_double),
        ("device_eta_factor", C.c_double),
        ("lambda_eta", C.c_double),
        ("eta_mode", C.c_int),
        ("max_iter", C.c_int),
        ("precision", C.c_int),
        ("circular", C.c_int),
        ("n_bloch_x", C.c_int),
        ("n_bloch_y", C.c_int),
        ("path_mode", C.c_int),
        ("dos_mode", C.c_int),
    ]


class NtSolveExtras(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in (
        "gl_adv", "gr_ret", "gl_plus", "gr_minus", "gml_adv", "gmr_ret", "sl_adv", "sr_ret",
        "G_D_LR", "G_D_RL", "G_L", "G_R", "sdos_l", "sdos_r", "ldos", "ddos")]


class TransportError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m phinfortran_amd.build` "
                          "(the transport path has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    I = C.c_int
    D = C.c_double
    pp = C.POINTER(NtParams)
    L.nt_params_default.argtypes = [pp]
    L.nt_params_default.restype = None
    L.nt_params_validate.argtypes = [pp, C.c_int, C.c_int]
    L.nt_params_validate.restype = C.c_int
    L.nt_create.argtypes = [C.POINTER(P), I]
    L.nt_destroy.argtypes = [P]
    L.nt_sync.argtypes = [P]
    L.nt_stream.argtypes = [P]
    L.nt_stream.restype = P
    L.nt_set_stream.argtypes = [P, P]
    L.nt_set_batch_limit.argtypes = [P, I]
    L.nt_last_error.argtypes = [P]
    L.nt_last_error.restype = C.c_char_p
    L.nt_status_string.argtypes = [I]
    L.nt_status_string.restype = C.c_char_p
    L.nt_surface_green_tdb.argtypes = [P, pp, I, I, D, P, P, P, P, P, P, P, P, P, P, P, P, P]
    L.nt_get_G_full_tdb.argtypes = [P, pp, I, I, I, I, I, D, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P]
    L.nt_G_block.argtypes = [P, pp, I, I, I, P, D, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P]
    solve_args = [P, pp, I, I, I, I, I, P, P, I, P, P, P, P, P, P, P, P, P, P, P, P, P, C.POINTER(NtSolveExtras)]
    L.nt_solve_tdb_batched_dev.argtypes = solve_args
    L.nt_solve_tdb_batched_noalloc_dev.argtypes = solve_args
    L.nt_solve_tdb_batched.argtypes = solve_args
    L.nt_solve_tdb_adjoint_batched_dev.argtypes = solve_args[:19] + [P, I, I, P, P, P, P, P, P]
    L.nt_buttiker_tdb_batched_dev.argtypes = solve_args[:2] + [C.POINTER(NtProbeDesc)] + solve_args[2:] + [
        C.POINTER(NtProbeOut)]
    L.nt_solve_tdb_multi_dev.argtypes = [P, pp, I, P]
    L.nt_hk_assemble_dev.argtypes = [P, I, P, I, P, P, I, P, P, I, I, P, I, C.c_longlong]
    L.nt_hk_assemble_orb_dev.argtypes = [P, I, P, I, P, P, I, I, P, P, I, I, P, I, C.c_longlong, I]
    L.nt_get_ldos.argtypes = [P, pp, I, I, I, P, D, P, P, P, P, P, P, P, P, I, P, P]
    L.nt_modes_tdb_batched_dev.argtypes = [P, pp, I, I, I, I, P, P, I] + [P] * 14 + [D, D, P,
                                                                              C.POINTER(NtModesOut), P]
    L.nt_unfold_pc_dev.argtypes = [P, I, P, P, I, C.POINTER(D), C.POINTER(D), I, I, P, P, I, P, P, P, P, P]
    L.nt_unfold_pc_tb_dev.argtypes = [P, I, P, P, I, C.POINTER(D), C.POINTER(D), I, P, I, P, P, P, P, P]
    L.nt_modes_unfold_tdb_batched_dev.argtypes = [P, pp, I, I, I, I, P, P, I] + [P] * 14 + [
        D, D, P, C.POINTER(NtModesOut), C.POINTER(NtUnfoldDesc), C.POINTER(NtUnfoldOut), P]
    L.nt_autotune_inverse.argtypes = [P, I, I, I, C.POINTER(I), C.POINTER(D)]
    L.nt_set_inverse_panel.argtypes = [P, I, I, I]
    L.nt_autotune_gemm.argtypes = [P, I, I, I, C.POINTER(I), C.POINTER(D)]
    L.nt_set_gemm_tile.argtypes = [P, I, I, I]
    L.nt_set_overlap.argtypes = [P, I]
    L.nt_set_workspace_poison.argtypes = [P, I]
    L.nt_set_point_split.argtypes = [P, I]
    L.nt_set_gemm_3m.argtypes = [P, I]
    L.nt_profile_enable.argtypes = [P, I]
    L.nt_profile_get.argtypes = [P, I, C.POINTER(C.c_longlong), C.POINTER(D), C.POINTER(D)]
    L.nt_profile_reset.argtypes = [P]
    L.nt_last_mean_iters.argtypes = [P]
    L.nt_last_mean_iters.restype = D
    L.nt_mixed_counts.argtypes = [P, P]
    L.nt_mixed_counts.restype = I
    L.nt_set_sweeps.argtypes = [P, I]
    L.nt_set_sweeps.restype = I
    L.nt_set_decimation_check.argtypes = [P, I]
    L.nt_set_decimation_check.restype = I
    L.nt_set_multi_workers.argtypes = [P, I]
    L.nt_set_multi_workers.restype = I
    L.nt_set_adjoint_checkpoint.argtypes = [P, I]
    L.nt_set_adjoint_checkpoint.restype = I
    L.nt_adjoint_bytes_per_point.argtypes = [P, I, I, I, I, I, I, P]
    L.nt_adjoint_bytes_per_point.restype = I
    LL = C.c_longlong
    L.nt_zgemm_batched_dev.argtypes = [P, C.c_char, C.c_char, I, I, I, NtC128, P, I, LL, P, I, LL, NtC128,
                                       P, I, LL, I]
    L.nt_zinv_batched_dev.argtypes = [P, I, P, LL, P, LL, P, I]
    L.nt_zgeev_batched_dev.argtypes = [P, I, P, LL, P, P, LL, P, I]
    L.nt_zsqrtm_batched_dev.argtypes = [P, I, P, LL, P, LL, P, I]
    _lib = L
    return L


def status_string(status: int) -> str:
    return lib().nt_status_string(int(status)).decode()
