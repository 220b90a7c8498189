"""ctypes mirror of include/muram_step.h and the loader of the HIP library.

The library is the product: ``muram_main_amd/libmuram_mi355x.so`` built from
``muram_main_amd/csrc`` for gfx950.  ``load()`` fails loudly when it is absent;
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmuram_mi355x.so")

MRM_OK = 0
ERRORS = {-1: "MRM_ERR_ARG", -2: "MRM_ERR_HIP", -3: "MRM_ERR_EOS", -4: "MRM_ERR_INVALID_DT",
          -5: "MRM_ERR_CFL", -6: "MRM_ERR_STATE", -7: "MRM_ERR_UNSUPPORTED"}
ERRORS_BY_NAME = {v: k for k, v in ERRORS.items()}
PREC_FP64, PREC_FP32_MIXED = 0, 1
U, U0, RES = 0, 1, 2
PRES, TEMP, NE, RHOI, AMB, QTOT, TAU, DIVB, PHI, SFLX = range(10)
# diagnostics arrays tvar1..tvar8, Qres, Qvis
TVAR1, QRES, QVIS = 10, 18, 19
DIAG_NAMES = ("tvar1", "tvar2", "tvar3", "tvar4", "tvar5", "tvar6", "tvar7", "tvar8", "Qres", "Qvis")
LAYOUT_AOS, LAYOUT_SOA = 0, 1

# every symbol include/muram_step.h declares
EXPORTS = [
    "mrm_create", "mrm_destroy", "mrm_last_error", "mrm_bufsize", "mrm_synchronize",
    "mrm_eos_load", "mrm_set_state", "mrm_get_state", "mrm_set_field", "mrm_get_field",
    "mrm_get_limits", "mrm_set_limits", "mrm_get_bc", "mrm_set_bc", "mrm_diag_zero",
    "mrm_cons_to_prim", "mrm_adjust_valf_max", "mrm_mhd_residual", "mrm_sync_timestep",
    "mrm_source_integrate_tcheck", "mrm_save_cons", "mrm_tcheck", "mrm_tvdlimit", "mrm_clean_divB",
    "mrm_exchange_grid", "mrm_set_boundary_conditions", "mrm_step",
    "mrm_io_read_prim", "mrm_io_write_prim", "mrm_io_write_header", "mrm_io_read_header",
    "mrm_io_write_backup", "mrm_io_read_backup", "mrm_io_read_var", "mrm_io_write_var",
    "mrm_adjoint_forward", "mrm_adjoint_backward",
    "mrm_timers", "mrm_get_timers", "mrm_timer_name",
    "mrm_branch_count", "mrm_get_branch_counts", "mrm_mhd_residual_noalloc",
    "mrm_set_boundary_conditions_pt", "mrm_potential_kernels", "mrm_potential_kernels_load",
    "mrm_get_potential_bext", "mrm_io_read_psf_kernels",
]

# potential-field kernels in the order mrm_potential_kernels takes them
# (potential_sd_fftw3.C:77-86)
PSF_KERNELS = ("z-1", "z-2", "x-0", "x-1", "x-2", "y-0", "y-1", "y-2")
NTIMER = 24

# limiter branch accounting: names of the count slots (MRM_BR_*) and of the
# checksum slots (MRM_BK_*), in slot order
BRANCH_COUNTS = ("cz", "cme", "slope_pos_l", "slope_pos_r", "slope_pos_m", "slope_neg_l", "slope_neg_r",
                 "slope_neg_m", "slope_zero", "qrho", "vfast", "mm_pos_lim", "mm_pos_sl", "mm_neg_lim",
                 "mm_neg_sl", "mm_zero", "clip", "cmax_limit", "cmax_left", "cmax_right", "rho_floor",
                 "e_floor")
BRANCH_SUMS = ("cz", "cme", "slope", "qrho", "vfast", "minmod", "clip", "cmax", "rho_floor", "e_floor")
BR_NSLOT = len(BRANCH_COUNTS) + 3 * len(BRANCH_SUMS)
# near-tie margin of the parity checks: choices within 1e-10 of their
# operands' scale of the decision boundary
BRANCH_TIE = 1e-10


def branch_dict(slots):
    """{'counts': {name: n}, 'checksums': {name: hex}, 'near_ties': {name: n},
    'exact_ties': {name: n}} of MRM_BR_NSLOT slots"""
    n, k = len(BRANCH_COUNTS), len(BRANCH_SUMS)
    return {"counts": {c: int(slots[i]) for i, c in enumerate(BRANCH_COUNTS)},
            "checksums": {c: f"{int(slots[n + i]):016x}" for i, c in enumerate(BRANCH_SUMS)},
            "near_ties": {c: int(slots[n + k + i]) for i, c in enumerate(BRANCH_SUMS)},
            "exact_ties": {c: int(slots[n + 2 * k + i]) for i, c in enumerate(BRANCH_SUMS)}}


def branch_mismatch(a, b):
    """differences between two slot arrays: of the decisive choices (count
    slots that differ, summed absolute count difference, checksums that
    differ) and of the ties (summed absolute differences of the near-tie and
    exact-tie counts, and b's totals of each)"""
    n, k = len(BRANCH_COUNTS), len(BRANCH_SUMS)
    diff = [abs(int(a[i]) - int(b[i])) for i in range(n)]
    near = range(n + k, n + 2 * k)
    exact = range(n + 2 * k, n + 3 * k)
    return {"slots_differing": sum(1 for d in diff if d), "count_diff": sum(diff),
            "checksums_differing": sum(1 for i in range(n, n + k) if int(a[i]) != int(b[i])),
            "near_tie_diff": sum(abs(int(a[i]) - int(b[i])) for i in near),
            "near_ties": sum(int(b[i]) for i in near),
            "exact_tie_diff": sum(abs(int(a[i]) - int(b[i])) for i in exact),
            "exact_ties": sum(int(b[i]) for i in exact),
            "decisive": sum(int(b[i]) for i in range(n))}


class GridDesc(C.Structure):
    _fields_ = [("gsize", C.c_int * 3), ("ghosts", C.c_int * 3), ("periods", C.c_int * 3),
                ("gxmin", C.c_double * 3), ("gxmax", C.c_double * 3)]


class PhysDesc(C.Structure):
    _fields_ = [("params", C.c_double * 9), ("bnd", C.c_double * 5), ("tvd_h", C.c_double * 4),
                ("tvd_cs", C.c_double * 4), ("tvd", C.c_double * 10),
                ("tvd_h_bnd", C.c_double * 2), ("tvd_visc_bnd", C.c_double * 2),
                ("tvd_eta_bnd", C.c_double * 2), ("tvd_coeff", C.c_double * 4),
                ("tchk", C.c_double * 4), ("dmp", C.c_double * 4), ("rt", C.c_double * 10),
                ("divB", C.c_double * 3), ("rt_ext_cor", C.c_int), ("pad_", C.c_int)]


class RunDesc(C.Structure):
    _fields_ = [("dt", C.c_double), ("time", C.c_double), ("CFL", C.c_double),
                ("CFL_tvd", C.c_double), ("outcad", C.c_double), ("autostep", C.c_int),
                ("need_diagnostics", C.c_int), ("globiter", C.c_int), ("iteration", C.c_int)]


class MrmError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Load the gfx950 library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MrmError(f"HIP library {path} is missing: run __graft_entry__.build() first "
                       "(there is no CPU fallback)")
    lib = C.CDLL(path)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    sig = {
        "mrm_create": ([C.POINTER(GridDesc), C.POINTER(PhysDesc), C.c_int, C.POINTER(P)], C.c_int),
        "mrm_destroy": ([P], C.c_int),
        "mrm_last_error": ([P], C.c_char_p),
        "mrm_bufsize": ([P, C.POINTER(C.c_long), C.POINTER(C.c_int)], C.c_int),
        "mrm_synchronize": ([P], C.c_int),
        "mrm_eos_load": ([P, C.c_void_p, C.c_size_t], C.c_int),
        "mrm_set_state": ([P, C.c_int, dp, C.c_int], C.c_int),
        "mrm_get_state": ([P, C.c_int, dp, C.c_int], C.c_int),
        "mrm_set_field": ([P, C.c_int, dp], C.c_int),
        "mrm_get_field": ([P, C.c_int, dp], C.c_int),
        "mrm_get_limits": ([P, dp], C.c_int),
        "mrm_set_limits": ([P, dp], C.c_int),
        "mrm_get_bc": ([P, dp, dp], C.c_int),
        "mrm_set_bc": ([P, C.c_double, C.c_double], C.c_int),
        "mrm_cons_to_prim": ([P], C.c_int),
        "mrm_adjust_valf_max": ([P, C.POINTER(RunDesc), dp], C.c_int),
        "mrm_mhd_residual": ([P, C.POINTER(RunDesc), dp], C.c_int),
        "mrm_sync_timestep": ([P, C.POINTER(RunDesc), C.c_double, C.c_double, dp], C.c_int),
        "mrm_source_integrate_tcheck": ([P, C.POINTER(RunDesc), C.c_int], C.c_int),
        "mrm_save_cons": ([P], C.c_int),
        "mrm_diag_zero": ([P], C.c_int),
        "mrm_tcheck": ([P, C.POINTER(RunDesc)], C.c_int),
        "mrm_tvdlimit": ([P, C.POINTER(RunDesc), C.c_double], C.c_int),
        "mrm_clean_divB": ([P, C.POINTER(RunDesc), dp], C.c_int),
        "mrm_exchange_grid": ([P], C.c_int),
        "mrm_set_boundary_conditions": ([P, C.POINTER(RunDesc), C.c_int], C.c_int),
        "mrm_step": ([P, C.POINTER(RunDesc), C.c_double, C.c_int], C.c_int),
        "mrm_io_read_prim": ([C.c_char_p, C.c_char_p, C.c_int, C.POINTER(GridDesc), dp], C.c_int),
        "mrm_io_write_prim": ([C.c_char_p, C.c_char_p, C.c_int, C.POINTER(GridDesc), dp], C.c_int),
        "mrm_io_write_header": ([C.c_char_p, C.c_int, C.POINTER(GridDesc), C.c_double, C.c_double, C.c_double],
                                C.c_int),
        "mrm_io_read_header": ([C.c_char_p, C.c_int, C.POINTER(C.c_int), dp, dp, dp, dp], C.c_int),
        "mrm_io_write_backup": ([C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_double], C.c_int),
        "mrm_io_read_backup": ([C.c_char_p, C.POINTER(C.c_int), dp, dp, dp], C.c_int),
        "mrm_io_read_var": ([C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.POINTER(GridDesc), dp], C.c_int),
        "mrm_io_write_var": ([C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.POINTER(GridDesc), dp], C.c_int),
        "mrm_adjoint_forward": ([P, C.POINTER(RunDesc), C.c_double, C.c_int, C.c_int], C.c_int),
        "mrm_adjoint_backward": ([P, dp, dp, C.c_int], C.c_int),
        "mrm_timers": ([P, C.c_int], C.c_int),
        "mrm_get_timers": ([P, dp, C.POINTER(C.c_long)], C.c_int),
        "mrm_timer_name": ([C.c_int], C.c_char_p),
        "mrm_branch_count": ([P, C.c_int, C.c_double], C.c_int),
        "mrm_mhd_residual_noalloc": ([C.POINTER(GridDesc), C.POINTER(PhysDesc), C.POINTER(RunDesc), C.c_double,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
        "mrm_get_branch_counts": ([P, C.POINTER(C.c_ulonglong)], C.c_int),
        "mrm_set_boundary_conditions_pt": ([P, C.POINTER(RunDesc), C.c_int, C.c_int], C.c_int),
        "mrm_potential_kernels": ([P, C.POINTER(C.c_float)], C.c_int),
        "mrm_potential_kernels_load": ([P, C.c_char_p], C.c_int),
        "mrm_get_potential_bext": ([P, dp], C.c_int),
        "mrm_io_read_psf_kernels": ([C.c_char_p, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib
