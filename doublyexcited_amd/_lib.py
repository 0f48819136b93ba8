"""ctypes binding of ``lib/libdex.so`` (the C-ABI declared in ``include/dex.h``).

The native library is required: importing this module without it raises, so
no caller can silently run anything else in its place.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIBPATH = Path(os.environ.get("DEX_LIB", Path(__file__).resolve().parent / "lib" / "libdex.so"))

DEX_OK = 0
DEX_ESIZE = 1
DEX_ECHOLESKY = 2
DEX_EGEVP = 3
DEX_ENOTCONVERGED = 4
DEX_EHIP = 5
DEX_EOOM = 6
DEX_EINVAL = 7
DEX_EADMISSION = 8
DEX_EUNSUPPORTED = 9
DEX_EZEROBLOCK = 10
DEX_FP64 = 0
DEX_FP32 = 1

# kernel classes of dex_get_kernel_stats
KERNEL_CLASSES = ("pass1", "pass2", "symadd", "gram", "combine", "coldots", "generic")

# every symbol include/dex.h declares
EXPORTS = [
    "dex_params_default", "dex_ctx_create", "dex_ctx_destroy", "dex_ctx_info", "dex_ctx_stop",
    "dex_get_kematrix", "dex_get_tinv", "dex_matvec", "dex_matvec_device", "dex_matvec_sym_device",
    "dex_matvec_noalloc", "dex_matvec_graph_create", "dex_graph_launch", "dex_graph_destroy",
    "dex_dots",
    "dex_dots_device", "dex_solve_opts_default", "dex_gplmr_eig", "dex_gplmr_eig_device",
    "dex_fill_gaussian_device", "dex_sync", "dex_residuals_device", "dex_gplmr_footprint",
    "dex_solve_footprint", "dex_precblock_device", "dex_diagsubspace_device", "dex_heplus_lanczos_device",
    "dex_status_string", "dex_last_error", "dex_set_kernel_timing", "dex_get_kernel_timing",
    "dex_get_kernel_stats", "dex_get_kernel_totals", "dex_reset_kernel_stats",
    "dex_host_kematrix", "dex_host_getinverse", "dex_host_solve_gevp",
    "dex_write_eigvecs", "dex_read_eigvecs_info", "dex_read_eigvecs", "dex_host_sixd_fix",
]


class DexParams(C.Structure):
    _fields_ = [("n", C.c_int32), ("delta", C.c_double), ("theta", C.c_double),
                ("charge", C.c_int32), ("nsmall", C.c_int32), ("tfactor", C.c_int32),
                ("precision", C.c_int32), ("device", C.c_int32)]


class DexInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("nn", C.c_int32), ("m", C.c_int64), ("d", C.c_int64),
                ("delta", C.c_double), ("theta", C.c_double), ("precision", C.c_int32),
                ("fast_path", C.c_int32)]


class DexResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("matvecs", C.c_int64),
                ("max_residual", C.c_double), ("t_total", C.c_double), ("t_matvec", C.c_double),
                ("t_prec", C.c_double), ("t_gram", C.c_double), ("t_update", C.c_double),
                ("t_smalleig", C.c_double), ("fallbacks", C.c_int32),
                ("fp64_restarts_used", C.c_int32)]


class DexSolveOpts(C.Structure):
    _fields_ = [("strict_cholesky", C.c_int32), ("verbose", C.c_int32),
                ("refresh_every", C.c_int32), ("fp64_restarts", C.c_int32),
                ("fp32_tol", C.c_double), ("symmetry", C.c_int32), ("lock", C.c_int32),
                ("inner_block", C.c_int32), ("guard", C.c_int32), ("inner_fp32", C.c_int32)]


class DexError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"dex status {status}: {msg}")
        self.status = status


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIBPATH.exists():
        raise ImportError(f"native library {LIBPATH} is missing; run `python -m doublyexcited_amd.build`")
    lib = C.CDLL(str(LIBPATH))
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    sig = {
        "dex_params_default": (None, [C.POINTER(DexParams)]),
        "dex_ctx_create": (C.c_int, [C.POINTER(DexParams), C.POINTER(P)]),
        "dex_ctx_destroy": (C.c_int, [P]),
        "dex_ctx_info": (C.c_int, [P, C.POINTER(DexInfo)]),
        "dex_get_kematrix": (C.c_int, [P, dp]),
        "dex_get_tinv": (C.c_int, [P, dp]),
        "dex_matvec": (C.c_int, [P, dp, dp, C.c_int32, C.c_int64]),
        "dex_matvec_device": (C.c_int, [P, P, P, C.c_int32, C.c_int64, C.c_int32]),
        "dex_matvec_sym_device": (C.c_int, [P, P, P, C.c_int32, C.c_int64, C.c_int32, C.c_int32]),
        "dex_matvec_noalloc": (C.c_int, [P, P, P, C.c_int32, C.c_int64, C.c_int32, C.c_int32]),
        "dex_matvec_graph_create": (C.c_int, [P, P, P, C.c_int32, C.c_int64, C.c_int32, C.c_int32,
                                              C.POINTER(P)]),
        "dex_graph_launch": (C.c_int, [P]),
        "dex_graph_destroy": (C.c_int, [P]),
        "dex_dots": (C.c_int, [P, dp, dp, C.c_int64, C.c_int32, C.c_int32, dp]),
        "dex_dots_device": (C.c_int, [P, P, P, C.c_int64, C.c_int32, C.c_int32, dp]),
        "dex_solve_opts_default": (None, [C.POINTER(DexSolveOpts)]),
        "dex_gplmr_eig": (C.c_int, [P, dp, dp, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                    C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                    C.c_int32, C.POINTER(DexSolveOpts), C.POINTER(DexResult)]),
        "dex_gplmr_eig_device": (C.c_int, [P, P, dp, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                           C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                           C.c_int32, C.POINTER(DexSolveOpts), C.POINTER(DexResult)]),
        "dex_fill_gaussian_device": (C.c_int, [P, P, C.c_int32, C.c_uint64, C.c_uint32, C.c_int32]),
        "dex_sync": (C.c_int, [P]),
        "dex_residuals_device": (C.c_int, [P, P, dp, C.c_int32, dp]),
        "dex_gplmr_footprint": (C.c_int64, [P, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
        "dex_solve_footprint": (C.c_int64, [C.POINTER(DexParams), C.c_int32, C.c_int32, C.c_int32,
                                            C.c_void_p]),
        "dex_diagsubspace_device": (C.c_int, [P, dp, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                              C.c_int32, C.c_void_p, dp, dp]),
        "dex_precblock_device": (C.c_int, [P, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                           C.c_double, C.c_double, C.c_int32, C.c_int32]),
        "dex_heplus_lanczos_device": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                                C.c_uint64, dp, dp, dp, C.POINTER(C.c_int32)]),
        "dex_status_string": (C.c_char_p, [C.c_int]),
        "dex_last_error": (C.c_char_p, []),
        "dex_set_kernel_timing": (C.c_int, [P, C.c_int32]),
        "dex_ctx_stop": (C.c_int, [P, C.c_int32]),
        "dex_get_kernel_timing": (C.c_int, [P, dp, dp]),
        "dex_get_kernel_stats": (C.c_int, [P, C.c_int32, dp, C.POINTER(C.c_int64), dp, dp]),
        "dex_get_kernel_totals": (C.c_int, [P, C.c_int32, C.POINTER(C.c_int64), dp, dp]),
        "dex_reset_kernel_stats": (C.c_int, [P]),
        "dex_host_kematrix": (C.c_int, [C.c_int32, C.c_double, dp]),
        "dex_write_eigvecs": (C.c_int, [C.c_char_p, dp, C.c_int32, C.c_int32]),
        "dex_read_eigvecs_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "dex_read_eigvecs": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, dp]),
        "dex_host_sixd_fix": (C.c_int, [dp, C.c_int32, dp, C.c_int32]),
        "dex_host_getinverse": (C.c_int, [C.c_int32, C.c_double, C.c_int32, C.c_int32, dp]),
        "dex_host_solve_gevp": (C.c_int, [C.c_int32, dp, dp, C.c_int32, C.c_int32, dp, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, allow=(DEX_OK,)) -> int:
    if status not in allow:
        lib = load()
        raise DexError(status, f"{lib.dex_status_string(status).decode()}: {lib.dex_last_error().decode()}")
    return status
