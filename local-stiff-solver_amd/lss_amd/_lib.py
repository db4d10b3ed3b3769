"""ctypes binding of liblss.so (C ABI declared in include/lss.h).

The library is loaded from this package's ``lib/`` directory.  There is no CPU
fallback: a missing library raises, so every solve runs the HIP kernels.
"""
from __future__ import annotations

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib")
LIB_PATH = os.path.join(LIB_DIR, "liblss.so")

STATUS = {
    0: "LSS_OK",
    1: "LSS_ERR_CONFIG",
    2: "LSS_ERR_SIZE",
    3: "LSS_ERR_NOT_FACTORIZED",
    4: "LSS_ERR_HIP",
    5: "LSS_ERR_NOT_SPD",
    6: "LSS_ERR_OOM",
    7: "LSS_ERR_UNSUPPORTED",
    8: "LSS_ERR_ARG",
}

# every entry point declared in include/lss.h
EXPORTS = [
    "lss_create", "lss_set_parameters", "lss_set_problematic_dofs", "lss_analyze_pattern",
    "lss_factorize_csc", "lss_factorize_csc_device", "lss_solve", "lss_solve_device",
    "lss_solve_multi_device", "lss_get_info", "lss_name", "lss_get_stiff_dofs", "lss_last_error",
    "lss_destroy", "lss_adjoint_pattern_grad", "lss_adjoint_chain", "lss_debug_row_norms", "lss_debug_gmm",
    "lss_debug_partition", "lss_debug_spmv", "lss_debug_base_apply", "lss_debug_precond_apply",
    "lss_residual_norm", "lss_residual_norm_noalloc", "lss_debug_amg", "lss_debug_nested_dissection",
]

_lib = None


class LssError(RuntimeError):
    """Raised for a non-zero lss_status; .code holds the status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32, i64 = ctypes.c_int32, ctypes.c_int64
    L.lss_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(P)]
    L.lss_set_parameters.argtypes = [P, ctypes.c_char_p]
    L.lss_set_problematic_dofs.argtypes = [P, P, i64]
    L.lss_analyze_pattern.argtypes = [P, i32, i64, P, P]
    L.lss_factorize_csc.argtypes = [P, i32, i64, P, P, P]
    L.lss_factorize_csc_device.argtypes = [P, i32, i64, P, P, P, P]
    L.lss_solve.argtypes = [P, P, P]
    L.lss_solve_device.argtypes = [P, P, P, P]
    L.lss_solve_multi_device.argtypes = [P, i32, P, P, P]
    L.lss_get_info.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.lss_name.restype = ctypes.c_char_p
    L.lss_name.argtypes = [P]
    L.lss_get_stiff_dofs.argtypes = [P, P, i64, ctypes.POINTER(i64), P, i64, ctypes.POINTER(i64)]
    L.lss_last_error.restype = ctypes.c_char_p
    L.lss_last_error.argtypes = [P]
    L.lss_destroy.restype = None
    L.lss_destroy.argtypes = [P]
    L.lss_adjoint_pattern_grad.argtypes = [i32, i64, P, P, P, P, P, P]
    L.lss_adjoint_chain.argtypes = [i64, P, P, P, P, P, P]
    L.lss_residual_norm.argtypes = [i32, i64, P, P, P, P, P, P, P]
    L.lss_residual_norm_noalloc.argtypes = [i32, i64, P, P, P, P, P, P, P]
    L.lss_debug_row_norms.argtypes = [P, P, i64]
    L.lss_debug_gmm.argtypes = [P, P, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i64)]
    L.lss_debug_partition.argtypes = [P, P, i64, P, i64, ctypes.POINTER(i32)]
    L.lss_debug_amg.argtypes = [P, i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i64),
                                ctypes.POINTER(i64), P, P, P, P, P, P, P]
    L.lss_debug_spmv.argtypes = [P, P, P]
    L.lss_debug_base_apply.argtypes = [P, P, P]
    L.lss_debug_precond_apply.argtypes = [P, P, P]
    L.lss_debug_nested_dissection.argtypes = [i32, P, P, i32, P, ctypes.POINTER(i32), ctypes.POINTER(i64), P, P, P,
                                              P, P]
    for name in EXPORTS:
        if name not in ("lss_name", "lss_last_error", "lss_destroy"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L
