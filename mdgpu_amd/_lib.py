"""ctypes binding of the C ABI declared in include/mdgpu_c.h.

The shared library is built in-tree (``mdgpu_amd/libmdgpu_amd.so``) by
``mdgpu_amd.build``.  Nothing here falls back to another implementation: if
the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MDGPU_AMD_LIB", os.path.join(_HERE, "libmdgpu_amd.so"))

MDGPU_OK = 0
MDGPU_ERR_ARG = 1
MDGPU_ERR_NOT_SPD = 2
MDGPU_ERR_DEVICE = 3
MDGPU_ERR_NOMEM = 4
MDGPU_ERR_NOT_CONVERGED = 5
SOLVE_TIMED_OUT = 1        # MDGPU_SOLVE_TIMED_OUT

FP64, FP32, BF16, BF16X3 = 0, 1, 2, 3
PART_PLAIN, PART_SUPERNODAL, PART_RELAXED = 0, 1, 2

SYM_N, SYM_NNODES, SYM_LNZ, SYM_NNZ_A, SYM_QLEN, SYM_MAXFRONT, SYM_NLEVELS, SYM_NODES_FLAT, SYM_NSEG = range(9)
SYM_NSIZES = 9

(ARR_AMD_PERM, ARR_PARENT, ARR_POST, ARR_COUNT, ARR_RELAXED_PERM, ARR_PERM, ARR_COUNT_FINAL,
 ARR_LP, ARR_ALOW_P, ARR_ALOW_I, ARR_NODES, ARR_FRONT_PARENT, ARR_Q_OFF, ARR_Q, ARR_LEVEL,
 ARR_SEG_COL, ARR_SEG_NNZ) = range(17)


class MdgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[status {code}] {msg}")
        self.code = code


class AnalyzeOpts(C.Structure):
    _fields_ = [("use_amd", C.c_int), ("partition", C.c_int), ("max_zeros", C.c_int)]


_lib = None

_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes); the C ABI surface that Python uses
_SIGNATURES = {
    "mdgpu_last_error": (C.c_char_p, []),
    "mdgpu_analyze": (C.c_int, [C.c_int, _I32P, _I32P, C.POINTER(AnalyzeOpts), C.POINTER(_P)]),
    "mdgpu_analyze_batch": (C.c_int, [C.c_int, _I32P, C.POINTER(_I32P), C.POINTER(_I32P), C.POINTER(AnalyzeOpts),
                                      C.c_int, C.POINTER(_P)]),
    "mdgpu_symbolic_destroy": (C.c_int, [_P]),
    "mdgpu_symbolic_save": (C.c_int, [_P, C.c_char_p]),
    "mdgpu_symbolic_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "mdgpu_symbolic_sizes": (C.c_int, [_P, _I64P]),
    "mdgpu_symbolic_flops": (C.c_int, [_P, _DP]),
    "mdgpu_symbolic_array": (C.c_int, [_P, C.c_int, _I32P, C.c_int64]),
}


def register(name, restype, argtypes):
    _SIGNATURES[name] = (restype, argtypes)


def lib():
    """Load the in-tree shared library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        # torch bundles a HIP runtime with the same SONAME as ROCm's; loading
        # it first makes this library bind to that one runtime, so device
        # pointers and stream handles from torch are valid here.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `python -m mdgpu_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != MDGPU_OK:
        msg = lib().mdgpu_last_error()
        raise MdgpuError(status, msg.decode() if msg else "")
    return status

(NUM_DEVICE_BYTES, NUM_LSTORE, NUM_ARENA, NUM_STEPS, NUM_LEVELS, NUM_PANEL_FRONTS, NUM_DIAG_JOBS, NUM_PRECISION,
 NUM_FALLBACKS) = range(9)
NUM_NINFO = 9
KIND_NAMES = ["assembly", "extend_add", "panel", "potrf_trsm", "inverse", "gemm"]

_SIGNATURES.update({
    "mdgpu_numeric_create": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "mdgpu_numeric_destroy": (C.c_int, [_P]),
    "mdgpu_numeric_set_graph": (C.c_int, [_P, C.c_int]),
    "mdgpu_numeric_info": (C.c_int, [_P, _I64P]),
    "mdgpu_plan_check": (C.c_int, [_P, C.c_int, _I64P]),
    "mdgpu_factorize": (C.c_int, [_P, _P, _P]),
    "mdgpu_factorize_profile": (C.c_int, [_P, _P, _P, _DP]),
    "mdgpu_factorize_steps": (C.c_int, [_P, _P, _P, _DP, C.c_int64]),
    "mdgpu_numeric_autotune": (C.c_int, [_P, _P, _P, _I64P]),
    "mdgpu_solve": (C.c_int, [_P, _P, _P, _P]),
    "mdgpu_solve_nrhs": (C.c_int, [_P, _P, _P, C.c_int, C.c_int64, _P]),
    "mdgpu_refine": (C.c_int, [_P, _P, _P, C.c_double, C.c_int, _DP, C.POINTER(C.c_int), _P]),
    "mdgpu_refine_segment_eta": (C.c_int, [_P, _DP, C.c_int]),
    "mdgpu_logdet": (C.c_int, [_P, _DP, _P]),
    "mdgpu_export_L": (C.c_int, [_P, _I32P, _I32P, _DP]),
    "mdgpu_export_L_segment": (C.c_int, [_P, C.c_int, _I32P, _I32P, _DP]),
    "mdgpu_solve_status": (C.c_int, [_P, C.POINTER(C.c_int), _P]),
    "mdgpu_row_union": (C.c_int, [_I32P, C.c_int, _I32P, C.c_int, _I32P, C.POINTER(C.c_int)]),
    "mdgpu_row_map": (C.c_int, [_I32P, C.c_int, _I32P, C.c_int, _I32P]),
    "mdgpu_backward_error": (C.c_int, [C.c_int, _P, _P, _P, _P, _P, _DP, _P]),
    "mdgpu_backward_error_noalloc": (C.c_int, [C.c_int, _P, _P, _P, _P, _P, _DP, _P]),
    "mdgpu_analyze_budgets": (C.c_int, [C.c_int, _I32P, _I32P, C.c_int, C.c_int, _I32P, C.POINTER(_P)]),
    "mdgpu_tune_partition": (C.c_int, [C.POINTER(_P), C.c_int, _P, C.c_int, C.c_int, C.c_int, _P, _DP,
                                       C.POINTER(C.c_int)]),
    "mdgpu_selinv": (C.c_int, [_P, _P]),
    "mdgpu_selinv_export": (C.c_int, [_P, _DP]),
    "mdgpu_logdet_grad": (C.c_int, [_P, _P, _P]),
    "mdgpu_solve_adjoint": (C.c_int, [_P, _P, _P, _P, _P, _P]),
})
