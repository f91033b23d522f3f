"""ctypes binding of libdfp.so (the C ABI declared in include/dfp.h)."""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# DFP_LIB selects another build of the same library (the host-sanitizer build
# that the CPU tests load); by default the in-tree libdfp.so
LIB_PATH = os.environ.get("DFP_LIB") or os.path.normpath(os.path.join(_HERE, "..", "lib", "libdfp.so"))
HEADER_PATH = os.path.normpath(os.path.join(_HERE, "..", "..", "include", "dfp.h"))

DFP_FP64 = 0
DFP_FP32_REFINE = 1
DFP_FIELDS_ON_DEVICE = 1
DFP_SAVE_P = 2
DFP_ASYNC = 4
DFP_SOLVE_ONLY = 8
DFP_CHECK_FINITE = 16

_dp = ctypes.POINTER(ctypes.c_double)


class DfpGrid(ctypes.Structure):
    _fields_ = [
        ("nxp", ctypes.c_int32), ("nym", ctypes.c_int32), ("nzp", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dz", ctypes.c_double),
        ("x", _dp), ("y", _dp), ("z", _dp), ("yg", _dp), ("zg", _dp),
        ("x_bc_type", ctypes.c_int32), ("y_bc_type", ctypes.c_int32), ("z_bc_type", ctypes.c_int32),
    ]


class DfpTuning(ctypes.Structure):
    _fields_ = [("z_split", ctypes.c_int32), ("x_columns", ctypes.c_int32), ("z_rows", ctypes.c_int32),
                ("z_fused", ctypes.c_int32)]


class DfpError(RuntimeError):
    """A nonzero return code of the C ABI (the analogue of the reference's gpu_abort)."""

    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        msg = f"{where}: {strerror(code)} ({code})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


_lib = None


def lib() -> ctypes.CDLL:
    """Load libdfp.so; raises if it has not been built (there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libdfp.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    sig = {
        "dfp_plan_create": ([ctypes.POINTER(DfpGrid), i32, i32, ctypes.POINTER(vp)], i32),
        "dfp_plan_destroy": ([vp], i32),
        "dfp_plan_create_tuned": ([ctypes.POINTER(DfpGrid), i32, i32, ctypes.POINTER(DfpTuning), ctypes.POINTER(vp)],
                                  i32),
        "dfp_plan_tuning": ([vp, ctypes.POINTER(DfpTuning)], i32),
        "dfp_batch_create": ([i32, i32, ctypes.POINTER(vp)], i32),
        "dfp_batch_destroy": ([vp], i32),
        "dfp_batch_set_stream": ([vp, vp], i32),
        "dfp_batch_set_occupancy": ([vp, i32, i32, i32], i32),
        "dfp_batch_use_graphs": ([vp, i32], i32),
        "dfp_batch_set_parallel": ([vp, i32], i32),
        "dfp_batch_project": ([vp, i32, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp),
                               ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(d), ctypes.POINTER(d),
                               ctypes.POINTER(i32), i32], i32),
        "dfp_batch_last_error": ([vp], ctypes.c_char_p),
        "dfp_project": ([vp, vp, vp, vp, vp, vp, d, d, i32], i32),
        "dfp_divergence": ([vp, vp, vp, vp, vp, i32], i32),
        "dfp_solve": ([vp, vp, vp, d, d, i32], i32),
        "dfp_correct": ([vp, vp, vp, vp, vp, i32], i32),
        "dfp_set_outflow": ([vp, d, d, vp, vp], i32),
        "dfp_rk_update": ([vp, i32, vp, vp, vp, vp, vp, vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp),
                           ctypes.POINTER(d), d, i32], i32),
        "dfp_apply_bc": ([vp, vp, vp, vp, vp, vp, vp, i32], i32),
        "dfp_substep": ([vp, i32, vp, vp, vp, vp, vp, vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp),
                         ctypes.POINTER(d), vp, vp, vp, vp, vp, d, i32], i32),
        "dfp_outflow_backflow": ([vp, vp, vp, vp, vp, d, d, i32], i32),
        "dfp_max_divergence": ([vp, vp, vp, vp, ctypes.POINTER(d), i32], i32),
        "dfp_max_divergence_noalloc": ([vp, vp, vp, vp, ctypes.POINTER(d), i32], i32),
        "dfp_set_stream": ([vp, vp], i32),
        "dfp_set_refine_sweeps": ([vp, i32], i32),
        "dfp_profile_enable": ([vp, i32], i32),
        "dfp_profile_read": ([vp, i32, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(i64), ctypes.POINTER(d)], i32),
        "dfp_profile_reset": ([vp], i32),
        "dfp_field_sizes": ([vp, ctypes.POINTER(i64)], i32),
        "dfp_synth_fields": ([vp, vp, vp, vp, ctypes.c_uint64, d], i32),
        "dfp_synth_fields_ex": ([vp, vp, vp, vp, ctypes.c_uint64, d, i32], i32),
        "dfp_strerror": ([i32], ctypes.c_char_p),
        "dfp_last_error": ([vp], ctypes.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def strerror(code: int) -> str:
    try:
        return lib().dfp_strerror(int(code)).decode()
    except Exception:
        return "error"


def check(rc: int, where: str, plan=None) -> None:
    if rc != 0:
        detail = ""
        if plan:
            try:
                detail = lib().dfp_last_error(plan).decode()
            except Exception:
                detail = ""
        raise DfpError(rc, where, detail)
