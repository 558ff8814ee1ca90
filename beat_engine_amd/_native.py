"""ctypes binding of libbeat_bem_hip.so (declarations in include/beat_bem.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbeat_bem_hip.so")

BEAT_OK = 0
BEAT_ERR_ARG = -1
BEAT_ERR_HIP = -2
BEAT_ERR_ALLOC = -3
BEAT_ERR_SINGULAR = -4
BEAT_ERR_CANCELLED = -5
PREC_FP64 = 0
PREC_FP32_REFINED = 1
PREC_FP32 = 2

REFINE_STATUS = {0: "none", 1: "converged", 2: "stagnated", 3: "step_limit", 4: "rounding_limited", 5: "nonfinite"}

EXPORTS = (
    "beat_ctx_create", "beat_ctx_destroy", "beat_last_error", "beat_mesh_upload", "beat_mesh_destroy",
    "beat_mesh_counts", "beat_bm_assemble", "beat_bm_solve", "beat_bm_solve_device", "beat_field_eval",
    "beat_dense_lu_solve", "beat_dense_lu_factor", "beat_refine_lu", "beat_mesh_upload_symmetric", "beat_mesh_symmetry",
    "beat_field_gradient_eval", "beat_operators_assemble", "beat_bm_sweep_device",
    "beat_bm_sweep", "beat_ctx_set_cancel_path", "beat_ctx_reclaim", "beat_ctx_memory_info", "beat_capabilities",
    "beat_adjoint_spl_gradient", "beat_bm_batch", "beat_autotune", "beat_ctx_debug_abort_waits",
    "beat_ctx_debug_force_exchange", "beat_ctx_regular_symmetric", "beat_field_eval_device_noalloc",
    "beat_ctx_set_tuning",
)

SYMMETRY_CODES = {"off": 0, "x": 1, "xy": 2, "ground": 3}


class RefineReport(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int),
        ("status", ctypes.c_int),
        ("working_status", ctypes.c_int),
        ("history_len", ctypes.c_int),
        ("rtol", ctypes.c_double),
        ("working_residuals", ctypes.POINTER(ctypes.c_double)),
        ("returned_residuals", ctypes.POINTER(ctypes.c_double)),
        ("history", ctypes.POINTER(ctypes.c_double)),
    ]


class SolveReport(ctypes.Structure):
    _fields_ = [
        ("assembly_s", ctypes.c_double),
        ("factor_s", ctypes.c_double),
        ("solve_s", ctypes.c_double),
        ("regular_kernel_s", ctypes.c_double),
        ("factorization_count", ctypes.c_int),
        ("lu_panels", ctypes.c_int),
        ("regular_pairs", ctypes.c_int64),
        ("singular_pairs", ctypes.c_int64),
        ("refine", RefineReport),
    ]


class SolveOptions(ctypes.Structure):
    _fields_ = [("max_steps", ctypes.c_int), ("rtol", ctypes.c_double)]


_lib = None


class BeatError(RuntimeError):
    pass


def capabilities(lib=None, ctx=None) -> dict:
    """The library's capability object (beat_capabilities)."""
    import json

    L = lib or load()
    n = L.beat_capabilities(ctx, None, 0)
    buf = ctypes.create_string_buffer(int(n))
    L.beat_capabilities(ctx, buf, n)
    return json.loads(buf.value.decode())


def load(path: str = LIB_PATH):
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"HIP library not built: {path} is missing (run __graft_entry__.build() or make -C beat_engine_amd/csrc)")
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    I = ctypes.c_int
    I64 = ctypes.c_int64
    D = ctypes.c_double
    L.beat_ctx_create.argtypes = [I, ctypes.POINTER(P)]
    L.beat_ctx_destroy.argtypes = [P]
    L.beat_ctx_destroy.restype = None
    L.beat_last_error.argtypes = [P]
    L.beat_last_error.restype = ctypes.c_char_p
    L.beat_mesh_upload.argtypes = [P, P, I64, P, I64, I, I, ctypes.POINTER(P)]
    L.beat_mesh_destroy.argtypes = [P]
    L.beat_mesh_destroy.restype = None
    L.beat_mesh_counts.argtypes = [P, P]
    L.beat_mesh_upload_symmetric.argtypes = [P, P, I64, P, I64, I, I, I, ctypes.POINTER(P)]
    L.beat_mesh_symmetry.argtypes = [P, P]
    L.beat_bm_assemble.argtypes = [P, P, I, D, D, D, P, I64, P, P]
    L.beat_bm_solve.argtypes = [P, P, I, D, D, D, P, I64, P, ctypes.POINTER(SolveOptions), ctypes.POINTER(SolveReport)]
    L.beat_bm_solve_device.argtypes = [P, P, I, D, D, D, P, I64, P, ctypes.POINTER(SolveOptions),
                                       ctypes.POINTER(SolveReport)]
    L.beat_field_eval.argtypes = [P, P, I, D, P, P, I64, P, I64, P]
    L.beat_field_gradient_eval.argtypes = [P, P, I, D, P, P, I64, P, I64, P]
    L.beat_field_eval_device_noalloc.argtypes = [P, P, I, D, P, P, I64, P, I64, P]
    L.beat_operators_assemble.argtypes = [P, P, I, D, P, P, P, P]
    L.beat_bm_sweep_device.argtypes = [P, P, I, I64, P, P, P, P, I64, P, ctypes.POINTER(SolveOptions), P]
    L.beat_bm_sweep.argtypes = [P, P, I, I64, P, P, P, P, I64, P, ctypes.POINTER(SolveOptions), P]
    L.beat_ctx_set_cancel_path.argtypes = [P, ctypes.c_char_p]
    L.beat_ctx_reclaim.argtypes = [P]
    L.beat_ctx_debug_abort_waits.argtypes = [P, I]
    L.beat_ctx_debug_force_exchange.argtypes = [P, I]
    L.beat_ctx_regular_symmetric.argtypes = [P, I]
    L.beat_ctx_memory_info.argtypes = [P, P]
    L.beat_capabilities.argtypes = [P, ctypes.c_char_p, I64]
    L.beat_capabilities.restype = I64
    L.beat_autotune.argtypes = [P, P, I, D, I64, P, P]
    L.beat_ctx_set_tuning.argtypes = [P, I, I64, P]
    L.beat_bm_batch.argtypes = [P, I64, P, I, P, P, P, P, I64, P, ctypes.POINTER(SolveOptions), P]
    L.beat_adjoint_spl_gradient.argtypes = [P, P, I64, P, P, P, P, P, I64, D, P, P, P]
    L.beat_dense_lu_solve.argtypes = [P, I, P, I64, P, I64, P]
    L.beat_dense_lu_factor.argtypes = [P, I, P, I64, P, P, P]
    L.beat_refine_lu.argtypes = [P, P, I64, P, I64, P, P, I, D, P, ctypes.POINTER(RefineReport)]
    if path == LIB_PATH:
        _lib = L
    return L
