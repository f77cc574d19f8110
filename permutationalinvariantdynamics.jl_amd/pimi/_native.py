"""ctypes binding of libpimi (the C ABI declared in include/pimi.h).

The library is required: importing a device entry point without it raises,
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libpimi.so")

PIMI_OK, PIMI_EINVAL, PIMI_EDIM, PIMI_ENOMEM, PIMI_EDEVICE, PIMI_EBUSY, PIMI_ENOTCONVERGED, PIMI_ERANGE = range(8)
PIMI_C64, PIMI_C32 = 0, 1
PRECOND_NONE, PRECOND_SECTOR, PRECOND_SECTOR_INNER = 0, 1, 2
CRITERION_REFERENCE, CRITERION_BACKWARD, CRITERION_WEIGHTED = 0, 1, 2
WEIGHT_NONE, WEIGHT_SECTOR = 0, 1
TERM_LOCAL_H, TERM_COLLECTIVE_H, TERM_LOCAL_JUMP, TERM_COLLECTIVE_JUMP, TERM_DIRECT_PRODUCT_H = range(5)

# Every symbol the header declares.
EXPORTS = [
    "pimi_last_error", "pimi_version", "pimi_context_create", "pimi_context_destroy", "pimi_compile",
    "pimi_plan_create_fused", "pimi_plan_info_get", "pimi_plan_dims", "pimi_trace_functional",
    "pimi_collective_operator", "pimi_iid_state", "pimi_plan_banded_export", "pimi_plan_banded_programs",
    "pimi_plan_export_sizes", "pimi_plan_export", "pimi_plan_destroy",
    "pimi_workspace_create", "pimi_workspace_destroy", "pimi_apply", "pimi_apply_adjoint", "pimi_steady_gmres", "pimi_expv_series",
    "pimi_workspace_last_apply_ms", "pimi_apply_bench", "pimi_copy_bench",
    "pimi_sector_preconditioner_factors", "pimi_precond_apply",
    "pimi_batch_create", "pimi_batch_destroy", "pimi_batch_job_bytes", "pimi_batch_submit", "pimi_batch_run",
    "pimi_batch_counts", "pimi_batch_graph_counts", "pimi_batch_next_finished", "pimi_batch_fetch", "pimi_batch_submit_expv",
    "pimi_batch_fetch_expv", "pimi_batch_apply", "pimi_batch_apply_bench",
    "pimi_batch_stats", "pimi_plan_prepare", "pimi_precond_bench",
    "pimi_trajectory_channels", "pimi_conditional_action", "pimi_conditional_action_noalloc",
    "pimi_conditional_rk4", "pimi_channel_intensities",
    "pimi_apply_jumps", "pimi_arnoldi", "pimi_arnoldi_vectors", "pimi_evolve_rk4",
    "pimi_workspace_autotune_apply", "pimi_batch_autotune_apply",
]


class PIMIError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"pimi status {status}: {message}")
        self.status = status


class NotConverged(PIMIError):
    pass


class Term(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("op", C.c_void_p), ("op2", C.c_void_p),
                ("rate", C.c_double), ("hbar", C.c_double)]


class ModelDesc(C.Structure):
    _fields_ = [("N", C.c_int32), ("d", C.c_int32), ("nterms", C.c_int32), ("reserved", C.c_int32),
                ("terms", C.POINTER(Term))]


class CSCSet(C.Structure):
    _fields_ = [("cp_base", C.c_void_p), ("colptr", C.c_void_p), ("rows", C.c_void_p), ("vals", C.c_void_p)]


class FusedDesc(C.Structure):
    _fields_ = [("ns", C.c_int32), ("has_h", C.c_int32), ("has_q", C.c_int32), ("ncg", C.c_int32),
                ("nob", C.c_int32), ("reserved", C.c_int32), ("dims", C.c_void_p), ("H", CSCSet), ("Q", CSCSet),
                ("K", CSCSet), ("cg_scale", C.c_void_p), ("ob_scale", C.c_void_p), ("ob_br_base", C.c_void_p),
                ("br_out", C.c_void_p), ("br_in", C.c_void_p), ("br_r", C.c_void_p),
                ("br_use_support", C.c_void_p), ("br_sup_base", C.c_void_p), ("sup_rows", C.c_void_p),
                ("sup_cols", C.c_void_p), ("sup_vals", C.c_void_p), ("br_dense_base", C.c_void_p),
                ("dense", C.c_void_p)]


class PlanInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("nsectors", C.c_int32), ("max_block", C.c_int32), ("nterms", C.c_int64),
                ("stencil_entries", C.c_int64), ("table_bytes_c64", C.c_int64), ("table_bytes_c32", C.c_int64),
                ("apply_form", C.c_int32), ("banded_groups", C.c_int32)]


class GmresOpts(C.Structure):
    _fields_ = [("krylovdim", C.c_int32), ("preconditioner", C.c_int32), ("maxiter", C.c_int64), ("atol", C.c_double),
                ("rtol", C.c_double), ("operator_scale", C.c_double), ("inner_rtol", C.c_double),
                ("max_refinements", C.c_int64), ("criterion", C.c_int32), ("inner_steps", C.c_int32),
                ("weighting", C.c_int32), ("reserved", C.c_int32)]


class ExpvOpts(C.Structure):
    _fields_ = [("krylovdim", C.c_int32), ("reserved", C.c_int32), ("atol", C.c_double), ("rtol", C.c_double),
                ("safety", C.c_double), ("max_steps", C.c_int64)]


class BatchOpts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("max_jobs", C.c_int32), ("memory_bytes", C.c_int64),
                ("flags", C.c_int32), ("reserved", C.c_int32)]


BATCH_TIMING = 1
OP_CLASSES = ("precond_gemm", "apply", "dots", "update", "cycle_end", "fp64", "other")


class SolveInfo(C.Structure):
    _fields_ = [("converged", C.c_int32), ("reserved", C.c_int32), ("iterations", C.c_int64),
                ("restarts", C.c_int64), ("applications", C.c_int64), ("accepted", C.c_int64),
                ("rejected", C.c_int64), ("factorizations", C.c_int64), ("refinements", C.c_int64),
                ("residual", C.c_double), ("normalized_residual", C.c_double), ("trace_error", C.c_double),
                ("operator_scale", C.c_double), ("linear_residual", C.c_double), ("estimated_error", C.c_double),
                ("seconds", C.c_double), ("relative_residual", C.c_double),
                ("weighted_relative_residual", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib = None


def lib():
    """Load libpimi; raises if the native library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libpimi is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.pimi_last_error.restype = C.c_char_p
        L.pimi_version.restype = C.c_char_p
        for name in EXPORTS:
            f = getattr(L, name)
            if name not in ("pimi_last_error", "pimi_version", "pimi_context_destroy", "pimi_plan_destroy",
                            "pimi_workspace_destroy", "pimi_batch_destroy"):
                f.restype = C.c_int32
        L.pimi_context_destroy.restype = None
        L.pimi_plan_destroy.restype = None
        L.pimi_workspace_destroy.restype = None
        L.pimi_batch_destroy.restype = None
        _lib = L
    return _lib


def check(status):
    if status != PIMI_OK:
        msg = lib().pimi_last_error().decode()
        if status == PIMI_ENOTCONVERGED:
            raise NotConverged(status, msg)
        raise PIMIError(status, msg)


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)
