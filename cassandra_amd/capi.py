"""ctypes binding of libcassandra_amd.so (include/cassandra_amd.h).

The library is loaded from the package directory only; a missing library is an
error (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcassandra_amd.so")

CAD_OK = 0
ERRORS = {
    -1: "argument", -2: "size mismatch", -3: "pattern", -4: "device", -5: "out of memory", -6: "breakdown",
}

ERR_ARGUMENT, ERR_SIZE_MISMATCH, ERR_PATTERN, ERR_DEVICE, ERR_OUT_OF_MEMORY, ERR_BREAKDOWN = -1, -2, -3, -4, -5, -6
SOLVER_LSQR, SOLVER_PCG, SOLVER_CG, SOLVER_LSMR = 0, 1, 2, 3
PREC_FP64, PREC_FP32, PREC_MIXED = 0, 1, 2
LAYOUT_ROW_MAJOR, LAYOUT_SLOT_MAJOR = 0, 1


class CadError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cassandra_amd error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Schedule(C.Structure):
    """cad_schedule: the product schedule of a matrix."""
    _fields_ = [
        ("gather_f32_rpt", C.c_int), ("gather_f32_block", C.c_int), ("gather_f32_sub", C.c_int),
        ("gather_f64_rpt", C.c_int), ("gather_f64_block", C.c_int), ("gather_f64_sub", C.c_int),
        ("scatter_rpt", C.c_int), ("scatter_order", C.c_int),
        ("scatter_tiles_per_cu", C.c_double), ("scatter_partial_cap", C.c_double), ("scatter_balance", C.c_double),
    ]


class SolveOptions(C.Structure):
    _fields_ = [
        ("solver", C.c_int), ("precision", C.c_int), ("precondition", C.c_int), ("max_iterations", C.c_int),
        ("tol", C.c_double), ("inner_tol", C.c_double), ("max_refinements", C.c_int), ("check_every", C.c_int),
        ("use_graph", C.c_int), ("verbose", C.c_int),
    ]


class SolveReport(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("refinements", C.c_int), ("converged", C.c_int), ("explicit_checks", C.c_int),
        ("rel_normal_residual", C.c_double), ("residual_norm", C.c_double), ("seconds", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


vp = C.c_void_p
i64 = C.c_int64
u64 = C.c_uint64
PP = C.POINTER(C.c_void_p)

# name: (restype, argtypes)
SIGNATURES = {
    "cad_last_error": (C.c_char_p, []),
    "cad_version": (C.c_char_p, []),
    "cad_context_create": (C.c_int, [C.c_int, PP]),
    "cad_context_destroy": (C.c_int, [vp]),
    "cad_context_set_stream": (C.c_int, [vp, vp]),
    "cad_context_get_stream": (C.c_int, [vp, PP]),
    "cad_context_reset_stream": (C.c_int, [vp]),
    "cad_context_synchronize": (C.c_int, [vp]),
    "cad_malloc": (C.c_int, [vp, C.c_size_t, PP]),
    "cad_free": (C.c_int, [vp, vp]),
    "cad_memcpy_h2d": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "cad_memcpy_d2h": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "cad_memcpy_d2d": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "cad_patterns_create": (C.c_int, [vp, vp, C.c_int, PP]),
    "cad_patterns_destroy": (C.c_int, [vp]),
    "cad_patterns_info": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(i64), C.POINTER(C.c_int)]),
    "cad_patterns_pattern_info": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(u64), C.POINTER(i64),
                                            C.POINTER(C.c_int), C.POINTER(i64)]),
    "cad_patterns_variations": (C.c_int, [vp, vp]),
    "cad_patterns_indices_host": (C.c_int, [vp, vp, i64, vp]),
    "cad_index_extract": (C.c_int, [vp, vp, i64, vp, C.c_int]),
    "cad_matrix_from_positions": (C.c_int, [vp, vp, i64, PP]),
    "cad_matrix_from_positions_systems": (C.c_int, [vp, vp, vp, C.c_int, PP]),
    "cad_matrix_systems": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(i64), vp]),
    "cad_matrix_from_positions_mf": (C.c_int, [vp, vp, i64, PP]),
    "cad_matrix_autotune": (C.c_int, [vp, C.c_char_p, C.c_size_t]),
    "cad_matrix_get_schedule": (C.c_int, [vp, C.POINTER(Schedule)]),
    "cad_matrix_kernel_times": (C.c_int, [vp, C.c_int, vp, C.c_int]),
    "cad_matrix_set_schedule": (C.c_int, [vp, C.POINTER(Schedule)]),
    "cad_matrix_create": (C.c_int, [vp, vp, i64, C.c_int, i64, PP]),
    "cad_matrix_destroy": (C.c_int, [vp]),
    "cad_matrix_info": (C.c_int, [vp, C.POINTER(i64), C.POINTER(C.c_int), C.POINTER(i64), C.POINTER(C.c_int)]),
    "cad_matrix_download": (C.c_int, [vp, vp]),
    "cad_spmv_f32": (C.c_int, [vp, vp, vp]),
    "cad_spmv_f64": (C.c_int, [vp, vp, vp]),
    "cad_spmtv_f32": (C.c_int, [vp, vp, vp]),
    "cad_spmtv_f64": (C.c_int, [vp, vp, vp]),
    "cad_jacobi": (C.c_int, [vp, vp]),
    "cad_solve_options_default": (None, [C.POINTER(SolveOptions)]),
    "cad_solve": (C.c_int, [vp, vp, vp, C.POINTER(SolveOptions), C.POINTER(SolveReport)]),
    "cad_solve_systems": (C.c_int, [vp, vp, vp, C.POINTER(SolveOptions), C.POINTER(SolveReport), vp, vp]),
    "cad_rel_normal_residual": (C.c_int, [vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "cad_eval_forward_f32": (C.c_int, [vp, vp, i64, vp, vp]),
    "cad_eval_forward_f64": (C.c_int, [vp, vp, i64, vp, vp]),
    "cad_eval_backward_f32": (C.c_int, [vp, vp, i64, vp, vp]),
    "cad_eval_backward_f64": (C.c_int, [vp, vp, i64, vp, vp]),
    "cad_eval_backward_f32_noalloc": (C.c_int, [vp, vp, i64, vp, vp]),
    "cad_eval_train_step_f32": (C.c_int, [vp, vp, vp, vp, i64, C.c_float, C.c_float, C.c_float, C.c_float,
                                          C.c_float, i64, vp, vp, vp, vp]),
    "cad_eval_train_epoch_f32": (C.c_int, [vp, vp, vp, vp, i64, i64, C.c_float, C.c_float, C.c_float, C.c_float,
                                           C.c_float, i64, vp, vp, vp, vp]),
    "cad_patterns_expanded_size": (C.c_int, [vp, C.POINTER(i64)]),
    "cad_weights_expand": (C.c_int, [vp, vp, vp]),
    "cad_weights_compact": (C.c_int, [vp, vp, vp]),
    "cad_evaluator_create": (C.c_int, [vp, vp, PP]),
    "cad_evaluator_destroy": (C.c_int, [vp]),
    "cad_evaluator_set_weights": (C.c_int, [vp, vp]),
    "cad_evaluator_eval": (C.c_int, [vp, vp, i64, vp]),
    "cad_evaluator_eval_host": (C.c_int, [vp, vp, i64, vp]),
    "cad_random_play": (C.c_int, [vp, u64, i64, C.c_int, C.c_int, vp]),
    "cad_make_targets": (C.c_int, [vp, u64, u64, C.c_double, C.c_double, vp, vp]),
    "cad_flip": (u64, [C.c_int, u64]),
    "cad_possible_moves": (u64, [u64, u64]),
    "cad_flips": (u64, [u64, u64, C.c_int]),
    "cad_random_play_host": (C.c_int, [u64, i64, C.c_int, C.c_int, vp]),
}

_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m cassandra_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != CAD_OK:
        raise CadError(rc, lib().cad_last_error().decode())
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def default_options(**kw):
    o = SolveOptions()
    lib().cad_solve_options_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o
