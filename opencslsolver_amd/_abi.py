"""ctypes declarations of include/ocsl.h."""
from __future__ import annotations

import ctypes

c_handle = ctypes.c_void_p
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)

OCSL_OK = 0
OCSL_WARN_MAXITER = 1
OCSL_WARN_BREAKDOWN = 2
OCSL_WARN_NONFINITE = 4
OCSL_ERR_ARG = -1
OCSL_ERR_STRUCTURE = -2
OCSL_ERR_NOMEM = -3
OCSL_ERR_HIP = -4
OCSL_ERR_STATE = -5
OCSL_ERR_NODEVICE = -6

CG = 0
ILU_BICGSTAB = 1
MODE_REFERENCE = 0
MODE_TOLERANCE = 1
FP64 = 0
MIXED = 1
HOST = 0
DEVICE = 1
TRI_EXACT = 0
TRI_JACOBI = 1
BICGSTAB_ORDER_INPUT = 0
BICGSTAB_ORDER_COLOUR = 1


class Options(ctypes.Structure):
    _fields_ = [
        ("tol", ctypes.c_double),
        ("max_iter", ctypes.c_int),
        ("ilu_sweeps", ctypes.c_int),
        ("tri_mode", ctypes.c_int),
        ("tri_sweeps", ctypes.c_int),
        ("tri_weight", ctypes.c_double),
        ("replace_ratio", ctypes.c_double),
        ("check_every", ctypes.c_int),
        ("reorder", ctypes.c_int),
        ("use_graph", ctypes.c_int),
        ("block_size", ctypes.c_int),
        ("grid_blocks", ctypes.c_int),
        ("vec_blocks", ctypes.c_int),
        ("rows_per_lane", ctypes.c_int),
        ("cg_kernels", ctypes.c_int),
        ("x_period", ctypes.c_int),
        ("bicgstab_order", ctypes.c_int),
    ]


class Report(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("replacements", ctypes.c_int),
        ("history_len", ctypes.c_int),
        ("residual", ctypes.c_double),
        ("true_relres", ctypes.c_double),
        ("solve_ms", ctypes.c_double),
        ("bytes", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneEntry(ctypes.Structure):
    _fields_ = [("rows_per_lane", ctypes.c_int), ("block_size", ctypes.c_int), ("grid_blocks", ctypes.c_int),
                ("vec_blocks", ctypes.c_int), ("cg_kernels", ctypes.c_int), ("us_per_iteration", ctypes.c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    "ocsl_default_options": (None, [ctypes.POINTER(Options)]),
    "ocsl_device_count": (ctypes.c_int, []),
    "ocsl_create": (ctypes.c_int, [ctypes.POINTER(c_handle), ctypes.c_int64, _u64p, _u64p]),
    "ocsl_destroy": (ctypes.c_int, [c_handle]),
    "ocsl_get_size": (ctypes.c_int, [c_handle, _i64p, _i64p]),
    "ocsl_set_values": (ctypes.c_int, [c_handle, ctypes.c_void_p, ctypes.c_int]),
    "ocsl_set_rhs": (ctypes.c_int, [c_handle, ctypes.c_void_p, ctypes.c_int]),
    "ocsl_set_x": (ctypes.c_int, [c_handle, ctypes.c_void_p, ctypes.c_int]),
    "ocsl_get_x": (ctypes.c_int, [c_handle, ctypes.c_void_p, ctypes.c_int]),
    "ocsl_solve": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(Options), ctypes.POINTER(Report)]),
    "ocsl_solve_multi": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.POINTER(Options), ctypes.POINTER(Report)]),
    "ocsl_CG_solve": (ctypes.c_int, [c_handle, _f64p]),
    "ocsl_ILU_BICGSTAB_solve": (ctypes.c_int, [c_handle, _f64p]),
    "ocsl_ILU_BICGSTAB_solve_multi": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_int, _f64p]),
    "ocsl_get_history": (ctypes.c_int, [c_handle, _f64p, ctypes.c_int, _i32p]),
    "ocsl_ABSsum": (ctypes.c_int, [c_handle, ctypes.c_void_p, ctypes.c_int, _f64p]),
    "ocsl_sum": (ctypes.c_int, [c_handle, ctypes.c_void_p, ctypes.c_int, _f64p]),
    "ocsl_norm": (ctypes.c_int, [c_handle, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _f64p]),
    "ocsl_gemv": (ctypes.c_int, [c_handle, ctypes.c_double, ctypes.c_void_p, ctypes.c_double,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "ocsl_ILU_decompose": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]),
    "ocsl_set_reference_graphs": (ctypes.c_int, [c_handle, ctypes.c_int]),
    "ocsl_time_tri_solve": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "ocsl_set_tri_form": (ctypes.c_int, [c_handle, ctypes.c_int]),
    "ocsl_time_tri_solve_form": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.POINTER(ctypes.c_double)]),
    "ocsl_colour_order": (ctypes.c_int, [c_handle, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "ocsl_ILU_decompose_from": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_int]),
    "ocsl_ILU_forsub": (ctypes.c_int, [c_handle, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "ocsl_ILU_backsub": (ctypes.c_int, [c_handle, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "ocsl_vec_norm": (ctypes.c_int, [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _f64p]),
    "ocsl_vec_ABSsum_noalloc": (ctypes.c_int, [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "ocsl_vec_sum": (ctypes.c_int, [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, _f64p]),
    "ocsl_vec_ABSsum": (ctypes.c_int, [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, _f64p]),
    "ocsl_vec_product": (ctypes.c_int, [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_int]),
    "ocsl_vec_axpy": (ctypes.c_int, [ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_double,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "ocsl_time_cg_spmv": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         _f64p, _f64p]),
    "ocsl_time_cg_iteration": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_int, _f64p, _f64p]),
    "ocsl_get_stream": (ctypes.c_void_p, [c_handle]),
    "ocsl_tune_cg": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Options), _f64p,
                                    ctypes.POINTER(TuneEntry), ctypes.c_int, _i32p]),
    "ocsl_set_tuned_cg": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.POINTER(Options)]),
    "ocsl_adjoint": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.POINTER(Options), ctypes.POINTER(Report)]),
    "ocsl_sampled_outer": (ctypes.c_int, [c_handle, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int]),
    "ocsl_time_sampled_outer": (ctypes.c_int, [c_handle, ctypes.c_int, _f64p, _f64p]),
    "ocsl_batch_create": (ctypes.c_int, [ctypes.POINTER(c_handle), ctypes.c_int, _i64p,
                                         ctypes.POINTER(_u64p), ctypes.POINTER(_u64p)]),
    "ocsl_batch_destroy": (ctypes.c_int, [c_handle]),
    "ocsl_batch_size": (ctypes.c_int, [c_handle, _i32p, _i64p]),
    "ocsl_batch_set_values": (ctypes.c_int, [c_handle, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    "ocsl_batch_set_rhs": (ctypes.c_int, [c_handle, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    "ocsl_batch_solve": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.POINTER(Options),
                                        ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                        ctypes.POINTER(Report), _f64p]),
    "ocsl_batch_set_launch": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_int]),
    "ocsl_batch_get_launch": (ctypes.c_int, [c_handle, _i32p, _i32p]),
    "ocsl_batch_time_iterations": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_int, _f64p]),
    "ocsl_batch_tune": (ctypes.c_int, [c_handle, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p,
                                       ctypes.POINTER(TuneEntry), ctypes.c_int, _i32p]),
    "ocsl_build_structure": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_int64, _i64p, ctypes.c_void_p,
                                            ctypes.c_void_p]),
    "ocsl_time_stream_triad": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, _f64p, _f64p]),
    "ocsl_last_error": (ctypes.c_char_p, []),
}


def declare(lib: ctypes.CDLL) -> None:
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
