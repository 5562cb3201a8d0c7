"""ctypes binding of libmhx.so (include/mhx.h).

The shared library is built in-tree (mrhyde_amd/libmhx.so).  There is no
fallback: if it is missing or fails to load, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from .abi import *  # noqa: F401,F403
from .abi import BatchDesc, KrylovOpts, NewtonOpts, ProblemDesc, ServerDesc, SolveStats

LIB_PATH = Path(__file__).resolve().parent / "libmhx.so"


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
    lib = C.CDLL(str(LIB_PATH))
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    sig = {
        "mhx_version": (C.c_char_p, []),
        "mhx_last_error": (C.c_char_p, [vp]),
        "mhx_ctx_create": (C.c_int, [C.POINTER(ProblemDesc), C.POINTER(vp)]),
        "mhx_ctx_destroy": (None, [vp]),
        "mhx_sizes": (C.c_int, [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64), C.POINTER(i32)]),
        "mhx_crs_pattern": (C.c_int, [vp, vp, vp]),
        "mhx_dirichlet_mask": (C.c_int, [vp, vp]),
        "mhx_set_dirichlet": (C.c_int, [vp, vp, i32]),
        "mhx_dof_coords": (C.c_int, [vp, vp]),
        "mhx_assemble": (C.c_int, [vp, vp, i32, vp, vp]),
        "mhx_spmv": (C.c_int, [vp, vp, vp, vp, i32]),
        "mhx_linear_solve": (C.c_int, [vp, vp, vp, vp, C.POINTER(KrylovOpts), C.POINTER(SolveStats)]),
        "mhx_linear_solve_noalloc": (C.c_int, [vp, vp, vp, vp, C.POINTER(KrylovOpts), C.POINTER(SolveStats)]),
        "mhx_nonlinear_solve": (C.c_int, [vp, vp, C.POINTER(NewtonOpts), C.POINTER(SolveStats)]),
        "mhx_l2_error": (C.c_int, [vp, vp, i32, vp, vp, i32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "mhx_profile_enable": (C.c_int, [vp, i32]),
        "mhx_profile_read": (C.c_int, [vp, i32, vp, vp, vp]),
        "mhx_profile_reset": (C.c_int, [vp]),
        "mhx_synchronize": (C.c_int, [vp]),
        "mhx_set_stream": (C.c_int, [vp, vp]),
        "mhx_copy_probe": (C.c_int, [i64, i32, C.POINTER(C.c_double)]),
        "mhx_mfma_probe": (C.c_int, [i64, i32, i32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "mhx_batch_create": (C.c_int, [C.POINTER(BatchDesc), C.POINTER(vp)]),
        "mhx_batch_solve": (C.c_int, [vp, i32]),
        "mhx_batch_results": (C.c_int, [vp, vp, vp, vp, vp]),
        "mhx_batch_info": (C.c_int, [vp, C.POINTER(i64), C.POINTER(i32), C.POINTER(C.c_double)]),
        "mhx_batch_set_wg": (C.c_int, [vp, i32]),
        "mhx_batch_destroy": (None, [vp]),
        "mhx_set_option": (C.c_int, [vp, C.c_char_p, i64]),
        "mhx_get_option": (C.c_int, [vp, C.c_char_p, C.POINTER(i64)]),
        "mhx_state_ptr": (C.c_int, [vp, i32, C.POINTER(C.c_void_p), C.POINTER(i64)]),
        "mhx_state_upload": (C.c_int, [vp, i32, vp, i64]),
        "mhx_state_download": (C.c_int, [vp, i32, vp, i64]),
        "mhx_rng_fill": (C.c_int, [C.c_uint64, i64, i32, C.c_double, C.c_double, vp]),
        "mhx_footprint": (C.c_int, [C.POINTER(ProblemDesc), C.POINTER(KrylovOpts), C.POINTER(i64), C.POINTER(i64),
                                    C.POINTER(i64)]),
        "mhx_ctx_update": (C.c_int, [vp, C.POINTER(ProblemDesc)]),
        "mhx_server_create": (C.c_int, [C.POINTER(ServerDesc), C.POINTER(vp)]),
        "mhx_server_submit": (C.c_int, [vp, C.POINTER(ProblemDesc), C.POINTER(i64)]),
        "mhx_server_poll": (C.c_int, [vp, i64, i32, vp, C.POINTER(SolveStats), C.POINTER(C.c_double)]),
        "mhx_server_info": (C.c_int, [vp, vp]),
        "mhx_server_last_error": (C.c_char_p, [vp]),
        "mhx_server_destroy": (None, [vp]),
        "mhx_amg_setup": (C.c_int, [vp, vp, i32]),
        "mhx_set_time_state": (C.c_int, [vp, C.c_double, i32, vp, i32]),
        "mhx_set_stage_state": (C.c_int, [vp, C.c_double, i32, vp, i32, vp, vp, i32, vp, i32]),
        "mhx_write_matrix_market": (C.c_int, [vp, C.c_char_p, vp, i32]),
        "mhx_write_vector_market": (C.c_int, [vp, C.c_char_p, vp, i32]),
        "mhx_read_matrix_market": (C.c_int, [vp, C.c_char_p, vp, i32]),
        "mhx_read_vector_market": (C.c_int, [vp, C.c_char_p, vp, i32]),
        "mhx_amg_level": (C.c_int, [vp, i32, i32, C.POINTER(i64), C.POINTER(C.c_double), vp, vp, vp]),
        "mhx_expr_eval": (C.c_int, [C.POINTER(ProblemDesc), C.c_char_p, i64, vp, vp]),
        "mhx_expr_last_error": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()

# names declared in include/mhx.h (checked by the CPU test suite)
EXPORTED = [
    "mhx_ctx_create", "mhx_ctx_destroy", "mhx_last_error", "mhx_sizes", "mhx_crs_pattern",
    "mhx_dirichlet_mask", "mhx_set_dirichlet", "mhx_dof_coords", "mhx_assemble", "mhx_spmv", "mhx_linear_solve",
    "mhx_nonlinear_solve", "mhx_l2_error", "mhx_profile_enable", "mhx_profile_read", "mhx_profile_reset",
    "mhx_synchronize", "mhx_set_stream", "mhx_set_option", "mhx_get_option", "mhx_copy_probe", "mhx_mfma_probe", "mhx_batch_create", "mhx_batch_solve",
    "mhx_batch_results", "mhx_batch_info", "mhx_batch_set_wg", "mhx_batch_destroy","mhx_state_ptr", "mhx_state_upload", "mhx_state_download", "mhx_rng_fill", "mhx_version",
    "mhx_footprint", "mhx_ctx_update", "mhx_server_create", "mhx_server_submit", "mhx_server_poll",
    "mhx_server_info", "mhx_server_last_error", "mhx_server_destroy", "mhx_amg_setup", "mhx_amg_level",
    "mhx_set_time_state", "mhx_write_matrix_market", "mhx_write_vector_market", "mhx_read_matrix_market",
    "mhx_read_vector_market", "mhx_linear_solve_noalloc", "mhx_expr_eval", "mhx_expr_last_error",
    "mhx_set_stage_state",
]
