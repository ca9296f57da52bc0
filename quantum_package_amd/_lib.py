"""Loader of libqpx.so (in-tree).  Fails loudly when the library is missing:
there is no CPU fallback for the product path."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqpx.so")

c_dp = C.POINTER(C.c_double)
c_u64p = C.POINTER(C.c_uint64)
c_i64p = C.POINTER(C.c_int64)
c_ip = C.POINTER(C.c_int)
c_vp = C.c_void_p

# (name, restype, argtypes) of every entry point declared in include/*.h
SIGNATURES = [
    ("qpx_create", C.c_int, [C.c_int, C.POINTER(c_vp)]),
    ("qpx_make_s2_eigenfunction", C.c_int, [C.c_int, C.c_int64, c_u64p, C.c_int, C.c_int64, c_u64p, c_i64p]),
    ("qpx_s2_complete", C.c_int, [c_vp, C.c_int, C.c_int64, c_u64p, C.c_int, C.c_int64, c_u64p, c_i64p]),
    ("qpx_select_generators", C.c_int, [c_vp, C.c_int, c_dp, C.c_int, c_dp, C.c_int, c_ip, c_dp]),
    ("qpx_select_tasks", C.c_int, [c_vp, C.c_int, c_dp, C.c_int, c_dp, C.c_int64, C.c_int, c_ip, c_ip, c_dp,
                                   c_u64p, c_dp, c_i64p, c_dp]),
    ("qpx_destroy", C.c_int, [c_vp]),
    ("qpx_last_error", C.c_int, [c_vp, C.c_char_p, C.c_int]),
    ("qpx_set_verbosity", C.c_int, [c_vp, C.c_int]),
    ("qpx_set_hamiltonian", C.c_int, [c_vp, C.c_int, C.c_int, C.c_int, c_dp, c_dp, C.c_int64]),
    ("qpx_set_wavefunction", C.c_int, [c_vp, C.c_int, C.c_int64, c_u64p]),
    ("qpx_sigma", C.c_int, [c_vp, C.c_int, C.c_int, C.c_int64, c_dp, c_dp, c_dp]),
    ("qpx_sigma_device", C.c_int, [c_vp, C.c_int, C.c_int, c_vp, c_vp, c_vp, c_vp]),
    ("qpx_get_bilinear_order", C.c_int, [c_vp, c_i64p]),
    ("qpx_diag_h", C.c_int, [c_vp, c_dp]),
    ("qpx_davidson", C.c_int, [c_vp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_double, C.c_uint64, c_dp, C.c_int64, c_dp, c_dp, c_ip]),
    ("qpx_select", C.c_int, [c_vp, C.c_int, c_dp, C.c_int, C.c_int, c_dp, C.c_int64, C.c_int, c_dp, c_u64p,
                             c_dp, c_i64p]),
    ("qpx_select_stats", C.c_int, [c_vp, c_i64p]),
    # include/qpx_proposal.h
    ("qpx_select_blocks", C.c_int, [c_vp, C.c_int, c_dp, C.c_int, C.c_int, c_dp, C.c_int64,
                                    C.POINTER(C.c_uint8), c_dp, c_u64p, c_dp, c_i64p, c_dp, c_i64p]),
    ("qpx_select_count_blocks", C.c_int, [c_vp, C.c_int, c_i64p]),
    ("qpx_mlp_create", C.c_int, [C.c_int, C.c_int, C.c_uint64, C.POINTER(c_vp)]),
    ("qpx_mlp_destroy", C.c_int, [c_vp]),
    ("qpx_mlp_last_error", C.c_int, [c_vp, C.c_char_p, C.c_int]),
    ("qpx_mlp_train_step", C.c_int, [c_vp, C.c_int, c_ip, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                                     C.POINTER(C.c_float)]),
    ("qpx_mlp_predict", C.c_int, [c_vp, C.c_int, c_ip, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("qpx_mlp_predict_static", C.c_int, [c_vp, C.c_int, c_ip, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("qpx_mlp_train_epochs", C.c_int, [c_vp, C.c_int, c_ip, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                                       C.c_int, C.c_int, C.c_uint64, C.POINTER(C.c_float)]),
    ("qpx_mlp_param", C.c_int, [c_vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int64]),
    ("qpx_mlp_gemm_check", C.c_int, [c_vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                     C.POINTER(C.c_float)]),
]

# include/qpx_debug.h: test and benchmark hooks
DEBUG = [
    ("qpx_debug_force_sigma_path", C.c_int, [c_vp, C.c_int, c_ip]),
    ("qpx_debug_sigma_kernel_times", C.c_int, [c_vp, C.c_int, C.c_int, c_vp, c_vp, c_vp, c_vp,
                                               C.POINTER(C.c_float), C.c_int]),
    ("qpx_debug_sparse_stats", C.c_int, [c_vp, c_i64p]),
    ("qpx_debug_lds_bandwidth", C.c_int, [c_vp, c_dp]),
    ("qpx_debug_select_block_ticks", C.c_int, [c_vp, c_i64p, C.c_int64]),
    ("qpx_debug_mlp_gemm", C.c_int, [c_vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                     C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]),
]

_LIB = None


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libqpx.so not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES + DEBUG:
        f = getattr(L, name)  # AttributeError: the library misses a declared entry point
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L
