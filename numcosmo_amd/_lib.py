"""ctypes binding of libnckde.so (the C-ABI declared in include/nckde.h).

There is no fallback: if the HIP library is missing or cannot be loaded,
``lib()`` raises, and every object of the package fails with it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnckde.so")

_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_h = C.c_void_p
_sz = C.c_size_t

# name -> (restype, argtypes); the complete exported surface of nckde.h
SIGNATURES = {
    "nckde_last_error": (C.c_char_p, []),
    "nckde_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_double, C.c_uint, C.c_int]),
    "nckde_destroy": (C.c_int, [_h]),
    "nckde_set_over_smooth": (C.c_int, [_h, C.c_double]),
    "nckde_set_shrink": (C.c_int, [_h, C.c_double]),
    "nckde_set_local_frac": (C.c_int, [_h, C.c_double]),
    "nckde_set_use_rot_href": (C.c_int, [_h, C.c_int]),
    "nckde_set_cov_type": (C.c_int, [_h, C.c_int]),
    "nckde_set_cov_fixed": (C.c_int, [_h, _dp]),
    "nckde_set_nearPD_maxiter": (C.c_int, [_h, C.c_uint]),
    "nckde_set_use_threads": (C.c_int, [_h, C.c_int]),
    "nckde_get_use_threads": (C.c_int, [_h, C.POINTER(C.c_int)]),
    "nckde_get_shrink": (C.c_int, [_h, C.POINTER(C.c_double)]),
    "nckde_get_local_frac": (C.c_int, [_h, C.POINTER(C.c_double)]),
    "nckde_get_use_rot_href": (C.c_int, [_h, C.POINTER(C.c_int)]),
    "nckde_get_nearPD_maxiter": (C.c_int, [_h, C.POINTER(C.c_uint)]),
    "nckde_get_cov_type": (C.c_int, [_h, C.POINTER(C.c_int)]),
    "nckde_get_cov_fixed": (C.c_int, [_h, _dp]),
    "nckde_get_over_smooth": (C.c_int, [_h, C.POINTER(C.c_double)]),
    "nckde_set_split_frac": (C.c_int, [_h, C.c_double]),
    "nckde_get_split_frac": (C.c_int, [_h, C.POINTER(C.c_double)]),
    "nckde_set_print_fit": (C.c_int, [_h, C.c_int]),
    "nckde_get_print_fit": (C.c_int, [_h, C.POINTER(C.c_int)]),
    "nckde_set_cv_type": (C.c_int, [_h, C.c_int]),
    "nckde_get_cv_type": (C.c_int, [_h, C.POINTER(C.c_int)]),
    "nckde_get_cv_stats": (C.c_int, [_h, C.POINTER(C.c_uint), C.POINTER(C.c_double)]),
    "nckde_get_mixed_check": (C.c_int, [_h, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "nckde_set_mixed_tolerance": (C.c_int, [_h, C.c_double]),
    "nckde_set_graphs": (C.c_int, [_h, C.c_int]),
    "nckde_set_eval_split": (C.c_int, [_h, C.c_uint]),
    "nckde_set_nnls_umethod": (C.c_int, [_h, C.c_int]),
    "nckde_get_nnls_umethod": (C.c_int, [_h, C.POINTER(C.c_int)]),
    "nckde_set_nnls_refine": (C.c_int, [_h, C.c_uint]),
    "nckde_get_nnls_refine": (C.c_int, [_h, C.POINTER(C.c_uint)]),
    "nckde_get_nnls_rungs": (C.c_int, [_h, C.POINTER(C.c_uint)]),
    "nckde_add_obs": (C.c_int, [_h, _dp]),
    "nckde_add_obs_batch": (C.c_int, [_h, _dp, _sz, _sz]),
    "nckde_reset": (C.c_int, [_h]),
    "nckde_prepare": (C.c_int, [_h]),
    "nckde_prepare_interp": (C.c_int, [_h, _dp, _sz]),
    "nckde_compute_IM": (C.c_int, [_h, _dp]),
    "nckde_eval_m2lnp": (C.c_int, [_h, _dp, _sz, _sz, _dp, C.c_int]),
    "nckde_eval": (C.c_int, [_h, _dp, _sz, _sz, _dp, C.c_int]),
    "nckde_eval_m2lnp_device": (C.c_int, [_h, C.c_void_p, _sz, C.c_void_p, C.c_int, C.c_void_p]),
    "nckde_eval_m2lnp_device_noalloc": (C.c_int, [_h, C.c_void_p, _sz, C.c_void_p, C.c_int, C.c_void_p]),
    "nckde_eval_m2lnp_multi": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(_sz),
                                         C.c_void_p, C.POINTER(C.c_void_p), _sz, C.c_int]),
    "nckde_noalloc_scratch_bytes": (_sz, []),
    "nckde_kernel_choose": (C.c_int, [_h, _dp, _sz, _up]),
    "nckde_sample": (C.c_int, [_h, _dp, _dp, C.c_void_p, _sz, _dp, C.c_void_p]),
    "nckde_get_dim": (C.c_int, [_h, C.POINTER(C.c_uint)]),
    "nckde_get_sample_size": (C.c_int, [_h, C.POINTER(_sz)]),
    "nckde_get_n_kernels": (C.c_int, [_h, C.POINTER(_sz)]),
    "nckde_get_n_samples": (C.c_int, [_h, C.POINTER(_sz)]),
    "nckde_get_href": (C.c_int, [_h, C.POINTER(C.c_double)]),
    "nckde_get_rnorm": (C.c_int, [_h, C.POINTER(C.c_double)]),
    "nckde_get_lnnorm": (C.c_int, [_h, _sz, C.POINTER(C.c_double)]),
    "nckde_get_weights": (C.c_int, [_h, _dp]),
    "nckde_get_cov_decomp": (C.c_int, [_h, _sz, _dp]),
    "nckde_get_nnls_state": (C.c_int, [_h, C.c_void_p, C.c_void_p, C.c_void_p]),
    "nckde_get_Ki": (C.c_int, [_h, _sz, _dp, _dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "nckde_get_full_cov": (C.c_int, [_h, _dp]),
    "nckde_get_full_cov_decomp": (C.c_int, [_h, _dp]),
    "nckde_get_sample": (C.c_int, [_h, _dp]),
    "nckde_get_invUsample": (C.c_int, [_h, _dp]),
    "nckde_get_knn": (C.c_int, [_h, C.POINTER(_sz), C.c_void_p]),
    "nckde_serialize": (C.c_int, [_h, C.c_int, C.c_void_p, _sz, C.POINTER(_sz)]),
    "nckde_deserialize": (C.c_int, [C.POINTER(C.c_void_p), C.c_void_p, _sz]),
    "nckde_set_profiling": (C.c_int, [_h, C.c_int]),
    "nckde_get_kernel_ms": (C.c_int, [_h, C.POINTER(C.c_float)]),
    "nckde_apes_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint, C.c_uint, C.c_int, C.c_int, C.c_uint64]),
    "nckde_apes_destroy": (C.c_int, [_h]),
    "nckde_apes_set_over_smooth": (C.c_int, [_h, C.c_double]),
    "nckde_apes_set_shrink": (C.c_int, [_h, C.c_double]),
    "nckde_apes_set_local_frac": (C.c_int, [_h, C.c_double]),
    "nckde_apes_set_random_walk_prob": (C.c_int, [_h, C.c_double]),
    "nckde_apes_set_use_interp": (C.c_int, [_h, C.c_int]),
    "nckde_apes_set_exploration": (C.c_int, [_h, C.c_uint]),
    "nckde_apes_set_precision": (C.c_int, [_h, C.c_int]),
    "nckde_apes_set_bounds": (C.c_int, [_h, _dp, _dp]),
    "nckde_apes_init": (C.c_int, [_h, _dp, _dp]),
    "nckde_apes_run": (C.c_int, [_h, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]),
    "nckde_apes_run_from": (C.c_int, [_h, C.c_uint, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]),
    "nckde_apes_get_state": (C.c_int, [_h, _dp, _dp]),
    "nckde_apes_get_stats": (C.c_int, [_h, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), _dp]),
    "nckde_apes_set_trace": (C.c_int, [_h, C.c_void_p, _sz]),
    "nckde_apes_get_trace_rows": (C.c_int, [_h, C.POINTER(C.c_size_t)]),
    "nckde_server_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint, _sz]),
    "nckde_server_destroy": (C.c_int, [_h]),
    "nckde_server_submit": (C.c_int, [_h, C.c_void_p, C.POINTER(C.c_uint64)]),
    "nckde_server_wait": (C.c_int, [_h, C.c_uint64, C.c_void_p]),
    "nckde_server_clock": (C.c_double, [_h]),
    "nckde_request_bytes": (C.c_int, [C.c_void_p, C.POINTER(_sz)]),
}


class Request(C.Structure):
    """nckde_request"""
    _fields_ = [("op", C.c_int), ("h", C.c_void_p), ("m", C.c_size_t), ("X", C.c_void_p), ("m2lnp", C.c_void_p),
                ("u", C.c_void_p), ("z", C.c_void_p), ("chi2", C.c_void_p), ("out", C.c_void_p),
                ("idx", C.c_void_p), ("precision", C.c_int)]


class RequestInfo(C.Structure):
    """nckde_request_info"""
    _fields_ = [("status", C.c_int), ("t_submit", C.c_double), ("t_start", C.c_double), ("t_end", C.c_double),
                ("bytes", C.c_size_t), ("message", C.c_char * 256), ("batch", C.c_uint)]

# nckde_m2lnL_batch_fn
M2LNL_BATCH_FN = C.CFUNCTYPE(C.c_int, C.POINTER(C.c_double), C.c_size_t, C.c_uint, C.POINTER(C.c_double),
                             C.c_void_p)


class NckdeError(RuntimeError):
    """A failing nckde call; the message is the library's (the reference's
    g_error text where the reference aborts)."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m numcosmo_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int):
    if status != 0:
        raise NckdeError(status, lib().nckde_last_error().decode(errors="replace"))
