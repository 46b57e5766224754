"""ctypes binding of libskema_mi355x.so (the C ABI in include/skema_mi355x.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C skema_amd/csrc``).  There is no fallback: if the library is missing
every call raises ``SkemaError``.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libskema_mi355x.so")

SKEMA_OK = 0
STATUS_NAMES = {
    1: "SKEMA_ERR_ARG",
    2: "SKEMA_ERR_SHAPE",
    3: "SKEMA_ERR_HIP",
    4: "SKEMA_ERR_OOM",
    5: "SKEMA_ERR_NUMERIC",
    6: "SKEMA_ERR_STATE",
    7: "SKEMA_ERR_UNSUPPORTED",
}

MAP_GAUSS, MAP_SPARSE_SIGN = 0, 1
KERNEL_NONE, KERNEL_GAUSSRBF = 0, 1
PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_FP8, PREC_BF16 = 0, 1, 2, 3, 4
DIST_NORMAL, DIST_UNIFORM = 0, 1
SSIGN_ROWS, SSIGN_COLS = 0, 1

PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3, "fp8": PREC_FP8, "bf16": PREC_BF16}


class SkemaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class skema_params(C.Structure):
    _fields_ = [
        ("m", C.c_int64),
        ("n", C.c_int64),
        ("rank", C.c_int64),
        ("range", C.c_int64),
        ("core", C.c_int64),
        ("window", C.c_int64),
        ("eta", C.c_double),
        ("nu", C.c_double),
        ("dim_redux", C.c_int32),
        ("issparse", C.c_int32),
        ("kernel", C.c_int32),
        ("precision", C.c_int32),
        ("gamma", C.c_double),
        ("nfeat", C.c_int64),
    ]


_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int
_d = C.c_double
_pp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes)
SIGNATURES = {
    "skema_params_default": (None, [C.POINTER(skema_params)]),
    "skema_last_error": (C.c_char_p, []),
    "skema_version": (_i32, []),
    "skema_ctx_create": (_i32, [_i32, _pp]),
    "skema_ctx_destroy": (_i32, [_vp]),
    "skema_ctx_set_stream": (_i32, [_vp, _vp]),
    "skema_ctx_set_async": (_i32, [_vp, _i32]),
    "skema_ctx_synchronize": (_i32, [_vp]),
    "skema_ctx_set_profiling": (_i32, [_vp, _i32]),
    "skema_ctx_set_profiling_only": (_i32, [_vp, C.c_char_p]),
    "skema_ctx_kernel_stats": (_i32, [_vp, C.c_char_p, C.POINTER(_i64), C.POINTER(_d)]),
    "skema_ctx_reset_stats": (_i32, [_vp]),
    "skema_map_dense": (_i32, [_vp, _i64, _i64, _vp, _i64, _i32, _pp]),
    "skema_map_dense_device": (_i32, [_vp, _i64, _i64, _vp, _i64, _i32, _pp]),
    "skema_map_gauss_seeded": (_i32, [_vp, _i64, _i64, C.c_uint64, _i32, _i32, _pp]),
    "skema_map_csr": (_i32, [_vp, _i64, _i64, _i64, _vp, _vp, _vp, _i32, _pp]),
    "skema_map_sparse_sign_seeded": (_i32, [_vp, _i64, _i64, C.c_uint64, _i32, _i64, _i32, _pp]),
    "skema_map_destroy": (_i32, [_vp]),
    "skema_map_info": (_i32, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i32), C.POINTER(_i32)]),
    "skema_map_download": (_i32, [_vp, _vp, _vp, _vp]),
    "skema_map_lmap": (_i32, [_vp, _d, _vp, _i64, _i64, _i64, _d, C.c_char, C.c_char, _i64, _i64, _vp, _i64]),
    "skema_map_rmap": (_i32, [_vp, _d, _vp, _i64, _i64, _i64, _d, C.c_char, C.c_char, _i64, _i64, _vp, _i64]),
    "skema_rbf_window": (_i32, [_vp, _vp, _i64, _i64, _i64, _d, _i64, _i64, _vp, _i64]),
    "skema_sketch_create": (_i32, [_vp, C.POINTER(skema_params), _vp, _vp, _vp, _vp, _pp]),
    "skema_sketch_destroy": (_i32, [_vp]),
    "skema_sketch_sizes": (_i32, [_vp, C.POINTER(_i64), C.POINTER(_i64)]),
    "skema_sketch_reset": (_i32, [_vp]),
    "skema_update_dense": (_i32, [_vp, _vp, _i64, _i64, _i64]),
    "skema_update_csr": (_i32, [_vp, _vp, _vp, _vp, _i64, _i64]),
    "skema_update_rbf": (_i32, [_vp, _vp, _i64, _i64, _i64]),
    "skema_linear_update_dense": (_i32, [_vp, _vp, _i64]),
    "skema_linear_update_csr": (_i32, [_vp, _vp, _vp, _vp]),
    "skema_linear_update_rbf": (_i32, [_vp, _vp, _i64]),
    "skema_get_state": (_i32, [_vp, _vp, _vp, _vp]),
    "skema_set_state": (_i32, [_vp, _vp, _vp, _vp]),
    "skema_state_device": (_i32, [_vp, _pp, _pp, _pp]),
    "skema_low_rank_approx": (_i32, [_vp, _vp, _vp, _vp]),
    "skema_low_rank_approx_device": (_i32, [_vp, _vp, _i64, _vp, _vp, _i64]),
    "skema_get_checkpoint": (_i32, [_vp, C.c_char_p, _vp, C.POINTER(_i64), C.POINTER(_i64)]),
    "skema_sketch_timings_json": (_i32, [_vp, C.c_char_p, C.c_size_t]),
    "skema_residual_fro_dense": (_i32, [_vp, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _i64,
                                        C.POINTER(_d), C.POINTER(_d)]),
    "skema_residual_fro_dense_noalloc": (_i32, [_vp, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64,
                                                _i64, C.POINTER(_d), C.POINTER(_d)]),
    "skema_spd_create": (_i32, [_vp, C.POINTER(skema_params), _vp, _pp]),
    "skema_spd_destroy": (_i32, [_vp]),
    "skema_spd_update_dense": (_i32, [_vp, _vp, _i64, _i64, _i64]),
    "skema_spd_update_csr": (_i32, [_vp, _vp, _vp, _vp, _i64, _i64]),
    "skema_spd_update_rbf": (_i32, [_vp, _vp, _i64, _i64, _i64]),
    "skema_spd_linear_update_dense": (_i32, [_vp, _vp, _i64]),
    "skema_spd_linear_update_csr": (_i32, [_vp, _vp, _vp, _vp]),
    "skema_spd_linear_update_rbf": (_i32, [_vp, _vp, _i64]),
    "skema_spd_get_state": (_i32, [_vp, _vp]),
    "skema_spd_timings_json": (_i32, [_vp, C.c_char_p, C.c_size_t]),
    "skema_spd_low_rank_approx": (_i32, [_vp, _i32, _vp, _vp]),
    "skema_spd_low_rank_approx_device": (_i32, [_vp, _i32, _vp, _i64, _vp]),
    "skema_residuals_dense": (_i32, [_vp, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "skema_residuals_csr": (_i32, [_vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _i64, _vp]),
    "skema_residuals_rbf": (_i32, [_vp, _vp, _i64, _i64, _i64, _d, _i64, _vp, _i64, _vp, _vp, _i64, _i64, _vp]),
    "skema_residuals_sym_dense": (_i32, [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp]),
    "skema_residuals_sym_csr": (_i32, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    "skema_residuals_sym_rbf": (_i32, [_vp, _vp, _i64, _i64, _i64, _d, _i64, _vp, _i64, _vp, _i64, _vp]),
    "skema_rbf_sketch_loss_grad": (_i32, [_vp, _vp, _i64, _i64, _i64, _d, _vp, _vp, _vp, _i64, _i32,
                                          C.POINTER(_d), _vp, _i64, C.POINTER(_d)]),
}

_lib: Optional[C.CDLL] = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the library (no GPU call is made by loading)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SkemaError(7, f"{path} is not built; run __graft_entry__.build() (make -C skema_amd/csrc)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != SKEMA_OK:
        msg = load().skema_last_error()
        raise SkemaError(code, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
