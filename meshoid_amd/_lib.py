"""ctypes binding of libmeshoid_hip.so (C ABI in include/meshoid_hip.h).

The library is the only compute path: if it is missing or cannot be loaded,
importing this module raises ImportError (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MESHOID_HIP_LIB", os.path.join(HERE, "libmeshoid_hip.so"))

ERR_VALUE, ERR_HIP, ERR_SINGULAR, ERR_UNSUPPORTED, ERR_KEY = -1, -2, -3, -4, -5
ST_AMBIGUOUS, ST_BAND_OVERFLOW, ST_HSML_NOCONV, ST_SINGULAR, ST_COUNT, ST_SEED_MISS = (
    1, 2, 4, 8, 16, 32)
ST_RANKDEF = 64
ST_RADIUS_MISS = 128
KNN_KEEP_DIST = 1
KNN_SMALL_RECORD = 2  # test hook: every wavefront takes the leaf-record overflow path
NOALLOC_INTEGRATE_MAX_K = 64

EXPORTED = [
    "mhip_version", "mhip_ctx_create", "mhip_ctx_destroy", "mhip_ctx_set_stream",
    "mhip_ctx_synchronize", "mhip_last_error", "mhip_index_create", "mhip_index_build",
    "mhip_index_destroy", "mhip_index_info", "mhip_nlist_create", "mhip_nlist_destroy",
    "mhip_knn_hsml", "mhip_nlist_export", "mhip_apply_D", "mhip_gradient_weights",
    "mhip_hsml_iter", "mhip_kernel", "mhip_prof_enable", "mhip_prof_read", "mhip_seed_features",
    "mhip_seed_predict", "mhip_mlp_train_step", "mhip_mlp_pack", "mhip_mlp_blob_bytes",
    "mhip_kernel_weights", "mhip_kernel_average", "mhip_kernel_variance", "mhip_integrate",
    "mhip_batch_create", "mhip_batch_run", "mhip_batch_info", "mhip_batch_destroy",
    "mhip_batch_tuning", "mhip_batch_force",
    "mhip_index_order", "mhip_nearest", "mhip_unique_inverse", "mhip_reconstruct", "mhip_nlist_subset",
    "mhip_surface_density", "mhip_grid_density", "mhip_radtransfer", "mhip_nlist_export_f64",
    "mhip_nlist_hsml", "mhip_integrate_noalloc",
]

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} not found: build it with `python -m meshoid_amd.build` "
        "(hipcc, gfx950). meshoid_amd has no CPU fallback.")

lib = ctypes.CDLL(LIB_PATH)

_p = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double
_pp = ctypes.POINTER(ctypes.c_void_p)

_SIGS = {
    "mhip_version": ([], _i),
    "mhip_ctx_create": ([_i, _p, _pp], _i),
    "mhip_ctx_destroy": ([_p], _i),
    "mhip_ctx_set_stream": ([_p, _p], _i),
    "mhip_ctx_synchronize": ([_p], _i),
    "mhip_last_error": ([_p], ctypes.c_char_p),
    "mhip_index_create": ([_pp], _i),
    "mhip_index_build": ([_p, _p, _p, _i, _i64, _i, _d], _i),
    "mhip_index_destroy": ([_p], _i),
    "mhip_index_info": ([_p, _p, _p, _p, _p], _i),
    "mhip_nlist_create": ([_pp], _i),
    "mhip_nlist_destroy": ([_p], _i),
    "mhip_knn_hsml": ([_p, _p, _p, _i64, _i, _p, _p, _p, _p, _p, _p, _p, _i, _p], _i),
    "mhip_nlist_export": ([_p, _p, _p, _p], _i),
    "mhip_nlist_export_f64": ([_p, _p, _p, _p], _i),
    "mhip_nlist_hsml": ([_p, _p, _p], _i),
    "mhip_apply_D": ([_p, _p, _p, _i, _i64, _i, _i, _i, _p, _p, _p], _i),
    "mhip_gradient_weights": ([_p, _p, _i, _i64, _i, _p, _i64, _i, _p, _p, _d, _i, _i, _p, _p],
                              _i),
    "mhip_hsml_iter": ([_p, _p, _i, _i64, _i, _i, _d, _p, _p], _i),
    "mhip_kernel": ([_p, _p, _i, _i64, _p], _i),
    "mhip_prof_enable": ([_p, _i], _i),
    "mhip_prof_read": ([_p, ctypes.c_char_p, _i64], _i),
    "mhip_seed_features": ([_p, _p, _p], _i),
    "mhip_seed_predict": ([_p, _p, _p, _p], _i),
    "mhip_mlp_train_step": ([_p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, _p, _i, ctypes.c_float], _i),
    "mhip_mlp_pack": ([_p, _p, _p, _p, _p], _i),
    "mhip_mlp_blob_bytes": ([], _i64),
    "mhip_kernel_weights": ([_p, _p, _i, _p, _i, _i64, _i, _p], _i),
    "mhip_kernel_average": ([_p, _p, _p, _i, _i64, _p], _i),
    "mhip_kernel_variance": ([_p, _p, _p, _i, _i64, _i, _p], _i),
    "mhip_integrate": ([_p, _p, _i, _p, _i, _i64, _i, _p], _i),
    "mhip_integrate_noalloc": ([_p, _p, _i, _p, _i, _i64, _i, _p], _i),
    "mhip_batch_create": ([_p, _i, _p, _p, _i, _i, _pp], _i),
    "mhip_batch_run": ([_p, _p, _p, _p, _p, _p, _p, _p, _i], _i),
    "mhip_batch_info": ([_p, _p, _p, _p, _p, _p], _i),
    "mhip_batch_destroy": ([_p], _i),
    "mhip_batch_tuning": ([_p, _p, _p, _p, _p, _p, _p, _p], _i),
    "mhip_batch_force": ([_p, _i, _i], _i),
    "mhip_index_order": ([_p, _p, _p], _i),
    "mhip_nearest": ([_p, _p, _p, _i64, _i, _p], _i),
    "mhip_nlist_subset": ([_p, _p, _p, _i64, _p, _p, _p, _p, _p, _p], _i),
    "mhip_surface_density": ([_p, _p, _i, _p, _i, _i64, _i, _p, _i, _d, _d, _d, _d, _d, _i, _i, _p], _i),
    "mhip_grid_density": ([_p, _p, _i, _p, _i, _i64, _p, _i, _d, _d, _d, _d, _d, _i, _i, _p], _i),
    "mhip_radtransfer": ([_p, _p, _p, _p, _p, _p, _i64, _i, _i, _d, _d, _d, _p, _p], _i),
    "mhip_unique_inverse": ([_p, _p, _i64, _i64, _p, _p, _p], _i),
    "mhip_reconstruct": ([_p, _p, _i64, _i, _d, _p, _i, _p, _p, _p, _i, _i, _p, _p, _i, _p], _i),
}
for _name, (_args, _res) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.argtypes = _args
    _fn.restype = _res


def raise_for(rc: int, ctx=None, what: str = ""):
    """Map a MHIP_ERR_* code to the exception the reference raises."""
    if rc == 0:
        return
    msg = lib.mhip_last_error(ctx).decode() if ctx else what
    if what and msg != what:
        msg = f"{what}: {msg}"
    if rc == ERR_VALUE:
        raise ValueError(msg)
    if rc == ERR_SINGULAR:
        raise np.linalg.LinAlgError(msg)
    if rc == ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == ERR_KEY:
        raise KeyError(msg)
    raise RuntimeError(msg)
