"""ctypes table of the surrogate's device kernels (include/mpf_nn.h) in
libmpf.so.  Pointers are device addresses (ints from torch tensors'
data_ptr()); the product path fails loudly without the library."""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import MPFError

NN_EXPORTS = (
    "mpf_nn_conv_fwd", "mpf_nn_conv_bwd_w", "mpf_nn_weights_bf16", "mpf_nn_relu_bwd", "mpf_nn_bias_grad",
    "mpf_nn_pool", "mpf_nn_upsample", "mpf_nn_copy_channels", "mpf_nn_adam", "mpf_nn_slot_features",
    "mpf_nn_slot_pool", "mpf_nn_head_in", "mpf_nn_slot_targets", "mpf_nn_mse", "mpf_nn_slot_apply",
    "mpf_nn_band_maxdiff", "mpf_nn_band_residual", "mpf_nn_to_bf16", "mpf_nn_fill", "mpf_nn_sum",
)

_bound = None


def lib():
    global _bound
    if _bound is not None:
        return _bound
    L = _lib.lib()
    vp, i, sz, f = C.c_void_p, C.c_int, C.c_size_t, C.c_float
    ip = C.POINTER(C.c_int)
    sig = {
        "mpf_nn_conv_fwd": [vp, vp, vp, vp, vp, ip, i, i, i, ip, i, vp],
        "mpf_nn_conv_bwd_w": [vp, vp, vp, ip, i, i, i, ip, vp],
        "mpf_nn_weights_bf16": [vp, vp, i, i, i, i, vp],
        "mpf_nn_relu_bwd": [vp, vp, sz, vp],
        "mpf_nn_bias_grad": [vp, vp, i, i, vp],
        "mpf_nn_pool": [vp, vp, ip, i, i, vp],
        "mpf_nn_upsample": [vp, vp, ip, i, i, vp],
        "mpf_nn_copy_channels": [vp, i, i, vp, i, i, i, i, i, vp],
        "mpf_nn_adam": [vp, vp, vp, vp, sz, f, f, f, f, i, vp],
        "mpf_nn_slot_features": [vp, i, i, i, i, i, i, i, vp, vp, vp, vp, vp, vp, vp],
        "mpf_nn_slot_pool": [vp, vp, vp, i, i, i, i, i, vp],
        "mpf_nn_head_in": [vp, i, vp, i, vp, i, i, i, vp],
        "mpf_nn_slot_targets": [vp, vp, i, i, i, i, i, i, vp, f, vp, vp, vp],
        "mpf_nn_mse": [vp, i, vp, vp, sz, f, vp, vp, vp],
        "mpf_nn_slot_apply": [vp, i, i, i, i, i, i, i, vp, vp, i, f, vp],
        "mpf_nn_band_maxdiff": [vp, vp, i, i, i, i, i, i, vp, vp],
        "mpf_nn_band_residual": [vp, vp, vp, i, i, i, i, i, i, f, vp, vp],
        "mpf_nn_to_bf16": [vp, vp, sz, vp],
        "mpf_nn_fill": [vp, f, sz, vp],
        "mpf_nn_sum": [vp, sz, vp, vp],
    }
    for name in NN_EXPORTS:
        fn = getattr(L, name)
        fn.argtypes = sig[name]
        fn.restype = C.c_int
    _bound = L
    return L


def call(name: str, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise MPFError(f"{name} failed ({'bad arguments' if rc == -1 else 'launch error'})")


def ivec(*v):
    return (C.c_int * len(v))(*v)
