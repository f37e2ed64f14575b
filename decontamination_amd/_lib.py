"""ctypes binding of libsomhip.so (declared in include/somhip.h).

The library is loaded from ``decontamination_amd/lib/libsomhip.so`` (built by
``python -m decontamination_amd.build``). There is no fallback: if the library or a gfx950 device is
missing, the calls raise.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsomhip.so")

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_UNSUPPORTED = -3
ERR_NODEVICE = -4
ERR_STATE = -5

TOPOLOGY_SQUARE = 0
TOPOLOGY_HEXAGONAL = 1

PREC_EXACT = 0
PREC_FP32 = 1
PREC_BF16_RERANK = 2

RERANK_REF32 = 0
RERANK_FP64 = 1

NOALLOC_MAX_ROWS = 1 << 21  # somhip_find_bmus_noalloc's row limit (SOMHIP_NOALLOC_MAX_ROWS)

SERVE_WINNERS = 0
SERVE_ACTIVATION = 1

TIMERS = {"bmu_exact": 0, "bmu_fp32": 1, "bmu_bf16": 2, "bmu_fallback": 3, "prep": 4, "errors": 5,
          "absmax": 6, "scatter": 7, "update": 8, "iter": 9}

PRECISIONS = {"exact": PREC_EXACT, "fp32": PREC_FP32, "bf16_rerank": PREC_BF16_RERANK}
ARITHMETICS = {"ref32": RERANK_REF32, "fp64": RERANK_FP64}

P = ctypes.c_void_p
I = ctypes.c_int
I64 = ctypes.c_int64
SZ = ctypes.c_size_t
F = ctypes.c_float
D = ctypes.c_double
PP = ctypes.POINTER(ctypes.c_void_p)

# name -> argtypes (every function returns int)
SIGNATURES = {
    "somhip_device_count": [ctypes.POINTER(ctypes.c_int)],
    "somhip_ctx_create": [I, PP],
    "somhip_ctx_destroy": [P],
    "somhip_ctx_synchronize": [P],
    "somhip_ctx_set_stream": [P, P],
    "somhip_ctx_get_stream": [P, PP],
    "somhip_ctx_set_timing": [P, I],
    "somhip_ctx_get_timing": [P, I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)],
    "somhip_malloc": [P, SZ, PP],
    "somhip_free": [P, P],
    "somhip_memcpy_h2d": [P, P, P, SZ],
    "somhip_memcpy_d2h": [P, P, P, SZ],
    "somhip_memcpy_d2d": [P, P, P, SZ],
    "somhip_memset": [P, P, I, SZ],
    "somhip_uploader_create": [P, I, PP],
    "somhip_uploader_destroy": [P],
    "somhip_uploader_put": [P, P, SZ, P, SZ, PP, PP],
    "somhip_model_create": [P, I, I, I, I, P, PP],
    "somhip_model_destroy": [P],
    "somhip_model_set_weights": [P, P, I],
    "somhip_model_get_weights": [P, P],
    "somhip_model_weights_ptr": [P, PP],
    "somhip_find_bmus": [P, P, I64, I, I, P],
    "somhip_find_bmus_noalloc": [P, P, I64, I, I, P],
    "somhip_count_bmus": [P, P, P, I64, I, I, P],
    "somhip_compute_errors": [P, P, P, I64, I, I, P],
    "somhip_gndm": [P, P, P, I64, I, P, P, P, P],
    "somhip_covariance_chunk": [P, P, I, P, I64, I, P],
    "somhip_accum_create": [P, PP],
    "somhip_accum_destroy": [P],
    "somhip_accum_reset": [P],
    "somhip_batch_epoch_accumulate": [P, P, P, I64, I, I, P],
    "somhip_batch_epoch_update": [P, I64, I64, F, P],
    "somhip_batch_iter_accumulate": [P, P, P, I64, I64, I64, F, I, I],
    "somhip_batch_iter_update": [P, P],
    "somhip_model_create_f64": [P, I, I, I, I, P, PP],
    "somhip_model_set_weights_f64": [P, P, I],
    "somhip_model_get_weights_f64": [P, P],
    "somhip_find_bmus_f64": [P, P, I64, P],
    "somhip_count_bmus_f64": [P, P, P, I64, P],
    "somhip_compute_errors_f64": [P, P, P, I64, P],
    "somhip_batch_epoch_accumulate_f64": [P, P, P, I64, P],
    "somhip_batch_epoch_update_f64": [P, I64, I64, D, P],
    "somhip_batch_iter_accumulate_f64": [P, P, P, I64, I64, I64, D],
    "somhip_batch_iter_update_f64": [P, P],
    "somhip_host_alloc": [SZ, PP],
    "somhip_host_free": [P],
    "somhip_server_create": [P, SZ, I, PP],
    "somhip_server_destroy": [P],
    "somhip_server_add_model": [P, P, ctypes.POINTER(ctypes.c_int)],
    "somhip_server_submit": [P, I, I, P, P, I64, P, ctypes.POINTER(ctypes.c_int64)],
    "somhip_server_wait": [P, I64, ctypes.c_double, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)],
    "somhip_server_stats": [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                            ctypes.POINTER(ctypes.c_size_t)],
}
STRING_FUNCS = ("somhip_last_error", "somhip_version")

_lib = None
_lock = threading.Lock()


class SomHipError(RuntimeError):
    """Raised when a libsomhip call fails (HIP error, unsupported shape, missing device)."""


def load():
    """Load libsomhip.so; raises if it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise SomHipError(
                f"libsomhip.so not found at {LIB_PATH}; build it with `python -m decontamination_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, argtypes in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
        for name in STRING_FUNCS:
            fn = getattr(L, name)
            fn.argtypes = []
            fn.restype = ctypes.c_char_p
        _lib = L
        return L


def exported_symbols():
    return list(SIGNATURES) + list(STRING_FUNCS)


def last_error():
    return load().somhip_last_error().decode(errors="replace")


def check(rc, what=""):
    """Map a libsomhip return code to the reference's exception types."""
    if rc == OK:
        return
    msg = f"{what}: {last_error()}" if what else last_error()
    if rc == ERR_ARG:
        raise ValueError(msg)
    if rc == ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise SomHipError(f"libsomhip error {rc}: {msg}")


def call(name, *args):
    check(getattr(load(), name)(*args), name)


def vp(x):
    """ctypes void pointer of a numpy array (host) or an int device address."""
    if x is None:
        return None
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    if isinstance(x, np.ndarray):
        return x.ctypes.data_as(ctypes.c_void_p)
    if hasattr(x, "ptr"):
        return ctypes.c_void_p(x.ptr)
    raise TypeError(f"cannot take a pointer of {type(x)}")


def device_count():
    c = ctypes.c_int(0)
    call("somhip_device_count", ctypes.byref(c))
    return c.value
