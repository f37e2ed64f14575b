"""CPU oracle for the batch-SOM hot path.

TEST INFRASTRUCTURE ONLY: only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker or the timed CPU
baseline. The product package (``decontamination_amd``) never imports it.

Contents:
  * ``lib()``: ctypes binding of ``som_oracle.c`` (built by ``oracle/Makefile``), the C
    restatement of the reference's CPU kernels in two arithmetic modes, ``REF32`` (the
    reference's own FP32-difference / FP64-sum arithmetic and FP32 accumulators) and ``FP64``.
  * ``som_ref``: the restated host driver of ``SOM_Batch.train`` / ``SOM_Abstract`` on top of
    those kernels, plus a fast factorised FP64 oracle for large inputs.

Every function is pinned in ``tests/test_oracle.py`` against the reference's golden vectors
(test/test_som_batch.py:30-89, test/test_som_abstract.py:27-64), brute force on tiny inputs
and closed forms. Parity status per function is listed in that file's header.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libsomoracle.so")

REF32 = 0
FP64 = 1

_lib = None


def build():
    """Compile the C oracle (gcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: run `make -C {HERE}` ({LIB_PATH})")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    I64 = ctypes.c_int64
    D = ctypes.c_double
    F = ctypes.c_float
    sigs = {
        "orc_find_bmus": [I, P, I, I, P, I64, P],
        "orc_count_bmus": [I, P, I, I, P, P, I64, P],
        "orc_compute_errors": [I, P, P, I, I, P, P, I64, D, P],
        "orc_top2": [I, P, P, I, I, P, P, I64, P, P],
        "orc_train_epoch_accumulate": [I, P, P, P, P, P, P, I64, I64, I64, F, I, I],
        "orc_train_iter_accumulate": [I, P, P, P, P, P, P, I64, I64, I64, F, I, I],
        "orc_divide": [I, P, P, I, I, P],
        "orc_train_epoch_accumulate_omp": [P, P, P, P, P, P, I64, I64, I64, F, I, I],
        "orc_compute_errors_omp": [P, P, I, I, P, P, I64, D, P],
        "orc_welford_update": [P, P, P, P, P, P, I64, I],
        "orc_top2_omp": [I, P, I, I, P, I64, P, P, P],
        "orc_count_bmus_omp": [I, P, I, I, P, P, I64, P, P, I],
    }
    for name, argtypes in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = argtypes
        fn.restype = None
    _lib = L
    return L


def ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)
