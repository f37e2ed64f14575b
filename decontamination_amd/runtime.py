"""Host runtime over libsomhip: contexts, device arrays, resident models and accumulators.

Replaces the reference's launch layer (decontamination/jit/__init__.py:96-412): ``DeviceArray``,
``device_array_{empty,zeros}``, per-launch ``cu.to_device`` copies and ``copy_to_host``. Here a
dataset is uploaded once and stays resident in HBM across epochs; weights live in a ``Model`` on
the device.
"""
import ctypes
import threading

import numpy as np

from . import _lib
from ._lib import call, vp

_contexts = {}
_ctx_lock = threading.Lock()


class Context:
    """One device + one HIP stream (libsomhip context)."""

    def __init__(self, device=0):
        _lib.load()
        h = ctypes.c_void_p()
        call("somhip_ctx_create", int(device), ctypes.byref(h))
        self.handle = h
        self.device = int(device)

    def synchronize(self):
        call("somhip_ctx_synchronize", self.handle)

    def set_stream(self, stream_ptr):
        call("somhip_ctx_set_stream", self.handle, None if stream_ptr is None else ctypes.c_void_p(stream_ptr))

    def stream(self):
        s = ctypes.c_void_p()
        call("somhip_ctx_get_stream", self.handle, ctypes.byref(s))
        return s.value

    def set_timing(self, enable=True):
        call("somhip_ctx_set_timing", self.handle, 1 if enable else 0)

    def timing(self):
        """{kernel: (total_ms, launches)} of the kernels timed since set_timing(True)."""
        out = {}
        for name, kid in _lib.TIMERS.items():
            ms = ctypes.c_double()
            cnt = ctypes.c_int64()
            call("somhip_ctx_get_timing", self.handle, kid, ctypes.byref(ms), ctypes.byref(cnt))
            if cnt.value:
                out[name] = (ms.value, cnt.value)
        return out

    def close(self):
        if self.handle:
            _lib.load().somhip_ctx_destroy(self.handle)
            self.handle = None


def get_context(device=0):
    """Process-wide context for ``device`` (created on first use)."""
    with _ctx_lock:
        c = _contexts.get(device)
        if c is None:
            c = Context(device)
            _contexts[device] = c
        return c


def gpu_available():
    try:
        return _lib.device_count() > 0
    except _lib.SomHipError:
        return False


class DeviceArray:
    """Owned device buffer with a numpy shape/dtype (the reference's DeviceArray)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        call("somhip_malloc", ctx.handle, max(self.nbytes, 1), ctypes.byref(p))
        self.ptr = p.value

    @classmethod
    def empty(cls, ctx, shape, dtype):
        return cls(ctx, shape, dtype)

    @classmethod
    def zeros(cls, ctx, shape, dtype):
        a = cls(ctx, shape, dtype)
        a.fill_zero()
        return a

    @classmethod
    def from_host(cls, ctx, array, dtype=None):
        arr = np.ascontiguousarray(array, dtype=dtype)
        a = cls(ctx, arr.shape, arr.dtype)
        a.copy_from_host(arr)
        return a

    def fill_zero(self):
        call("somhip_memset", self.ctx.handle, ctypes.c_void_p(self.ptr), 0, self.nbytes)

    def copy_from_host(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        if arr.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        call("somhip_memcpy_h2d", self.ctx.handle, ctypes.c_void_p(self.ptr), vp(arr), self.nbytes)

    def copy_to_host(self):
        out = np.empty(self.shape, self.dtype)
        call("somhip_memcpy_d2h", self.ctx.handle, vp(out), ctypes.c_void_p(self.ptr), self.nbytes)
        return out

    def view_ptr(self, offset_elems=0):
        return self.ptr + int(offset_elems) * self.dtype.itemsize

    def free(self):
        if self.ptr:
            call("somhip_free", self.ctx.handle, ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                _lib.load().somhip_free(self.ctx.handle, ctypes.c_void_p(self.ptr))
                self.ptr = 0
        except Exception:
            pass


class Uploader:
    """Ring of device slots for the chunk protocol (somhip_uploader_*): ``put`` copies a chunk on a
    copy stream of its own while the kernels of the previous chunk run, and orders the context's
    stream after the copy. The host arrays of the last ``slots`` chunks are kept referenced, so a
    pinned buffer is not released while its copy may still be in flight."""

    def __init__(self, ctx, slots=2):
        h = ctypes.c_void_p()
        call("somhip_uploader_create", ctx.handle, int(slots), ctypes.byref(h))
        self.ctx = ctx
        self.handle = h
        self.slots = int(slots)
        self._held = [None] * self.slots
        self._next = 0

    def put(self, x, rho=None):
        """Device pointers (x_ptr, rho_ptr or None) of the uploaded chunk."""
        xd = ctypes.c_void_p()
        rd = ctypes.c_void_p()
        call("somhip_uploader_put", self.handle, vp(x), x.nbytes, None if rho is None else vp(rho),
             0 if rho is None else rho.nbytes, ctypes.byref(xd), ctypes.byref(rd))
        self._held[self._next] = (x, rho)
        self._next = (self._next + 1) % self.slots
        return xd.value, (None if rho is None else rd.value)

    def close(self):
        if self.handle:
            _lib.load().somhip_uploader_destroy(self.handle)
            self.handle = None
            self._held = [None] * self.slots

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _opt(ptr):
    return None if ptr is None else ctypes.c_void_p(ptr)


class Model:
    """Weights (mn, dim) + topography (mn, 2) resident on the device, FP32 (the default) or FP64
    (``dtype=np.float64``: the float64 entry points, whose search is the exact FP64 scan and which
    ignore the precision / arithmetic codes)."""

    def __init__(self, ctx, m, n, dim, topology, topography, dtype=np.float32):
        self.dtype = np.dtype(dtype).type
        if self.dtype not in (np.float32, np.float64):
            raise ValueError("device models hold float32 or float64 weights")
        self.f64 = self.dtype is np.float64
        topo = np.ascontiguousarray(topography, dtype=self.dtype)
        if topo.shape != (m * n, 2):
            raise ValueError("topography must have shape (m*n, 2)")
        h = ctypes.c_void_p()
        t = _lib.TOPOLOGY_SQUARE if topology == "square" else _lib.TOPOLOGY_HEXAGONAL
        call("somhip_model_create_f64" if self.f64 else "somhip_model_create", ctx.handle, int(m), int(n), int(dim),
             t, vp(topo), ctypes.byref(h))
        self.ctx = ctx
        self.handle = h
        self.m, self.n, self.dim = int(m), int(n), int(dim)
        self.mn = self.m * self.n
        self.topology = topology
        self._sfx = "_f64" if self.f64 else ""

    def set_weights(self, w):
        fn = "somhip_model_set_weights" + self._sfx
        if isinstance(w, DeviceArray):
            call(fn, self.handle, ctypes.c_void_p(w.ptr), 1)
        elif isinstance(w, int):
            call(fn, self.handle, ctypes.c_void_p(w), 1)
        else:
            arr = np.ascontiguousarray(w, dtype=self.dtype).reshape(self.mn, self.dim)
            call(fn, self.handle, vp(arr), 0)

    def get_weights(self):
        out = np.empty((self.mn, self.dim), self.dtype)
        call("somhip_model_get_weights" + self._sfx, self.handle, vp(out))
        return out

    def weights_ptr(self):
        if self.f64:
            raise ValueError("float64 models have no FP32 weights pointer")
        p = ctypes.c_void_p()
        call("somhip_model_weights_ptr", self.handle, ctypes.byref(p))
        return p.value

    # -- inference (device pointers in, device pointers out) ------------------------------------
    def find_bmus(self, x_ptr, n, out_ptr, precision, arithmetic):
        if self.f64:
            call("somhip_find_bmus_f64", self.handle, ctypes.c_void_p(x_ptr), int(n), ctypes.c_void_p(out_ptr))
        else:
            call("somhip_find_bmus", self.handle, ctypes.c_void_p(x_ptr), int(n), precision, arithmetic,
                 ctypes.c_void_p(out_ptr))

    def count_bmus(self, x_ptr, rho_ptr, n, result_ptr, precision, arithmetic):
        if self.f64:
            call("somhip_count_bmus_f64", self.handle, ctypes.c_void_p(x_ptr), _opt(rho_ptr), int(n),
                 ctypes.c_void_p(result_ptr))
        else:
            call("somhip_count_bmus", self.handle, ctypes.c_void_p(x_ptr), _opt(rho_ptr), int(n), precision,
                 arithmetic, ctypes.c_void_p(result_ptr))

    def compute_errors(self, x_ptr, rho_ptr, n, errors_ptr, precision, arithmetic):
        if self.f64:
            call("somhip_compute_errors_f64", self.handle, ctypes.c_void_p(x_ptr), _opt(rho_ptr), int(n),
                 ctypes.c_void_p(errors_ptr))
        else:
            call("somhip_compute_errors", self.handle, ctypes.c_void_p(x_ptr), _opt(rho_ptr), int(n), precision,
                 arithmetic, ctypes.c_void_p(errors_ptr))

    def close(self):
        if self.handle:
            _lib.load().somhip_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Accumulator:
    """Training accumulator (per-BMU sums in epoch mode, numerator/denominator in vector mode)."""

    def __init__(self, model):
        h = ctypes.c_void_p()
        call("somhip_accum_create", model.handle, ctypes.byref(h))
        self.handle = h
        self.model = model

    def reset(self):
        call("somhip_accum_reset", self.handle)

    # sigma0 is passed in the model's type (the reference casts it to dtype: som_batch.py:315)
    def epoch_accumulate(self, x_ptr, rho_ptr, n, precision, arithmetic, errors_ptr=None):
        if self.model.f64:
            call("somhip_batch_epoch_accumulate_f64", self.handle, ctypes.c_void_p(x_ptr), _opt(rho_ptr), int(n),
                 _opt(errors_ptr))
        else:
            call("somhip_batch_epoch_accumulate", self.handle, ctypes.c_void_p(x_ptr), _opt(rho_ptr), int(n),
                 precision, arithmetic, _opt(errors_ptr))

    def epoch_update(self, cur_epoch, n_epochs, sigma0, history_ptr=None):
        if self.model.f64:
            call("somhip_batch_epoch_update_f64", self.handle, int(cur_epoch), int(n_epochs),
                 float(np.float64(sigma0)), _opt(history_ptr))
        else:
            call("somhip_batch_epoch_update", self.handle, int(cur_epoch), int(n_epochs),
                 float(np.float32(sigma0)), _opt(history_ptr))

    def iter_accumulate(self, x_ptr, rho_ptr, n, cur_vector, n_vectors, sigma0, precision, arithmetic):
        if self.model.f64:
            call("somhip_batch_iter_accumulate_f64", self.handle, ctypes.c_void_p(x_ptr), _opt(rho_ptr), int(n),
                 int(cur_vector), int(n_vectors), float(np.float64(sigma0)))
        else:
            call("somhip_batch_iter_accumulate", self.handle, ctypes.c_void_p(x_ptr), _opt(rho_ptr), int(n),
                 int(cur_vector), int(n_vectors), float(np.float32(sigma0)), precision, arithmetic)

    def iter_update(self, history_ptr=None):
        call("somhip_batch_iter_update" + ("_f64" if self.model.f64 else ""), self.handle, _opt(history_ptr))

    def close(self):
        if self.handle:
            _lib.load().somhip_accum_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
