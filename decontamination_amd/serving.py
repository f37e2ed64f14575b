"""Serving: several trained SOMs resident on the GPU, BMU and activation-map requests of uneven sizes
continuously batched into shared launches (libsomhip's server; see include/somhip.h).

    server = SOMServer([som_a, som_b], memory_budget=64 << 30)
    t = server.submit(0, 'winners', x)            # x: (N, dim) float32, ideally from pinned_empty()
    winners, latency_ms = server.wait(t)
    t = server.submit(1, 'activation', x, rho)
    act_map, latency_ms = server.wait(t)          # (m, n) float32, like get_activation_map

Results equal ``get_winners`` / ``get_activation_map`` of the same SOM with precision 'bf16_rerank'
bit for bit, whatever else is in flight.
"""
import ctypes
import threading
import weakref

import numpy as np

from . import _lib
from ._lib import call
from .runtime import Model, get_context

KINDS = {"winners": _lib.SERVE_WINNERS, "activation": _lib.SERVE_ACTIVATION}


def _free_pinned(ptr):
    try:
        _lib.load().somhip_host_free(ctypes.c_void_p(ptr))
    except Exception:
        pass


def pinned_empty(shape, dtype):
    """numpy array in pinned (page-locked) host memory, freed when the array is garbage collected."""
    dtype = np.dtype(dtype)
    shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
    nbytes = max(int(np.prod(shape, dtype=np.int64)) * dtype.itemsize, 1)
    p = ctypes.c_void_p()
    call("somhip_host_alloc", nbytes, ctypes.byref(p))
    buf = (ctypes.c_byte * nbytes).from_address(p.value)
    arr = np.frombuffer(buf, dtype=np.uint8, count=nbytes)[: int(np.prod(shape, dtype=np.int64)) * dtype.itemsize]
    arr = arr.view(dtype).reshape(shape)
    weakref.finalize(buf, _free_pinned, p.value)
    return arr


class SOMServer:
    """Resident SOMs + continuous-batching request server on one GPU."""

    def __init__(self, soms, memory_budget=64 << 30, arithmetic="ref32", device=0):
        if arithmetic not in _lib.ARITHMETICS:
            raise ValueError(f"invalid arithmetic `{arithmetic}`")
        soms = list(soms)
        if any(som.dtype is not np.float32 for som in soms):
            raise NotImplementedError("serving supports float32 SOMs")
        self.ctx = get_context(device)
        h = ctypes.c_void_p()
        call("somhip_server_create", self.ctx.handle, int(memory_budget), _lib.ARITHMETICS[arithmetic],
             ctypes.byref(h))
        self.handle = h
        self.soms = list(soms)
        self.models = []
        for som in self.soms:
            mdl = Model(self.ctx, som.m, som.n, som.dim, som.topology, som._topography)
            mdl.set_weights(som.weights)
            mid = ctypes.c_int()
            call("somhip_server_add_model", self.handle, mdl.handle, ctypes.byref(mid))
            self.models.append(mdl)
        self._refs = {}
        self._lock = threading.Lock()

    def submit(self, model_index, kind, x, rho=None, out=None):
        """Queue a request; returns a ticket. ``x``/``rho``/``out`` must not be modified until done."""
        som = self.soms[model_index]
        if kind not in KINDS:
            raise ValueError(f"invalid request kind `{kind}`")
        x = np.asarray(x)
        if x.dtype != np.float32 or not x.flags.c_contiguous:
            x = np.ascontiguousarray(x, dtype=np.float32)
        if x.ndim != 2 or x.shape[1] != som.dim:
            raise ValueError(f"requests for model {model_index} must have shape (N, {som.dim})")
        n = x.shape[0]
        if rho is not None:
            rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
            if rho.shape[0] != n:
                raise ValueError("`dataset` and `dataset_weights` must be aligned")
        if out is None:
            out = np.empty(n, np.int32) if kind == "winners" else np.empty(som.m * som.n, np.float64)
        t = ctypes.c_int64()
        call("somhip_server_submit", self.handle, int(model_index), KINDS[kind], _lib.vp(x), _lib.vp(rho), int(n),
             _lib.vp(out), ctypes.byref(t))
        with self._lock:
            self._refs[t.value] = (x, rho, out, kind, model_index)
        return t.value

    def poll(self, ticket):
        return self.wait(ticket, timeout=0.0, _poll=True)

    def wait(self, ticket, timeout=None, _poll=False):
        """Block until the request is done; returns (result, latency_ms). Winners come back as int32 (N,),
        activation maps as (m, n) float32 like ``get_activation_map``. With timeout, returns None on timeout."""
        done = ctypes.c_int(0)
        lat = ctypes.c_double(0.0)
        step = 3600.0 if timeout is None else float(timeout)
        while True:
            call("somhip_server_wait", self.handle, int(ticket), step, ctypes.byref(done), ctypes.byref(lat))
            if done.value or timeout is not None or _poll:
                break
        if not done.value:
            return None
        with self._lock:
            x, rho, out, kind, mi = self._refs.pop(ticket)
        if kind == "activation":
            som = self.soms[mi]
            return out.astype(som.dtype).reshape(som.m, som.n), lat.value
        return out, lat.value

    def stats(self):
        c, l, t, f = (ctypes.c_int64() for _ in range(4))
        mem = ctypes.c_size_t()
        call("somhip_server_stats", self.handle, ctypes.byref(c), ctypes.byref(l), ctypes.byref(t), ctypes.byref(f),
             ctypes.byref(mem))
        return {"completed": c.value, "launches": l.value, "tiles": t.value, "in_flight": f.value,
                "memory_in_use": mem.value}

    def close(self):
        if self.handle:
            _lib.load().somhip_server_destroy(self.handle)
            self.handle = None
            for m in self.models:
                m.close()
            self.models = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
