"""GPU: somhip_find_bmus_noalloc (the allocation-free winners entry point) through the C ABI gives the same
winners as somhip_find_bmus, for every precision and rerank arithmetic, on two calls in a row (the second
with fewer rows, so a fallback list left by the first call must not leak into it), and rejects inputs above
its row limit."""
import ctypes

import numpy as np
import pytest

import decontamination_amd as dam
from decontamination_amd import _lib
from decontamination_amd.runtime import DeviceArray, Model, get_context

pytestmark = pytest.mark.gpu

CASES = [  # (m, n, dim, topology)
    (64, 64, 16, "hexagonal"),
    (10, 12, 32, "square"),
]
MODES = [("exact", "ref32"), ("fp32", "ref32"), ("bf16_rerank", "ref32"), ("bf16_rerank", "fp64")]


def _data(rng, rows, dim):
    x = rng.standard_normal((rows, dim)).astype(np.float32)
    x[5::997] = np.nan  # rows without a finite distance: the BF16 path sends them to its fallback list
    x[7::1009, 0] = np.inf
    return x


def _winners(model, name, xd, n, prec, arith):
    out = DeviceArray.empty(model.ctx, (max(n, 1),), np.int32)
    _lib.call(name, model.handle, ctypes.c_void_p(xd.ptr), int(n), _lib.PRECISIONS[prec], _lib.ARITHMETICS[arith],
              ctypes.c_void_p(out.ptr))
    got = out.copy_to_host()[:n]
    out.free()
    return got


@pytest.mark.parametrize("m,n,dim,topology", CASES)
def test_noalloc_equals_find_bmus_twice(m, n, dim, topology):
    assert dam.gpu_available(), "no GPU visible: the GPU tests must run on an MI355X"
    ctx = get_context(0)
    rng = np.random.default_rng(m * 100 + dim)
    som = dam.SOM_Batch(m, n, dim, topology=topology)
    model = Model(ctx, m, n, dim, topology, som._topography)
    model.set_weights(rng.standard_normal((m * n, dim)).astype(np.float32))
    for prec, arith in MODES:
        for rows in (300_001, 1_003):  # the second call is smaller
            print(f"noalloc {m}x{n}/D{dim} {prec}/{arith} rows={rows}", flush=True)
            x = _data(rng, rows, dim)
            xd = DeviceArray.from_host(ctx, x)
            want = _winners(model, "somhip_find_bmus", xd, rows, prec, arith)
            got = _winners(model, "somhip_find_bmus_noalloc", xd, rows, prec, arith)
            xd.free()
            assert np.array_equal(got, want), (prec, arith, rows, int(np.sum(got != want)))
            assert np.all(got[5::997] == 0)  # no finite distance: default index 0
    model.close()


def test_noalloc_row_limit_and_empty_input():
    ctx = get_context(0)
    model = Model(ctx, 4, 4, 16, "square", dam.SOM_Batch(4, 4, 16, topology="square")._topography)
    model.set_weights(np.ones((16, 16), np.float32))
    rows = _lib.NOALLOC_MAX_ROWS + 1
    xd = DeviceArray.zeros(ctx, (rows, 16), np.float32)  # real buffers of that size, though the call rejects it
    out = DeviceArray.empty(ctx, (rows,), np.int32)
    with pytest.raises(ValueError, match="at most"):
        _lib.call("somhip_find_bmus_noalloc", model.handle, ctypes.c_void_p(xd.ptr), rows, _lib.PREC_BF16_RERANK,
                  _lib.RERANK_REF32, ctypes.c_void_p(out.ptr))
    _lib.call("somhip_find_bmus_noalloc", model.handle, None, 0, _lib.PREC_BF16_RERANK, _lib.RERANK_REF32, None)
    xd.free()
    out.free()
    model.close()
