"""GPU: chunked datasets through the uploader ring (somhip_uploader_*), whose copies overlap the
kernels of the previous chunk.

Bars: BMU indices bit-exact against the oracle per chunk, activation counts exact, and chunked
training bit-identical to the same chunks uploaded synchronously one at a time (the overlap must not
change a single bit), across ragged chunk sizes (empty, tiny, growing, shrinking) and a generator
that reuses one host buffer for every chunk.
"""
import numpy as np
import pytest

import decontamination_amd as dam
from decontamination_amd import _lib
from decontamination_amd.runtime import Accumulator, DeviceArray, Model, Uploader, get_context
from oracle import FP64, REF32, build as build_oracle
from oracle import som_ref as R

pytestmark = pytest.mark.gpu

SIZES = [1000, 50_000, 3, 0, 120_000, 70_000, 160_000, 17, 9_000]


@pytest.fixture(scope="module", autouse=True)
def _setup():
    build_oracle()
    assert dam.gpu_available(), "no GPU visible: the GPU tests must run on an MI355X"


def _chunks(seed, dim):
    rng = np.random.default_rng(seed)
    xs = [rng.standard_normal((s, dim)).astype(np.float32) for s in SIZES]
    rs = [rng.random(s) for s in SIZES]
    return xs, rs


@pytest.mark.parametrize("precision", ["exact", "bf16_rerank"])
@pytest.mark.parametrize("slots", [2, 3])
def test_uploader_ring_winners_and_counts(precision, slots):
    m, n, dim = 16, 16, 16
    print(f"uploader ring {precision} slots={slots}", flush=True)
    rng = np.random.default_rng(7)
    W = rng.standard_normal((m * n, dim)).astype(np.float32)
    xs, _ = _chunks(8, dim)
    ctx = get_context(0)
    model = Model(ctx, m, n, dim, "hexagonal", R.topography(m, n, "hexagonal"))
    model.set_weights(W)
    prec = _lib.PRECISIONS[precision]
    up = Uploader(ctx, slots=slots)
    outs = [DeviceArray.empty(ctx, (max(len(x), 1),), np.int32) for x in xs]
    act = DeviceArray.zeros(ctx, (m * n,), np.float64)
    for x, o in zip(xs, outs):
        xp, _ = up.put(x)
        model.find_bmus(xp, len(x), o.ptr, prec, _lib.RERANK_REF32)
        model.count_bmus(xp, None, len(x), act.ptr, prec, _lib.RERANK_REF32)
    ctx.synchronize()
    total = np.zeros(m * n)
    for x, o in zip(xs, outs):
        want = R.find_bmus(W, x, REF32)
        got = o.copy_to_host()[:len(x)]
        assert np.array_equal(got, want)
        total += np.bincount(want, minlength=m * n)
    assert np.array_equal(act.copy_to_host(), total)
    up.close()


def test_reused_host_buffer():
    """A generator that overwrites one host array for every chunk: each chunk must be the one seen."""
    m, n, dim = 12, 12, 8
    rng = np.random.default_rng(21)
    W = rng.standard_normal((m * n, dim)).astype(np.float32)
    xs, _ = _chunks(22, dim)
    buf = np.empty((max(SIZES), dim), np.float32)

    def builder():
        def gen():
            for x in xs:
                buf[:len(x)] = x
                yield buf[:len(x)]
        return gen

    som = dam.SOM_Batch(m, n, dim, topology="square")
    som._weights[...] = W
    act = som.get_activation_map(lambda: builder())
    want = R.activation_map(W, np.concatenate(xs), None, REF32, m, n)
    assert np.array_equal(act, np.asarray(want, np.float32).reshape(m, n))


@pytest.mark.parametrize("precision", ["fp32", "bf16_rerank"])
def test_streamed_training_equals_synchronous(precision):
    m, n, dim, epochs = 20, 20, 16, 3
    print(f"streamed training {precision}", flush=True)
    rng = np.random.default_rng(31)
    W0 = rng.standard_normal((m * n, dim)).astype(np.float32)
    xs, rs = _chunks(32, dim)
    gen = lambda arrs: (lambda: lambda: iter(arrs))

    som = dam.SOM_Batch(m, n, dim, precision=precision)
    som._weights[...] = W0
    som.train(gen(xs), gen(rs), n_epochs=epochs, use_best_epoch=False)

    # the same epochs with one synchronous upload per chunk and no overlap
    ctx = get_context(0)
    model = Model(ctx, m, n, dim, "hexagonal", R.topography(m, n, "hexagonal"))
    model.set_weights(W0)
    acc = Accumulator(model)
    prec = _lib.PRECISIONS[precision]
    hist = DeviceArray.empty(ctx, (m * n, dim), np.float32)
    errors = DeviceArray.zeros(ctx, (epochs, 3), np.float64)
    for e in range(epochs):
        acc.reset()
        for x, r in zip(xs, rs):
            dx = DeviceArray.from_host(ctx, x) if len(x) else None
            dr = DeviceArray.from_host(ctx, r) if len(r) else None
            if len(x):
                acc.epoch_accumulate(dx.ptr, dr.ptr, len(x), prec, _lib.RERANK_REF32,
                                     errors.view_ptr((e - 1) * 3) if e else None)
            ctx.synchronize()
        acc.epoch_update(e, epochs, som.sigma, hist.ptr)
    want = hist.copy_to_host()
    assert np.array_equal(som.weights, want)


def test_streamed_training_matches_oracle():
    """Ragged weighted chunks against the oracle's chunked training in FP64 (FP64 distances and
    accumulators, so the comparison is not limited by the reference's FP32 accumulators over 23k rows):
    equal to about one FP32 ulp."""
    m, n, dim, epochs = 8, 8, 8, 3
    rng = np.random.default_rng(41)
    sizes = [700, 0, 5, 9000, 2300, 11_000, 1]
    xs = [rng.standard_normal((s, dim)).astype(np.float32) for s in sizes]
    rs = [rng.random(s) for s in sizes]
    W0 = rng.standard_normal((m * n, dim)).astype(np.float32)
    gen = lambda arrs: (lambda: lambda: iter(arrs))
    som = dam.SOM_Batch(m, n, dim, arithmetic="fp64")
    som._weights[...] = W0
    som.train(gen(xs), gen(rs), n_epochs=epochs, use_best_epoch=False)
    Wr, _, qe, te = R.train(W0, m, n, gen(xs), gen(rs), n_epochs=epochs, use_best_epoch=False, mode=FP64)
    assert np.allclose(som.weights, Wr, rtol=1e-6, atol=1e-7)
    assert np.allclose(som.quantization_errors, qe, rtol=1e-6)
    assert np.allclose(som.topographic_errors, te, rtol=1e-9, atol=0)


def test_large_chunk_wraps_the_staging_ring_and_pinned_chunks():
    """A 160 MB chunk takes five 32 MB pinned pieces from a ring of four (the ring wraps inside one
    put, so a piece is refilled only after its DMA); chunks in pinned memory skip the staging. Both give
    the BMUs of the resident path, which are pinned against the oracle on a sample of rows."""
    from decontamination_amd.serving import pinned_empty
    m, n, dim, N = 24, 24, 16, 2_600_000
    print("large chunk through the staging ring", flush=True)
    rng = np.random.default_rng(71)
    X = rng.standard_normal((N, dim), dtype=np.float32)
    som = dam.SOM_Batch(m, n, dim)
    som._weights[...] = rng.standard_normal((m * n, dim)).astype(np.float32)
    want = som.get_winners(X)  # one pageable chunk of 166 MB: staged, ring wraps
    ctx = get_context(0)
    xd = DeviceArray.from_host(ctx, X)
    out = DeviceArray.empty(ctx, (N,), np.int32)
    som._device_model().find_bmus(xd.ptr, N, out.ptr, _lib.PREC_BF16_RERANK, _lib.RERANK_REF32)
    assert np.array_equal(want, out.copy_to_host())
    idx = rng.choice(N, 20_000, replace=False)
    assert np.array_equal(want[idx], R.find_bmus(som.weights, X[idx], REF32))
    # the same rows as pinned chunks (copied directly) and as pageable chunks, through a generator
    pinned = [pinned_empty((s, dim), np.float32) for s in (1_000_000, 1_600_000)]
    pinned[0][...] = X[:1_000_000]
    pinned[1][...] = X[1_000_000:]
    act_pinned = som.get_activation_map(lambda: lambda: iter(pinned))
    act_pageable = som.get_activation_map(lambda: lambda: iter([X[:1_000_000], X[1_000_000:]]))
    want_act = np.bincount(want, minlength=m * n).reshape(m, n).astype(np.float32)
    assert np.array_equal(act_pinned, want_act) and np.array_equal(act_pageable, want_act)


def test_two_uploaders_share_the_staging_ring():
    """The pinned staging pieces are process-wide: uploaders that take turns (and one destroyed while
    the other goes on) must each land their own bytes. 48 MB chunks take two pieces per put, so the
    second uploader refills pieces whose last DMA belonged to the first."""
    m, n, dim, N = 16, 16, 16, 750_000
    print("two uploaders on the shared staging ring", flush=True)
    rng = np.random.default_rng(72)
    som = dam.SOM_Batch(m, n, dim)
    som._weights[...] = rng.standard_normal((m * n, dim)).astype(np.float32)
    model = som._device_model()
    ctx = get_context(0)
    chunks = [rng.standard_normal((N, dim), dtype=np.float32) for _ in range(6)]
    want = [som.get_winners(c) for c in chunks]
    a, b = Uploader(ctx, slots=2), Uploader(ctx, slots=2)
    outs = [DeviceArray.empty(ctx, (N,), np.int32) for _ in chunks]
    for k, c in enumerate(chunks):
        if k == 4:
            a.close()  # b goes on alone over pieces a used last
        up = b if k % 2 or a.handle is None else a
        xp, _ = up.put(c)
        model.find_bmus(xp, N, outs[k].ptr, _lib.PREC_BF16_RERANK, _lib.RERANK_REF32)
    b.close()
    for k in range(len(chunks)):
        assert np.array_equal(outs[k].copy_to_host(), want[k]), k
