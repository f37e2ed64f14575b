"""GPU: float64 SOMs (the reference's dtype=np.float64, som_abstract.py:46 and som_batch.py:289-315),
through the C ABI's float64 entry points.

Bars: winners bit-exact against the float64 oracle's (data that FP32 cannot hold, exact ties, NaN rows);
activation maps and errors to FP64 rounding (relative 1e-12); training, epoch and vector-count modes,
on the separable closed-form grids and on an arbitrary topography, to FP64 rounding of the oracle's
per-sample sums (relative 1e-10 of max(|w|, 1e-3)); chunked training equal to resident training to FP64
rounding (the fixed-point scales follow each chunk); float32 and float64 models refused by each other's
entry points.
"""
import ctypes

import numpy as np
import pytest

import decontamination_amd as dam
from decontamination_amd import _lib
from decontamination_amd.runtime import Accumulator, DeviceArray, Model, get_context
from decontamination_amd.serving import SOMServer
from oracle import build as build_oracle
from oracle import som_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _setup():
    build_oracle()
    assert dam.gpu_available(), "no GPU visible: the GPU tests must run on an MI355X"


def _case(m, n, dim, N, seed):
    rng = np.random.default_rng(seed)
    W = rng.standard_normal((m * n, dim)) / 3.0  # values FP32 cannot hold
    X = rng.standard_normal((N, dim)) / 7.0
    W[m * n // 2] = W[1]  # duplicated neuron: the lower index wins
    X[:5] = W[1]
    X[3, 0] = np.nan
    X[9] = np.nan
    return W, X


def _rel(got, want):
    return float(np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-3)))


@pytest.mark.parametrize("shape", [(7, 9, 5, "hexagonal"), (12, 10, 16, "square"), (9, 11, 40, "hexagonal"),
                                   (5, 6, 64, "square")])
def test_float64_inference_vs_oracle(shape):
    m, n, dim, topology = shape
    print(f"float64 inference {m}x{n}/D{dim} {topology}", flush=True)
    W, X = _case(m, n, dim, 3000, 40 + dim)
    som = dam.SOM_Batch(m, n, dim, dtype=np.float64, topology=topology)
    som._weights[...] = W
    assert np.array_equal(som.get_winners(X), R.find_bmus_f64(W, X))
    rho = np.random.default_rng(dim).random(X.shape[0]) * 3.0
    rho[::7] = 0.0
    rho[1::11] = -1.0
    act = som.get_activation_map(X, rho)
    assert act.dtype == np.float64
    assert np.allclose(act.reshape(-1), R.activation_map_f64(W, X, rho), rtol=1e-12, atol=0)
    assert np.array_equal(som.get_activation_map(X).reshape(-1), R.activation_map_f64(W, X))
    topo = R.topography(m, n, topology, np.float64)
    for r in (None, rho):
        q, t = som.compute_errors(X, r)
        eq, et = R.errors_f64(W, topo, X, r, topology)
        assert np.isclose(q, eq, rtol=1e-12, atol=0) and np.isclose(t, et, rtol=1e-12, atol=0), (q, eq, t, et)


def test_float64_hexagonal_threshold_uses_fp64_topography():
    """Grid neighbours (0, 2) and (0, 3) of a hexagonal map lie at 1.0000000000000004 in FP64 (a
    topographic error), at 0.99999988 in the FP32 topography of float32 SOMs (none)."""
    W = np.full((8, 2), 100.0)
    W[2] = [0.0, 0.0]
    W[3] = [0.1, 0.0]
    X = np.array([[0.04, 0.0]])
    s64 = dam.SOM_Batch(2, 4, 2, dtype=np.float64)
    s64._weights[...] = W
    assert s64.compute_errors(X) == (0.04, 1.0)
    s32 = dam.SOM_Batch(2, 4, 2, precision="exact", arithmetic="fp64")
    s32._weights[...] = W
    assert s32.compute_errors(X)[1] == 0.0


@pytest.mark.parametrize("topology", ["square", "hexagonal"])
def test_float64_training_vs_oracle(topology):
    """Three epochs (separable update on the closed-form grid) and the vector-count mode over three chunks,
    against the oracle's per-sample FP64 sums; chunked training equals resident training to FP64 rounding."""
    m, n, dim, N, E = 8, 9, 6, 4000, 3
    print(f"float64 training {topology}", flush=True)
    W0, X = _case(m, n, dim, N, 7)
    rho = np.random.default_rng(8).random(N) * 2.0
    rho[::13] = 0.0
    topo = R.topography(m, n, topology, np.float64)
    sigma = 3.0
    som = dam.SOM_Batch(m, n, dim, dtype=np.float64, topology=topology, sigma=sigma)
    som._weights[...] = W0
    som.train(X, rho, n_epochs=E, use_best_epoch=False)
    W = W0.copy()
    for e in range(E):
        W = R.epoch_f64(W, topo, X, rho, sigma, e, E)
        assert _rel(som._history[e], W) <= 1e-10, (e, _rel(som._history[e], W))
    assert som.weights.dtype == np.float64
    eq, et = R.errors_f64(W, topo, X, rho, topology)
    assert np.isclose(som.quantization_errors[-1], eq, rtol=1e-10)
    assert np.isclose(som.topographic_errors[-1], et, rtol=1e-10)

    chunks = lambda: lambda: (X[i:i + 1500] for i in range(0, N, 1500))  # noqa: E731
    wchunks = lambda: lambda: (rho[i:i + 1500] for i in range(0, N, 1500))  # noqa: E731
    sc = dam.SOM_Batch(m, n, dim, dtype=np.float64, topology=topology, sigma=sigma)
    sc._weights[...] = W0
    sc.train(chunks, wchunks, n_epochs=E, use_best_epoch=False)
    assert np.allclose(sc.weights, som.weights, rtol=1e-12, atol=1e-14)

    sv = dam.SOM_Batch(m, n, dim, dtype=np.float64, topology=topology, sigma=sigma)
    sv._weights[...] = W0
    sv.train(chunks, wchunks, n_vectors=3500)
    want = R.iter_f64(W0, topo, X[:3500], rho[:3500], sigma, 0, 3500)
    assert _rel(sv.weights, want) <= 1e-10, _rel(sv.weights, want)


def test_float64_values_beyond_float32_range():
    """Data, weights and densities above the float32 range (3.4e38) in a float64 SOM: the fixed-point
    scales come from FP64 bounds, so the activation map and training follow the oracle."""
    m, n, dim, N, E, sigma = 6, 7, 5, 3000, 2, 2.0
    print("float64 values beyond the float32 range", flush=True)
    W0, X = _case(m, n, dim, N, 13)
    W0 *= 1e40
    X *= 1e40
    rho = np.random.default_rng(14).random(N) * 1e45
    rho[::11] = 0.0
    topo = R.topography(m, n, "hexagonal", np.float64)
    som = dam.SOM_Batch(m, n, dim, dtype=np.float64, sigma=sigma)
    som._weights[...] = W0
    assert np.array_equal(som.get_winners(X), R.find_bmus_f64(W0, X))
    act = som.get_activation_map(X, rho)
    assert np.allclose(act.reshape(-1), R.activation_map_f64(W0, X, rho), rtol=1e-12, atol=0)
    som.train(X, rho, n_epochs=E, use_best_epoch=False)
    W = W0.copy()
    for e in range(E):
        W = R.epoch_f64(W, topo, X, rho, sigma, e, E)
        assert _rel(som._history[e], W) <= 1e-10, (e, _rel(som._history[e], W))


@pytest.mark.parametrize("shape", [(64, 64, 16, "hexagonal"), (100, 100, 16, "hexagonal"), (100, 100, 32, "square")])
def test_float64_multi_slice_scatter(shape):
    """float64 maps whose per-neuron table needs several LDS slices: the bucketed scatter over FP64 rows at
    64x64/D16 and 100x100/D32 and the sorted one at 100x100/D16. One weighted epoch against the
    factorised FP64 oracle on the winners of the same weights, which are checked against the float64
    oracle on a subsample."""
    m, n, dim, topology = shape
    print(f"float64 multi-slice scatter {m}x{n}/D{dim} {topology}", flush=True)
    N = 30_000
    rng = np.random.default_rng(90 + dim + m)
    W0 = rng.standard_normal((m * n, dim)) / 3.0
    X = rng.standard_normal((N, dim)) / 2.0
    rho = rng.random(N) * 2.0
    rho[::17] = 0.0
    som = dam.SOM_Batch(m, n, dim, dtype=np.float64, topology=topology)
    som._weights[...] = W0
    c1 = som.get_winners(X)
    assert np.array_equal(c1[:600], R.find_bmus_f64(W0, X[:600]))
    som.train(X, rho, n_epochs=1, use_best_epoch=False)
    want = R.factorised_epoch_f64(W0, R.topography(m, n, topology, np.float64), X, rho, max(m, n) / 2.0, 0, 1, c1)
    assert _rel(som.weights, want) <= 1e-10, _rel(som.weights, want)


def test_float64_dense_update_on_arbitrary_topography():
    """A topography that is not the closed-form grid takes the dense H.S update."""
    m, n, dim, N = 7, 8, 5, 3000
    W0, X = _case(m, n, dim, N, 9)
    topo = np.random.default_rng(10).random((m * n, 2)) * 6.0
    ctx = get_context(0)
    model = Model(ctx, m, n, dim, "hexagonal", topo, dtype=np.float64)
    model.set_weights(W0)
    acc = Accumulator(model)
    xd = DeviceArray.from_host(ctx, X)
    acc.reset()
    acc.epoch_accumulate(xd.ptr, None, N, _lib.PREC_EXACT, _lib.RERANK_FP64)
    acc.epoch_update(1, 4, 2.5)
    assert _rel(model.get_weights(), R.epoch_f64(W0, topo, X, None, 2.5, 1, 4)) <= 1e-10


def test_float32_and_float64_models_refuse_each_other():
    ctx = get_context(0)
    L = _lib.load()
    m32 = Model(ctx, 2, 2, 3, "square", R.topography(2, 2, "square"))
    m64 = Model(ctx, 2, 2, 3, "square", R.topography(2, 2, "square", np.float64), dtype=np.float64)
    out = DeviceArray.empty(ctx, (4,), np.int32)
    x32 = DeviceArray.from_host(ctx, np.zeros((4, 3), np.float32))
    x64 = DeviceArray.from_host(ctx, np.zeros((4, 3), np.float64))
    rc = L.somhip_find_bmus_f64(m32.handle, ctypes.c_void_p(x64.ptr), 4, ctypes.c_void_p(out.ptr))
    assert rc == _lib.ERR_ARG and b"float32 model" in L.somhip_last_error()
    rc = L.somhip_find_bmus(m64.handle, ctypes.c_void_p(x32.ptr), 4, _lib.PREC_EXACT, _lib.RERANK_FP64,
                            ctypes.c_void_p(out.ptr))
    assert rc == _lib.ERR_ARG and b"float64 model" in L.somhip_last_error()
    with pytest.raises(NotImplementedError):
        SOMServer([dam.SOM_Batch(2, 2, 3, dtype=np.float64)])


def test_float64_decontamination_pipeline():
    """Decontamination_SOM(dtype=np.float64): PCA initialisation, float64 training, winners, then the
    reference's own gndm sequence (decontamination_som.py:600-667): the two activation maps, gnd and gndm."""
    m, n, dim, N = 6, 7, 5, 20_000
    print("float64 decontamination pipeline", flush=True)
    rng = np.random.default_rng(12)
    X = rng.standard_normal((N, dim)) * [1.0, 2.0, 0.5, 1.0, 3.0]
    density = rng.poisson(3.0, N).astype(np.float64)
    coverage = rng.random(N) * 0.5 + 0.5
    d = dam.Decontamination_SOM(256, np.arange(N), coverage, X, density, m, n, dim, True, dtype=np.float64)
    d.train(n_epochs=2, show_progress_bar=False)
    assert d.weights.dtype == np.float64
    assert np.array_equal(d.winners, R.find_bmus_f64(d.weights, X))
    d.compute_gndm(show_progress_bar=False)
    corrected = density / coverage
    cat = R.activation_map_f64(d.weights, X, corrected).reshape(m, n)
    foot = R.activation_map_f64(d.weights, X).reshape(m, n)
    assert d.catalog_activation_map.dtype == np.float64
    assert np.allclose(d.catalog_activation_map, cat, rtol=1e-12, atol=0)
    assert np.array_equal(d.footprint_activation_map, foot)
    gnd = np.divide(cat, foot, out=np.full(cat.shape, np.nan), where=foot > 0.0)
    assert np.allclose(d.gnd, gnd, rtol=1e-12, atol=0, equal_nan=True)
    assert np.allclose(d.gndm, gnd.reshape(-1)[d.winners], rtol=1e-12, atol=0, equal_nan=True)
