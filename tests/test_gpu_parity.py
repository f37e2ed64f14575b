"""GPU parity: the HIP path (through the C ABI) against the reference's golden vectors and the CPU
oracle on the same seeded inputs.

Bars: BMU indices and hit counts bit-exact against the oracle's reference arithmetic ('exact' and
'bf16_rerank' precisions; 'fp32' is the expanded-form GEMM, whose near-tie choices are reported
against a tolerance instead); QE to 1e-12 relative; TE and activation counts exact; trained
centroids within the reference's own rtol 1e-4 of the oracle and of the golden vectors, and within
1e-6 relative of the FP64 factorised oracle on the same BMUs (one-step check).
"""
import json
import os

import numpy as np
import pytest

import decontamination_amd as dam
from oracle import FP64, REF32, build as build_oracle
from oracle import som_ref as R

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
PRECS = ["exact", "fp32", "bf16_rerank"]
EXACT_PRECS = ["exact", "bf16_rerank"]


@pytest.fixture(scope="module", autouse=True)
def _setup():
    build_oracle()
    assert dam.gpu_available(), "no GPU visible: the GPU tests must run on an MI355X"


def golden_data(n):
    return np.random.default_rng(seed=0).random((n, 4), np.float32)


def golden_som(cls=dam.SOM_Batch, **kw):
    som = cls(4, 4, 4, dtype=np.float32, topology="square", **kw)
    som.init_rand(seed=0)
    return som


# ---------------------------------------------------------------- golden vectors -------------

@pytest.mark.parametrize("precision", PRECS)
def test_golden_winners(precision):
    som = golden_som(precision=precision)
    assert np.array_equal(som.get_winners(golden_data(8), threads_per_blocks=8), np.array(GOLD["winners"]))


@pytest.mark.parametrize("precision", PRECS)
def test_golden_activation_map(precision):
    som = golden_som(precision=precision)
    act = som.get_activation_map(golden_data(25_000), threads_per_blocks=8)
    assert act.dtype == np.float32 and act.shape == (4, 4)
    assert np.array_equal(act, np.array(GOLD["activation_map"], np.float32))


@pytest.mark.parametrize("precision", PRECS)
def test_golden_epoch_centroids(precision):
    som = golden_som(precision=precision)
    som.train(golden_data(1000), n_epochs=1, threads_per_blocks=64)
    assert np.allclose(som.centroids, np.array(GOLD["centroids_epoch"]), rtol=GOLD["centroids_epoch_rtol"])


@pytest.mark.parametrize("precision", PRECS)
def test_golden_iter_centroids(precision):
    som = golden_som(precision=precision)
    som.train(golden_data(1000), n_vectors=1000, threads_per_blocks=64)
    assert np.allclose(som.centroids, np.array(GOLD["centroids_iter"]), rtol=GOLD["centroids_iter_rtol"])


@pytest.mark.parametrize("precision", PRECS)
def test_golden_errors_smoke(precision):
    som = golden_som(precision=precision)
    q, t = som.compute_errors(golden_data(8), threads_per_blocks=8)
    assert q == 0.0 or np.isfinite(q)


# ---------------------------------------------------------------- BMU parity -----------------

CASES = [
    # (m, n, dim, N, topology, seed)
    (1, 1, 4, 50, "square", 1),
    (1, 2, 3, 200, "hexagonal", 2),
    (4, 4, 4, 1000, "square", 3),
    (5, 7, 8, 3001, "hexagonal", 4),
    (10, 10, 8, 5000, "square", 5),
    (8, 12, 16, 4099, "hexagonal", 6),
    (16, 16, 17, 2048, "hexagonal", 7),
    (64, 64, 16, 6000, "hexagonal", 8),
    (32, 32, 16, 9000, "hexagonal", 11),
    (45, 50, 32, 3000, "square", 12),
    (20, 30, 32, 3000, "square", 9),
    (7, 9, 40, 700, "square", 10),
    (12, 10, 48, 2500, "hexagonal", 15),
    (30, 30, 64, 3000, "square", 16),
    # coarse sections of several chunks: 3 at 80x80/D16; 10 four-block chunks at 100x100/D32, so
    # the BF16 search rescans 5 and records the other 5 in its main pass
    (80, 80, 16, 4000, "hexagonal", 13),
    (100, 100, 32, 3000, "hexagonal", 14),
]


def make_case(m, n, dim, N, seed, ties=True, nans=True):
    rng = np.random.default_rng(seed)
    W = rng.standard_normal((m * n, dim)).astype(np.float32)
    X = rng.standard_normal((N, dim)).astype(np.float32)
    if ties and m * n >= 4:
        W[m * n // 2] = W[1]  # duplicated neuron: equal distances, lower index must win
        X[: min(N, 5)] = W[1]
    if nans and N > 10:
        X[3, 0] = np.nan
        X[9] = np.nan
    return W, X


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}d{c[2]}n{c[3]}{c[4][0]}" for c in CASES])
@pytest.mark.parametrize("precision", PRECS)
def test_winners_vs_oracle(case, precision):
    m, n, dim, N, topo, seed = case
    W, X = make_case(m, n, dim, N, seed)
    som = dam.SOM_Batch(m, n, dim, topology=topo, precision=precision)
    som._weights[...] = W
    got = som.get_winners(X)
    exp = R.find_bmus(W, X, REF32)
    assert got.dtype == np.int32 and got.shape == (N,)
    if precision in EXACT_PRECS:
        assert np.array_equal(got, exp)
    else:
        agree = np.mean(got == exp)
        assert agree >= 0.99, agree
        assert np.array_equal(got[:5], exp[:5]) or m * n < 4  # exact ties resolved to the lower index


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}d{c[2]}n{c[3]}{c[4][0]}" for c in CASES])
@pytest.mark.parametrize("precision", EXACT_PRECS)
def test_errors_vs_oracle(case, precision):
    m, n, dim, N, topo, seed = case
    W, X = make_case(m, n, dim, N, seed)
    rho = np.random.default_rng(seed + 100).random(N)
    rho[::13] = 0.0
    som = dam.SOM_Batch(m, n, dim, topology=topo, precision=precision)
    som._weights[...] = W
    q, t = som.compute_errors(X, rho)
    eq, et = R.compute_errors(W, R.topography(m, n, topo), X, rho, topo, REF32)
    if m * n < 2:
        assert np.isnan(q) and np.isnan(t)
        return
    assert np.isclose(q, eq, rtol=1e-12, atol=0), (q, eq)
    assert np.isclose(t, et, rtol=1e-12, atol=0), (t, et)


@pytest.mark.parametrize("case", CASES[:8], ids=[f"{c[0]}x{c[1]}d{c[2]}n{c[3]}{c[4][0]}" for c in CASES[:8]])
@pytest.mark.parametrize("precision", EXACT_PRECS)
def test_activation_vs_oracle(case, precision):
    m, n, dim, N, topo, seed = case
    W, X = make_case(m, n, dim, N, seed)
    som = dam.SOM_Batch(m, n, dim, topology=topo, precision=precision)
    som._weights[...] = W
    got = som.get_activation_map(X)
    exp = R.activation_map(W, X, None, REF32, m, n)
    assert np.array_equal(got, exp.astype(np.float32))
    rho = np.random.default_rng(seed).random(N) * 3.0
    rho[::5] = -1.0
    got = som.get_activation_map(X, rho)
    exp = R.activation_map(W, X, rho, REF32, m, n)
    assert np.allclose(got, exp.astype(np.float32), rtol=1e-6, atol=0)


@pytest.mark.parametrize("shape", [(40, 40, 16), (30, 35, 32), (20, 24, 64)])
def test_bf16_window_overflow_falls_back_exactly(shape):
    """A map whose neurons are mostly exact duplicates puts hundreds of neurons inside every sample's
    window: the per-lane candidate lists overflow and every such sample goes through the exact
    fallback, which must return the lowest-index tie like the oracle."""
    m, n, dim = shape
    rng = np.random.default_rng(17)
    W = np.repeat(rng.standard_normal((4, dim)).astype(np.float32), m * n // 4 + 1, axis=0)[: m * n]
    W[::7] += rng.standard_normal((len(W[::7]), dim)).astype(np.float32)
    X = rng.standard_normal((3000, dim)).astype(np.float32)
    for precision in ("bf16_rerank",):
        som = dam.SOM_Batch(m, n, dim, precision=precision)
        som._weights[...] = W
        assert np.array_equal(som.get_winners(X), R.find_bmus(W, X, REF32))
        q, t = som.compute_errors(X)
        eq, et = R.compute_errors(W, R.topography(m, n), X, None, "hexagonal", REF32)
        assert np.isclose(q, eq, rtol=1e-12, atol=0) and np.isclose(t, et, rtol=1e-12, atol=0)


def test_bf16_rerank_beyond_16bit_slots():
    """A 256x257 map (65,792 neurons) has more neuron slots than the BF16 kernel's 16-bit candidate
    entries address; `bf16_rerank` then runs the exact kernel: the oracle's winners and errors, and
    training equal to `exact` training bit for bit."""
    m, n, dim, N = 256, 257, 8, 2000
    print(f"bf16_rerank beyond 16-bit slots {m}x{n}/D{dim}", flush=True)
    W, X = make_case(m, n, dim, N, 31)
    som = dam.SOM_Batch(m, n, dim, precision="bf16_rerank")
    som._weights[...] = W
    assert np.array_equal(som.get_winners(X), R.find_bmus(W, X, REF32))
    q, t = som.compute_errors(X)
    eq, et = R.compute_errors(W, R.topography(m, n), X, None, "hexagonal", REF32)
    assert np.isclose(q, eq, rtol=1e-12, atol=0) and np.isclose(t, et, rtol=1e-12, atol=0)
    trained = []
    for precision in ("exact", "bf16_rerank"):
        s = dam.SOM_Batch(m, n, dim, precision=precision)
        s._weights[...] = W
        s.train(X, n_epochs=2, use_best_epoch=False)
        trained.append(s)
    assert np.array_equal(trained[0].weights, trained[1].weights)
    assert np.array_equal(trained[0].quantization_errors, trained[1].quantization_errors)


@pytest.mark.parametrize("precision", EXACT_PRECS)
def test_fp64_arithmetic_matches_fp64_oracle(precision):
    W, X = make_case(12, 12, 16, 5000, 21)
    som = dam.SOM_Batch(12, 12, 16, precision=precision, arithmetic="fp64")
    som._weights[...] = W
    assert np.array_equal(som.get_winners(X), R.find_bmus(W, X, FP64))


# ---------------------------------------------------------------- training parity ------------

@pytest.mark.parametrize("precision", PRECS)
@pytest.mark.parametrize("topology", ["square", "hexagonal"])
def test_train_epochs_vs_oracle(precision, topology):
    m, n, dim, N = 6, 7, 5, 4000
    rng = np.random.default_rng(31)
    X = rng.standard_normal((N, dim)).astype(np.float32)
    rho = rng.random(N)
    W0 = rng.standard_normal((m * n, dim)).astype(np.float32)
    som = dam.SOM_Batch(m, n, dim, topology=topology, precision=precision)
    som._weights[...] = W0
    som.train(X, rho, n_epochs=3, use_best_epoch=False)
    Wr, hist, qe, te = R.train(W0, m, n, X, rho, n_epochs=3, topology=topology, use_best_epoch=False, mode=REF32)
    assert som._history.shape == (3, m * n, dim)
    if precision == "fp32":
        # expanded-form FP32 BMUs may differ from the reference formula at near ties
        assert np.allclose(som.weights, Wr, rtol=0, atol=2e-2)
        assert np.allclose(som.quantization_errors, qe, rtol=1e-3)
        return
    assert np.allclose(som.weights, Wr, rtol=1e-4, atol=1e-6)
    assert np.allclose(som.quantization_errors, qe, rtol=1e-5)
    assert np.allclose(som.topographic_errors, te, rtol=1e-9)


@pytest.mark.parametrize("precision", EXACT_PRECS)
def test_one_step_vs_fp64_factorised(precision):
    """Same weights in, one epoch: GPU centroids vs the FP64 factorised oracle on the same BMUs."""
    m, n, dim, N = 16, 16, 16, 20000
    rng = np.random.default_rng(41)
    X = rng.standard_normal((N, dim)).astype(np.float32)
    W0 = rng.standard_normal((m * n, dim)).astype(np.float32) * 0.3
    som = dam.SOM_Batch(m, n, dim, precision=precision, arithmetic="fp64")
    som._weights[...] = W0
    som.train(X, n_epochs=1, use_best_epoch=False)
    sigma = R.sigma_at(max(m, n) / 2.0, 0, 1)
    Wf, c1, _, _ = R.factorised_epoch_fp64(W0, R.topography(m, n), X, None, sigma)
    rel = np.abs(som.weights - Wf) / np.maximum(np.abs(Wf), 1e-3)
    assert rel.max() <= 1e-6, rel.max()


@pytest.mark.parametrize("shape", [(64, 64, 16, "hexagonal"), (100, 100, 32, "hexagonal"), (90, 70, 20, "square"),
                                   (60, 60, 40, "square"), (50, 50, 64, "hexagonal")])
def test_one_step_multi_slice_scatter_vs_fp64(shape):
    """Maps whose per-neuron table needs several LDS slices (the bucketed scatter): one weighted epoch
    with invalid rows and zero / negative / infinite weights, against the FP64 factorised oracle on the
    FP64 oracle's BMUs."""
    m, n, dim, topo = shape
    print(f"multi-slice scatter {m}x{n}/D{dim}", flush=True)
    N = 150_000
    rng = np.random.default_rng(71 + dim)
    X = rng.standard_normal((N, dim)).astype(np.float32)
    X[11::4999] = np.nan
    rho = rng.random(N) * 3.0
    rho[::97] = 0.0
    rho[1::101] = -1.0
    rho[2::103] = np.inf
    W0 = (rng.standard_normal((m * n, dim)) * 0.5).astype(np.float32)
    som = dam.SOM_Batch(m, n, dim, topology=topo, arithmetic="fp64")
    som._weights[...] = W0
    som.train(X, rho, n_epochs=1, use_best_epoch=False)
    sigma = R.sigma_at(max(m, n) / 2.0, 0, 1)
    ok = np.isfinite(X).all(1) & (rho > 0) & np.isfinite(rho)
    Xo, ro = X[ok], rho[ok].astype(np.float32).astype(np.float64)  # the density is rounded to the dtype
    c1, _, _ = R.top2_omp(W0, Xo, FP64)
    Wf, _, _, _ = R.factorised_epoch_fp64(W0, R.topography(m, n, topo), Xo, ro, sigma, bmu=c1, out_dtype=np.float64)
    rel = np.abs(som.weights - Wf) / np.maximum(np.abs(Wf), 1e-3)
    assert rel.max() <= 1e-6, rel.max()


def test_sorted_scatter_hot_neurons():
    """The sorted scatter (maps of 8 or more slices) with neuron runs far longer than one thread's share
    of entries: 70% of the rows sit at three neurons (one run spans hundreds of threads, so many threads
    flush into one row), the rest spread over the map; one weighted epoch against the FP64 oracle, and
    the same epoch from chunks of unequal size."""
    m, n, dim, N = 100, 100, 32, 400_000
    print("sorted scatter hot neurons 100x100/D32", flush=True)
    rng = np.random.default_rng(81)
    W0 = (rng.standard_normal((m * n, dim)) * 0.5).astype(np.float32)
    X = rng.standard_normal((N, dim)).astype(np.float32)
    hot = rng.random(N) < 0.7
    X[hot] = W0[[17, 4242, 9999]][rng.integers(0, 3, hot.sum())] + \
        (rng.standard_normal((hot.sum(), dim)) * 1e-3).astype(np.float32)
    rho = rng.random(N) * 2.0
    rho[::89] = 0.0
    sigma = R.sigma_at(50.0, 0, 1)
    c1, _, _ = R.top2_omp(W0, X, FP64)
    assert np.bincount(c1, minlength=m * n).max() > 0.2 * N
    Wf, _, _, _ = R.factorised_epoch_fp64(W0, R.topography(m, n), X, rho.astype(np.float32).astype(np.float64),
                                          sigma, bmu=c1, out_dtype=np.float64)
    gen = lambda a: (lambda: lambda: (a[i:j] for i, j in ((0, 150_000), (150_000, 390_000), (390_000, N))))  # noqa: E731
    for data, weights in ((X, rho), (gen(X), gen(rho))):
        som = dam.SOM_Batch(m, n, dim, arithmetic="fp64", sigma=50.0)
        som._weights[...] = W0
        som.train(data, weights, n_epochs=1, use_best_epoch=False)
        rel = np.abs(som.weights - Wf) / np.maximum(np.abs(Wf), 1e-3)
        assert rel.max() <= 1e-6, rel.max()


@pytest.mark.parametrize("shape", [(64, 64, 16, "hexagonal"), (32, 40, 32, "square"), (24, 24, 48, "hexagonal")])
def test_trained_map_bf16_equals_exact(shape):
    """On a map with trained structure (near-equal neighbour distances, the hard case for the BF16
    window): BF16 training equals exact training bit for bit, and the BF16 winners / errors on the
    trained map equal the oracle's."""
    from decontamination_amd import synth
    m, n, dim, topo = shape
    N = 20000
    X = synth.catalogue_rows(N, 8, dim - 10, seed=3)
    W0 = synth.uniform_weights(m * n, dim, seed=4)
    soms = {}
    for precision in ("exact", "bf16_rerank"):
        som = dam.SOM_Batch(m, n, dim, topology=topo, precision=precision)
        som._weights[...] = W0
        som.train(X, n_epochs=4, use_best_epoch=False)
        soms[precision] = som
    assert np.array_equal(soms["exact"].weights, soms["bf16_rerank"].weights)
    assert np.array_equal(soms["exact"].quantization_errors, soms["bf16_rerank"].quantization_errors)
    W = soms["exact"].weights
    assert np.array_equal(soms["bf16_rerank"].get_winners(X), R.find_bmus(W, X, REF32))
    q, t = soms["bf16_rerank"].compute_errors(X)
    eq, et = R.compute_errors(W, R.topography(m, n, topo), X, None, topo, REF32)
    assert np.isclose(q, eq, rtol=1e-12, atol=0) and np.isclose(t, et, rtol=1e-12, atol=0)


def test_determinism_and_chunking():
    m, n, dim, N = 10, 10, 8, 30000
    rng = np.random.default_rng(51)
    X = rng.standard_normal((N, dim)).astype(np.float32)
    rho = rng.random(N)
    W0 = rng.standard_normal((m * n, dim)).astype(np.float32)
    res = []
    for _ in range(2):
        som = dam.SOM_Batch(m, n, dim)
        som._weights[...] = W0
        som.train(X, rho, n_epochs=2, use_best_epoch=False)
        res.append(som.weights.copy())
    assert np.array_equal(res[0], res[1])
    gen = lambda a: (lambda: lambda: (a[i:i + 7000] for i in range(0, len(a), 7000)))
    som = dam.SOM_Batch(m, n, dim)
    som._weights[...] = W0
    som.train(gen(X), gen(rho), n_epochs=2, use_best_epoch=False)
    assert np.allclose(som.weights, res[0], rtol=1e-12, atol=1e-12)


def test_early_stop_and_best_epoch():
    m, n, dim, N = 5, 5, 4, 3000
    rng = np.random.default_rng(61)
    X = rng.random((N, dim)).astype(np.float32)
    W0 = rng.random((m * n, dim)).astype(np.float32)
    som = dam.SOM_Batch(m, n, dim)
    som._weights[...] = W0
    som.train(X, n_epochs=6, stop_quantization_error=1e9, stop_topographic_error=1e9, use_best_epoch=False)
    assert np.isfinite(som.quantization_errors[0]) and np.all(np.isnan(som.quantization_errors[1:]))
    Wr, hist, qe, te = R.train(W0, m, n, X, n_epochs=6, stop_quantization_error=1e9,
                               stop_topographic_error=1e9, use_best_epoch=False)
    assert np.allclose(som.weights, Wr, rtol=1e-4, atol=1e-6)
    som2 = dam.SOM_Batch(m, n, dim)
    som2._weights[...] = W0
    som2.train(X, n_epochs=4)
    assert np.array_equal(som2.weights, som2._history[np.nanargmin(som2.quantization_errors)])


def test_empty_and_invalid_inputs():
    som = dam.SOM_Batch(3, 3, 4)
    som.init_rand(0)
    assert som.get_winners(np.zeros((0, 4), np.float32)).shape == (0,)
    assert np.array_equal(som.get_activation_map(np.zeros((0, 4), np.float32)), np.zeros((3, 3), np.float32))
    with pytest.raises(ValueError):
        som.get_winners(np.zeros((5, 3), np.float32))


def test_cpu_configuration_in_full():
    """The reference's CPU-runnable configuration in full: 10x10 square map, D=8, 100,000 samples of a
    seeded 5-Gaussian mixture (means uniform in [-3, 3], sigma 0.5), weights uniform in [-1, 1], 10
    epochs from sigma0=5. With FP64 arithmetic the GPU path equals the FP64 oracle to about an FP32 ulp
    (no FP32 accumulator on either side). With the reference arithmetic its exact sums land closer to
    the FP64 oracle than the reference's own FP32-accumulator restatement does (max |dW| 2e-3 here)."""
    rng = np.random.default_rng(0)
    means = rng.uniform(-3, 3, (5, 8))
    labels = rng.integers(0, 5, 100_000)
    X = (means[labels] + 0.5 * rng.standard_normal((100_000, 8))).astype(np.float32)
    W0 = rng.uniform(-1, 1, (100, 8)).astype(np.float32)
    kw = dict(n_epochs=10, sigma=5.0, topology="square", use_best_epoch=False)
    W64, _, qe64, te64 = R.train(W0, 10, 10, X, mode=FP64, **kw)
    W32, _, qe32, _ = R.train(W0, 10, 10, X, mode=REF32, **kw)
    soms = {}
    for arith in ("fp64", "ref32"):
        som = dam.SOM_Batch(10, 10, 8, topology="square", sigma=5.0, arithmetic=arith)
        som._weights[...] = W0
        som.train(X, n_epochs=10, use_best_epoch=False)
        soms[arith] = som
    g = soms["fp64"]
    assert np.allclose(g.weights, W64, rtol=1e-6, atol=1e-6)
    assert np.allclose(g.quantization_errors, qe64, rtol=1e-7)
    assert np.allclose(g.topographic_errors, te64, rtol=1e-12, atol=0)
    r = soms["ref32"]
    assert np.abs(r.weights - W64).max() < np.abs(W32 - W64).max()
    assert np.allclose(r.quantization_errors, qe32, rtol=1e-3)
    assert np.array_equal(r.get_winners(X), R.find_bmus(r.weights, X, REF32))


@pytest.mark.parametrize("grid", ["closed_form", "custom"])
def test_update_paths_vs_fp64_oracle(grid):
    """One epoch through the C ABI against the oracle's FP64 epoch step on the same topography. A model
    with the reference's closed-form grid takes the separable update (rows then columns); any other
    topography, here random FP32 coordinates, takes the dense H.S kernel."""
    from decontamination_amd import _lib
    from decontamination_amd.runtime import Accumulator, DeviceArray, Model, get_context
    m, n, dim, N = 12, 15, 16, 30_000
    rng = np.random.default_rng(91)
    X = rng.standard_normal((N, dim)).astype(np.float32)
    W0 = rng.standard_normal((m * n, dim)).astype(np.float32) * 0.5
    topo = R.topography(m, n, "hexagonal")
    if grid == "custom":
        topo = (rng.random((m * n, 2)) * 10).astype(np.float32)
    ctx = get_context(0)
    model = Model(ctx, m, n, dim, "hexagonal", topo)
    model.set_weights(W0)
    acc = Accumulator(model)
    xd = DeviceArray.from_host(ctx, X)
    acc.reset()
    acc.epoch_accumulate(xd.ptr, None, N, _lib.PREC_EXACT, _lib.RERANK_FP64)
    acc.epoch_update(0, 3, 7.5)
    got = model.get_weights()
    num, den = R.epoch_accumulate(W0, topo, X, None, 0, 3, 7.5, FP64)
    want = R.divide(num, den, FP64)
    assert np.allclose(got, want, rtol=1e-6, atol=1e-7), np.abs(got - want).max()
