"""GPU: the steps on either side of the SOM path, through the C ABI.

Bars: the device covariance equals the oracle's sequential Welford (covariance.py:30-119) to FP64
rounding (relative 1e-10 of the largest entry) across dims 2..64, weights, invalid rows, chunking
and FP64 input, and reproduces the reference's golden SOM_PCA covariance and centroids; the
galaxy-number-density maps, gnd and gndm are bit-identical to the oracle's activation maps and
_compute_gndm on the same winners; the Decontamination_SOM pipeline (PCA init, training, winners,
gndm) agrees with the oracle chain at the reference's tolerances.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import decontamination_amd as dam
from decontamination_amd import _lib
from decontamination_amd.decontamination_som import compute_gndm
from decontamination_amd.runtime import DeviceArray, get_context
from h5stub import fake_h5py  # noqa: F401  (pytest fixture)
from oracle import REF32, build as build_oracle
from oracle import som_ref as R

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


@pytest.fixture(scope="module", autouse=True)
def _setup():
    build_oracle()
    assert dam.gpu_available(), "no GPU visible: the GPU tests must run on an MI355X"


def _close(got, want, rel=1e-10):
    scale = np.abs(want).max()
    return np.allclose(got, want, rtol=rel, atol=rel * scale)


# ---------------------------------------------------------------- covariance -----------------

@pytest.mark.parametrize("dim", [2, 4, 5, 16, 17, 32, 33, 48, 64])
def test_covariance_vs_oracle(dim):
    print(f"covariance dim={dim}", flush=True)
    rng = np.random.default_rng(100 + dim)
    N = 40_000
    scale = rng.random(dim) * 5 + 0.1
    offset = rng.standard_normal(dim) * 100
    X = (rng.standard_normal((N, dim)) * scale + offset).astype(np.float32)
    w = rng.random(N) * 2
    w[::9] = 0.0
    w[1::31] = -2.0
    w[2::37] = np.nan
    w[3::41] = np.inf
    X[4::43, dim // 2] = np.nan
    X[5::47, 0] = -np.inf
    for weights in (None, w):
        want = R.covariance(dim, X, weights)
        got = dam.Covariance.compute(dim, X, weights)
        assert got.shape == (dim, dim) and np.array_equal(got, got.T)
        assert _close(got, want), (dim, weights is None, np.abs(got - want).max(), np.abs(want).max())
    # chunked (ragged, with an empty chunk) and FP64 input
    cuts = [0, 3, 3, 17_000, 17_001, N]
    gx = lambda: lambda: (X[a:b].astype(np.float64) for a, b in zip(cuts[:-1], cuts[1:]))
    gw = lambda: lambda: (w[a:b] for a, b in zip(cuts[:-1], cuts[1:]))
    assert _close(dam.Covariance.compute(dim, gx, gw), R.covariance(dim, X, w))


def test_covariance_reference_tests_and_errors():
    data = np.random.default_rng(seed=0).random((25_000, 4), np.float32)
    for ddof in (1, 0):  # test/test_covariance.py:19-41
        assert np.allclose(dam.Covariance.compute(4, data, ddof=ddof), np.cov(data.T, ddof=ddof), rtol=1e-12, atol=0)
    with pytest.raises(ValueError):
        dam.Covariance.compute(4, data[:1])
    with pytest.raises(ValueError):
        dam.Covariance.compute(4, np.full((10, 4), np.nan, np.float32))
    with pytest.raises(ValueError):
        dam.Covariance.compute(4, data, ddof=-1)
    with pytest.raises(ValueError):
        dam.Covariance.compute(3, data)


def test_som_pca_golden():
    """test/test_som_pca.py:26-104 on the GPU covariance."""
    data = np.random.default_rng(seed=0).random((25_000, 4), np.float32)
    som = dam.SOM_PCA(4, 4, 4, dtype=np.float32, topology="square")
    som.train(data, min_weight=-1.0, max_weight=+1.0, scale_by_variance=False)
    assert np.allclose(som.cov_matrix, GOLD["pca_cov_matrix"])
    assert np.allclose(som.centroids, GOLD["pca_centroids_linear"])
    som.train(data, min_weight=-1.0, max_weight=+1.0, scale_by_variance=True, apply_cdf=True, cdf_gain=1.0)
    assert np.allclose(som.cov_matrix, GOLD["pca_cov_matrix"])
    assert np.allclose(som.centroids, GOLD["pca_centroids_cdf"])


# ---------------------------------------------------------------- galaxy number density ------

def _footprint(seed, N, dim):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, dim)).astype(np.float32)
    rho = rng.random(N) * 3
    rho[::13] = 0.0
    rho[1::17] = -1.0
    rho[2::19] = np.nan
    return X, rho


@pytest.mark.parametrize("m,n,dim,N", [(7, 9, 5, 30_000), (32, 32, 16, 200_000)])
def test_gndm_vs_oracle(m, n, dim, N):
    print(f"gndm {m}x{n} dim={dim} N={N}", flush=True)
    X, rho = _footprint(7 + m, N, dim)
    som = dam.SOM_Batch(m, n, dim)
    som._weights[...] = np.random.default_rng(8).standard_normal((m * n, dim)).astype(np.float32)
    W = som.weights
    winners = som.get_winners(X)
    assert np.array_equal(winners, R.find_bmus(W, X, REF32))
    cat, foot, n_gal, n_pix, gnd, gndm = compute_gndm(get_context(0), winners, rho, m, n)
    cat_o = R.activation_map(W, X, rho, REF32, m, n).astype(np.float32)
    foot_o = R.activation_map(W, X, None, REF32, m, n).astype(np.float32)
    assert np.array_equal(cat, cat_o) and np.array_equal(foot, foot_o)
    # the reference's own maps: get_activation_map gives the same arrays
    assert np.array_equal(cat, som.get_activation_map(X, rho)) and np.array_equal(foot, som.get_activation_map(X))
    n_gal_o, n_pix_o, gnd_o, gndm_o = R.gndm(cat_o, foot_o, winners, m, n)
    assert n_gal == n_gal_o and n_pix == n_pix_o
    assert np.array_equal(gnd, gnd_o, equal_nan=True) and np.array_equal(gndm, gndm_o, equal_nan=True)


def test_gndm_edges_through_abi():
    """Winners outside [0, mn) give NaN; an empty footprint gives zero maps and an all-NaN gnd."""
    ctx = get_context(0)
    mn = 6
    w = np.array([0, 5, 6, -1, 2, 2, 2147483647, 3, 1], np.int32)
    rho = np.array([1.0, 2.0, 5.0, 5.0, 0.5, 0.25, 9.0, 0.0, 4.0])
    d_w, d_r = DeviceArray.from_host(ctx, w), DeviceArray.from_host(ctx, rho)
    maps = DeviceArray.empty(ctx, (3, mn), np.float32)
    out = DeviceArray.empty(ctx, (len(w),), np.float32)
    _lib.call("somhip_gndm", ctx.handle, ctypes.c_void_p(d_w.ptr), ctypes.c_void_p(d_r.ptr), len(w), mn,
              ctypes.c_void_p(maps.view_ptr(0)), ctypes.c_void_p(maps.view_ptr(mn)),
              ctypes.c_void_p(maps.view_ptr(2 * mn)), ctypes.c_void_p(out.ptr))
    cat, foot, gnd = maps.copy_to_host()
    assert np.array_equal(cat, [1.0, 4.0, 0.75, 0.0, 0.0, 2.0])
    assert np.array_equal(foot, [1.0, 1.0, 2.0, 1.0, 0.0, 1.0])
    assert np.array_equal(gnd, np.array([1.0, 4.0, 0.375, 0.0, np.nan, 2.0], np.float32), equal_nan=True)
    want = np.array([1.0, 2.0, np.nan, np.nan, 0.375, 0.375, np.nan, 0.0, 4.0], np.float32)
    assert np.array_equal(out.copy_to_host(), want, equal_nan=True)
    cat, foot, n_gal, n_pix, gnd, gndm = compute_gndm(ctx, np.zeros(0, np.int32), np.zeros(0), 2, 3)
    assert not cat.any() and not foot.any() and np.all(np.isnan(gnd)) and gndm.shape == (0,)


# ---------------------------------------------------------------- pipeline --------------------

def test_decontamination_pipeline_vs_oracle():
    m, n, dim, N = 6, 6, 5, 20_000
    print("decontamination pipeline", flush=True)
    rng = np.random.default_rng(9)
    X = (rng.standard_normal((N, dim)) * [1, 2, 0.5, 1, 3]).astype(np.float32)
    density = rng.poisson(3.0, N).astype(np.float64)
    coverage = rng.random(N) * 0.5 + 0.5
    d = dam.Decontamination_SOM(256, np.arange(N), coverage, X, density, m, n, dim, True)
    d.train(n_epochs=3, show_progress_bar=False)
    corrected = density / coverage
    cov_o = R.covariance(dim, X, corrected)
    assert _close(d.pca.cov_matrix, cov_o.astype(np.float32), rel=1e-6)
    _, _, Wp = R.pca_weights(cov_o, 0.0, 1.0, m, n, dim)
    assert np.allclose(d.pca.weights, Wp.reshape(m * n, dim).astype(np.float32), rtol=1e-6, atol=1e-7)
    Wr, _, qe, _ = R.train(d.pca.weights, m, n, X, corrected, n_epochs=3, use_best_epoch=True)
    assert np.allclose(d.weights, Wr, rtol=1e-4, atol=1e-6)
    assert np.array_equal(d.winners, R.find_bmus(d.weights, X, REF32))
    d.compute_gndm(show_progress_bar=False)
    cat_o = R.activation_map(d.weights, X, corrected, REF32, m, n).astype(np.float32)
    foot_o = R.activation_map(d.weights, X, None, REF32, m, n).astype(np.float32)
    assert np.array_equal(d.catalog_activation_map, cat_o) and np.array_equal(d.footprint_activation_map, foot_o)
    n_gal, n_pix, gnd, gndm = R.gndm(cat_o, foot_o, d.winners, m, n)
    assert np.array_equal(d.gnd, gnd, equal_nan=True) and np.array_equal(d.gndm, gndm, equal_nan=True)
    assert d.n_gal == n_gal and d.n_pix == n_pix
    assert np.allclose(d.visibility, (1.0 + (gndm - n_gal / n_pix) / (n_gal / n_pix)) * coverage, equal_nan=True)


def test_decontamination_save_load_round_trip(fake_h5py):
    """Decontamination_SOM.save/load (decontamination_som.py:451-537) through the in-memory h5py stand-in:
    a model trained on the GPU, saved, loaded into a fresh object, re-infers bit-identical winners and
    galaxy-number-density products."""
    m, n, dim, N = 8, 8, 6, 30_000
    print("decontamination save/load round trip", flush=True)
    rng = np.random.default_rng(21)
    X = (rng.standard_normal((N, dim)) * [1, 2, 0.5, 1, 3, 1.5]).astype(np.float32)
    density = rng.poisson(3.0, N).astype(np.float64)
    coverage = rng.random(N) * 0.5 + 0.5
    d = dam.Decontamination_SOM(512, np.arange(N), coverage, X, density, m, n, dim, True)
    d.train(n_epochs=3, show_progress_bar=False)
    d.compute_gndm(show_progress_bar=False)
    d.save("dec.hdf5")
    assert set(fake_h5py) == {"dec.hdf5", "dec_pca.hdf5"}
    assert set(fake_h5py["dec.hdf5"]["dataset"].data) == {"nside", "footprint", "coverage", "footprint_systematics",
                                                          "galaxy_number_density", "winner"}
    e = dam.Decontamination_SOM(1, np.zeros(1), np.ones(1), np.zeros((1, dim), np.float32), np.zeros(1), m, n, dim,
                                True)
    e.load("dec.hdf5")
    assert e.nside == 512 and np.array_equal(e.footprint, np.arange(N)) and np.array_equal(e.coverage, coverage)
    assert np.array_equal(e.weights, d.weights) and np.array_equal(e.pca.weights, d.pca.weights)
    assert np.array_equal(e.winners, d.winners)
    assert np.array_equal(e.som.get_winners(e.footprint_systematics), d.winners)
    e.compute_gndm(show_progress_bar=False)
    assert np.array_equal(e.catalog_activation_map, d.catalog_activation_map)
    assert np.array_equal(e.footprint_activation_map, d.footprint_activation_map)
    assert np.array_equal(e.gnd, d.gnd, equal_nan=True) and np.array_equal(e.gndm, d.gndm, equal_nan=True)
    assert e.n_gal == d.n_gal and e.n_pix == d.n_pix
