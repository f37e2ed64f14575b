"""CPU: the oracle and host logic of the steps on either side of the SOM path.

* Covariance (Welford restatement, covariance.py:30-119): against numpy's np.cov (unweighted and
  integer frequency weights, where np.cov's normalisation is the reference's sum(w) - ddof), the
  two-pass closed form for real weights, the row rule (non-finite or non-positive weights and
  non-finite values skipped) and chunking.
* SOM_PCA grid (som_pca.py:116-177): the reference's golden covariance and centroids
  (test/test_som_pca.py), closed-form structure of the grid, and the product's host
  ``train_from_cov_matrix`` equal to the oracle.
* Galaxy number density (decontamination_som.py:600-628): hand-computed example, and the product's
  host ``_compute_gndm`` equal to the oracle.
"""
import json
import os

import numpy as np
import pytest

import decontamination_amd as dam
from decontamination_amd.decontamination_som import ds_divide
from oracle import build as build_oracle
from oracle import som_ref as R

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    build_oracle()


def pca_data():
    return np.random.default_rng(seed=0).random((25_000, 4), np.float32)


# ---------------------------------------------------------------- covariance oracle ----------

def test_welford_matches_numpy_cov():
    X = pca_data()
    # test/test_covariance.py:19-41: np.cov with ddof 1 and 0
    for ddof in (1, 0):
        assert np.allclose(R.covariance(4, X, ddof=ddof), np.cov(X.T, ddof=ddof), rtol=1e-12, atol=0)
    rng = np.random.default_rng(3)
    f = rng.integers(1, 5, X.shape[0])
    assert np.allclose(R.covariance(4, X, f.astype(np.float64)), np.cov(X.T, fweights=f, ddof=1), rtol=1e-12, atol=0)


def test_welford_real_weights_and_row_rule():
    rng = np.random.default_rng(4)
    X = (rng.standard_normal((5000, 6)) * [1, 2, 3, 0.5, 10, 1] + [100, -3, 0, 7, 1e3, 2]).astype(np.float64)
    w = rng.random(5000) * 3
    w[::7] = 0.0
    w[1::11] = -1.0
    w[2::13] = np.nan
    w[3::17] = np.inf
    X[5::19, 2] = np.nan
    X[6::23, 4] = np.inf
    ok = np.isfinite(w) & (w > 0) & np.all(np.isfinite(X), axis=1)
    Xv, wv = X[ok], w[ok]
    mu = (wv[:, None] * Xv).sum(0) / wv.sum()
    d = Xv - mu
    want = (wv[:, None] * d).T @ d / (wv.sum() - 1)
    got = R.covariance(6, X, w)
    assert np.allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
    # chunked through generator builders, including an empty chunk
    cuts = [0, 1, 1000, 1000, 3333, 5000]
    gx = lambda: lambda: (X[a:b] for a, b in zip(cuts[:-1], cuts[1:]))
    gw = lambda: lambda: (w[a:b] for a, b in zip(cuts[:-1], cuts[1:]))
    assert np.allclose(R.covariance(6, gx, gw), got, rtol=1e-12, atol=1e-13 * np.abs(want).max())


def test_welford_errors():
    with pytest.raises(ValueError):
        R.covariance(3, np.zeros((1, 3)))  # one row: sum(w) - ddof = 0
    with pytest.raises(ValueError):
        R.covariance(3, np.full((4, 3), np.nan))
    with pytest.raises(ValueError):
        R.covariance(3, np.zeros((4, 3)), ddof=-1)


# ---------------------------------------------------------------- SOM_PCA ---------------------

def test_pca_oracle_golden():
    c = R.covariance(4, pca_data())
    assert np.allclose(c.astype(np.float32), GOLD["pca_cov_matrix"])
    _, _, W = R.pca_weights(c, -1.0, 1.0, 4, 4, 4)
    assert np.allclose(W.astype(np.float32), GOLD["pca_centroids_linear"])
    _, _, W = R.pca_weights(c, -1.0, 1.0, 4, 4, 4, scale_by_variance=True, apply_cdf=True, cdf_gain=1.0)
    assert np.allclose(W.astype(np.float32), GOLD["pca_centroids_cdf"])


def test_pca_grid_structure():
    """Closed forms: the linear grid is v0 x_i + v1 y_j on the two leading eigenvectors; with the
    CDF every weight lies strictly inside (min, max) and the grid is symmetric about its centre."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((5, 5))
    cov = A @ A.T + np.diag([5.0, 4.0, 3.0, 2.0, 1.0])
    ev, evec, W = R.pca_weights(cov, 0.0, 1.0, 3, 4, 5)
    assert np.all(np.diff(ev) <= 0)
    assert np.allclose(cov @ evec, evec * ev, atol=1e-10)
    lx, ly = np.linspace(0, 1, 3), np.linspace(0, 1, 4)
    assert np.allclose(W, evec[:, 0] * lx[:, None, None] + evec[:, 1] * ly[None, :, None], rtol=0, atol=1e-15)
    _, _, Wc = R.pca_weights(cov, -2.0, 3.0, 4, 4, 5, apply_cdf=True, cdf_gain=2.0)
    assert np.all(Wc > -2.0) and np.all(Wc < 3.0)
    assert np.allclose(Wc + Wc[::-1, ::-1], 1.0, atol=1e-12)  # erf odd, grid symmetric about 0


def test_som_pca_host_equals_oracle():
    cov = R.covariance(4, pca_data())
    for kw in ({}, {"scale_by_variance": True}, {"scale_by_variance": True, "apply_cdf": True, "cdf_gain": 1.0},
               {"apply_cdf": True, "cdf_gain": 0.5}):
        som = dam.SOM_PCA(4, 4, 4, topology="square")
        som.train_from_cov_matrix(cov, min_weight=-1.0, max_weight=1.0, **kw)
        ev, evec, W = R.pca_weights(cov, -1.0, 1.0, 4, 4, 4, **kw)
        assert np.array_equal(som.weights, W.astype(np.float32).reshape(16, 4))
        assert np.array_equal(som.eigenvalues, ev.astype(np.float32))
        assert np.array_equal(som.cov_matrix, cov.astype(np.float32))
    with pytest.raises(ValueError):
        dam.SOM_PCA(4, 4, 1)
    with pytest.raises(ValueError):
        dam.SOM_PCA(4, 4, 4).train_from_cov_matrix(cov, min_weight=1.0, max_weight=1.0)
    with pytest.raises(ValueError):
        dam.SOM_PCA(4, 4, 4).train_from_cov_matrix(cov, apply_cdf=True, cdf_gain=0.0)


# ---------------------------------------------------------------- galaxy number density -------

def test_gndm_oracle_hand_example():
    cat = np.array([[2.0, 0.0], [3.0, 1.5]], np.float32)
    foot = np.array([[4.0, 0.0], [2.0, 3.0]], np.float32)
    winners = np.array([0, 3, 3, 2, 0, 1], np.int32)
    n_gal, n_pix, gnd, gndm = R.gndm(cat, foot, winners, 2, 2)
    assert n_gal == np.float32(6.5) and n_pix == np.float32(9.0)
    assert gnd[0, 0] == np.float32(0.5) and np.isnan(gnd[0, 1]) and gnd[1, 0] == np.float32(1.5)
    assert gnd[1, 1] == np.float32(0.5) and gnd.dtype == np.float32
    assert np.array_equal(gndm, np.array([0.5, 0.5, 0.5, 1.5, 0.5, np.nan], np.float32), equal_nan=True)


def test_decontamination_host_compute_gndm_equals_oracle():
    rng = np.random.default_rng(6)
    cat = rng.random((5, 6)).astype(np.float32) * 10
    foot = rng.integers(0, 4, (5, 6)).astype(np.float32)
    winners = rng.integers(0, 30, 100).astype(np.int32)
    d = dam.Decontamination_SOM(64, np.arange(100), np.ones(100), np.zeros((100, 3), np.float32), np.ones(100),
                                5, 6, 3, True)
    d._winners = winners
    gnd, gndm = d._compute_gndm(cat, foot)
    n_gal, n_pix, gnd_o, gndm_o = R.gndm(cat, foot, winners, 5, 6)
    assert np.array_equal(gnd, gnd_o, equal_nan=True) and np.array_equal(gndm, gndm_o, equal_nan=True)
    assert d.n_gal == n_gal and d.n_pix == n_pix


def test_ds_divide_and_constructor():
    a, b = np.array([1.0, 2.0]), np.array([2.0, 4.0])
    assert np.array_equal(ds_divide(a, b), [0.5, 0.5])
    ga = lambda: (c for c in (a[:1], a[1:]))
    gb = lambda: (c for c in (b[:1], b[1:]))
    assert [list(c) for c in ds_divide(ga, gb)()] == [[0.5], [0.5]]
    with pytest.raises(ValueError):
        ds_divide(a, gb)
    with pytest.raises(NotImplementedError):
        dam.Decontamination_SOM(64, a, a, np.zeros((2, 3)), a, 2, 2, 3, False)
    d = dam.Decontamination_SOM(64, a, b, np.zeros((2, 3), np.float32), a, 2, 3, 3, True, topology="square")
    assert (d.m, d.n, d.dim, d.topology, d.batch) == (2, 3, 3, "square", True)
    assert np.array_equal(d.corrected_galaxy_number_density, [0.5, 0.5])
    assert d.sigma == 1.5 and d.pca.dim == 3 and d.winners is None
