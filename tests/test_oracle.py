"""Pins the CPU oracle (oracle/) before anything is checked against it.

Parity status of each oracle function (what pins it):
  find_bmus / _find_bmu_xpu      golden winners (test_som_abstract.py:57-60), golden activation counts,
                                 brute force (numpy argmin with lexicographic ties, NaN rows).
  count_bmus                     golden activation map (test_som_abstract.py:41-46), sum invariant,
                                 integer-weight == duplication invariant.
  compute_errors / top2          brute force top-2 by numpy sort, closed forms (x_i = w_j gives QE = 0;
                                 constructed TE cases on square and hexagonal maps), rho linearity.
  train epoch step + divide      golden epoch centroids (test_som_batch.py:30-50, rtol 1e-4), closed
                                 forms (sigma -> 1e-12 clamp gives the k-means step; sigma huge gives
                                 the global mean), chunking invariance.
  train iter step                golden vector-count centroids (test_som_batch.py:63-83, rtol 1e-4).
  factorised FP64 epoch / top2   agreement with the C FP64 kernels on random inputs.
  top2_omp (parallel top-2, d1)  the sequential top2 above (ties, NaN rows), d1 against the mode's distance
                                 formula, golden winners.
  activation_map_omp             golden activation map at 1, 3 and 8 threads; the sequential map with
                                 real weights.
  trajectory_fp64                golden epoch centroids (one epoch); three epochs against the sequential
                                 FP64 restatement of SOM_Batch.train.
  HEALPix pixel-centre z         the reference's own local-coordinate formula restated independently
                                 (hp/__init__.py:376-419 with u = v = 0.5), equal-area symmetry.
QE/TE values on hexagonal maps have no reference number (the reference's test asserts nothing,
test_som_abstract.py:68-73): they are pinned by the closed forms and brute force above only.
"""
import json
import os

import numpy as np
import pytest

from oracle import FP64, REF32, build as build_oracle
from oracle import som_ref as R

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


@pytest.fixture(scope="module", autouse=True)
def _built():
    build_oracle()


def golden_data(n, seed=0):
    return np.random.default_rng(seed=seed).random((n, 4), np.float32)


@pytest.mark.parametrize("mode", [REF32, FP64])
def test_golden_epoch_centroids(mode):
    W0 = R.init_rand(4, 4, 4, seed=0)
    W, hist, qe, te = R.train(W0, 4, 4, golden_data(1000), n_epochs=1, topology="square", mode=mode)
    assert np.allclose(W.reshape(4, 4, 4), np.array(GOLD["centroids_epoch"]), rtol=GOLD["centroids_epoch_rtol"])


@pytest.mark.parametrize("mode", [REF32, FP64])
def test_golden_iter_centroids(mode):
    W0 = R.init_rand(4, 4, 4, seed=0)
    W, hist, qe, te = R.train(W0, 4, 4, golden_data(1000), n_vectors=1000, topology="square", mode=mode)
    assert np.allclose(W.reshape(4, 4, 4), np.array(GOLD["centroids_iter"]), rtol=GOLD["centroids_iter_rtol"])


def test_golden_activation_and_winners():
    W0 = R.init_rand(4, 4, 4, seed=0)
    act = R.activation_map(W0, golden_data(25_000), m=4, n=4)
    assert np.array_equal(act, np.array(GOLD["activation_map"], dtype=np.float64))
    assert np.array_equal(R.find_bmus(W0, golden_data(8)), np.array(GOLD["winners"]))


def test_init_rand_matches_dataset_prefix():
    # init_rand(0) draws the same stream as the golden datasets: its rows are data points
    assert np.array_equal(R.init_rand(4, 4, 4, seed=0), golden_data(16))


def _brute_top2(W, X):
    """Independent top-2: numpy FP64-sum of FP32 squared differences, lexicographic (d, index)."""
    diff = (X[:, None, :].astype(np.float32) - W[None, :, :].astype(np.float32)).astype(np.float32)
    d = (diff * diff).astype(np.float32).astype(np.float64)
    d = np.add.reduce(d, axis=2)  # FP64 accumulation
    d = np.where(np.isnan(d), np.inf, d)
    idx = np.broadcast_to(np.arange(W.shape[0]), d.shape)
    order = np.lexsort((idx, d), axis=1)
    c1 = order[:, 0]
    c2 = order[:, 1] if W.shape[0] > 1 else np.full(X.shape[0], -1)
    d1 = np.take_along_axis(d, order[:, :1], 1)[:, 0]
    c1 = np.where(np.isfinite(d1), c1, -1)
    return c1, c2, d1


def test_find_bmus_brute_force_with_ties_and_nans():
    rng = np.random.default_rng(3)
    W = rng.integers(0, 3, (40, 3)).astype(np.float32)  # many exact duplicates -> ties
    X = rng.integers(0, 3, (500, 3)).astype(np.float32)
    X[7] = np.nan
    c1, c2, _ = _brute_top2(W, X)
    got = R.find_bmus(W, X)
    exp = np.where(c1 < 0, 0, c1)
    assert np.array_equal(got, exp)
    g1, g2 = R.top2(W, R.topography(5, 8, "square"), X)
    assert np.array_equal(g1[c1 >= 0], c1[c1 >= 0])
    assert np.array_equal(g2[c1 >= 0], c2[c1 >= 0])
    assert g1[7] == -1 and g2[7] == -1


def test_errors_closed_forms_square():
    # samples placed exactly on neurons: QE = 0. Each sample's second BMU is chosen by construction.
    m, n = 3, 3
    topo = R.topography(m, n, "square")
    W = (topo * 10.0).astype(np.float32)  # neuron j at 10 * grid position, far apart
    x0 = W[4] + np.array([1.0, 0.0], np.float32)  # near centre neuron, second = (1,2) right neighbour
    x1 = W[0] + np.array([0.0, 0.0], np.float32)  # on neuron 0; second = neighbours at distance 10 (tie -> 1)
    X = np.stack([x0, x1]).astype(np.float32)
    e = R.errors_raw(W, topo, X, None, "square")
    assert e[2] == 2.0
    assert np.isclose(e[0], 1.0)  # sqrt(1) + sqrt(0)
    # neighbours (1,1)-(1,2) and (0,0)-(0,1) are at squared distance 1 <= 2: no topographic error
    assert e[1] == 0.0
    # move x1 so that its second BMU is the diagonal-far neuron (2,2): |t|^2 = 8 > 2
    W2 = W.copy()
    W2[8] = W2[0] + np.array([0.5, 0.5], np.float32)
    e2 = R.errors_raw(W2, topo, X[1:], None, "square")
    assert e2[1] == 1.0 and e2[2] == 1.0


def test_errors_hex_threshold_uses_fp32_topography():
    # hexagonal diagonal neighbours: squared topographic distance in the reference arithmetic can be
    # slightly above 1.0 -> counted as topographic error (strict '>')
    topo = R.topography(4, 4, "hexagonal")
    t = topo.astype(np.float32)
    d = ((t[1] - t[0]).astype(np.float32) ** 2).astype(np.float64).sum()
    W = np.zeros((16, 2), np.float32)
    W[0] = [0.0, 0.0]
    W[1] = [1.0, 0.0]
    W[2:] = 100.0
    X = np.array([[0.1, 0.0]], np.float32)
    e = R.errors_raw(W, topo, X, None, "hexagonal")
    assert e[1] == (1.0 if d > 1.0 else 0.0)


def test_errors_weight_linearity_and_skips():
    rng = np.random.default_rng(5)
    W = rng.random((20, 3), np.float32)
    topo = R.topography(4, 5, "hexagonal")
    X = rng.random((300, 3), np.float32)
    rho = rng.random(300)
    rho[::7] = 0.0
    rho[::11] = -1.0
    e1 = R.errors_raw(W, topo, X, rho, "hexagonal")
    e2 = R.errors_raw(W, topo, X, 2.0 * rho, "hexagonal")
    assert np.allclose(e2, 2.0 * e1, rtol=1e-12)
    keep = rho > 0
    e3 = R.errors_raw(W, topo, X[keep], rho[keep], "hexagonal")
    assert np.allclose(e1, e3, rtol=1e-12)


def test_activation_duplication_invariant():
    rng = np.random.default_rng(6)
    W = rng.random((12, 5), np.float32)
    X = rng.random((400, 5), np.float32)
    k = rng.integers(0, 4, 400)
    a = R.activation_map(W, X, k.astype(np.float64))
    b = R.activation_map(W, np.repeat(X, k, axis=0))
    assert np.array_equal(a, b)
    assert a.sum() == k.sum()


def test_epoch_sigma_clamp_is_kmeans_step():
    rng = np.random.default_rng(7)
    X = rng.random((2000, 3), np.float32)
    W0 = rng.random((9, 3), np.float32)
    W, *_ = R.train(W0, 3, 3, X, n_epochs=1, sigma=1e-20, topology="square", use_best_epoch=False, mode=FP64)
    bmu = R.find_bmus(W0, X, FP64)
    for j in range(9):
        sel = X[bmu == j].astype(np.float64)
        exp = sel.mean(0) if len(sel) else np.zeros(3)
        assert np.allclose(W[j], exp, rtol=1e-6, atol=1e-7)


def test_epoch_sigma_huge_is_global_mean():
    rng = np.random.default_rng(8)
    X = rng.random((1500, 4), np.float32)
    W0 = rng.random((6, 4), np.float32)
    W, *_ = R.train(W0, 2, 3, X, n_epochs=1, sigma=1e8, use_best_epoch=False, mode=FP64)
    assert np.allclose(W, X.astype(np.float64).mean(0)[None, :], rtol=1e-6)


def test_chunked_equals_single_chunk():
    rng = np.random.default_rng(9)
    X = rng.random((999, 4), np.float32)
    rho = rng.random(999)
    W0 = rng.random((16, 4), np.float32)

    def gen(a, s):
        return lambda: lambda: (a[i:i + s] for i in range(0, len(a), s))

    a = R.train(W0, 4, 4, X, rho, n_epochs=2, mode=FP64)
    b = R.train(W0, 4, 4, gen(X, 100), gen(rho, 100), n_epochs=2, mode=FP64)
    assert np.allclose(a[0], b[0], rtol=1e-12)
    assert np.allclose(a[2], b[2], rtol=1e-12)


def test_factorised_fp64_matches_c_fp64():
    rng = np.random.default_rng(10)
    m, n, d = 5, 6, 7
    X = rng.standard_normal((3000, d)).astype(np.float32)
    rho = rng.random(3000)
    W0 = rng.standard_normal((m * n, d)).astype(np.float32)
    topo = R.topography(m, n, "hexagonal")
    sigma = R.sigma_at(max(m, n) / 2.0, 0, 3)
    Wf, c1, c2, d1 = R.factorised_epoch_fp64(W0, topo, X, rho, sigma)
    num, den = R.epoch_accumulate(W0, topo, X, rho, 0, 3, np.float32(max(m, n) / 2.0), FP64)
    Wc = R.divide(num, den, FP64)
    assert np.allclose(Wf, Wc, rtol=1e-6, atol=1e-7)
    g1, g2 = R.top2(W0, topo, X, None, FP64)
    assert np.array_equal(c1, g1) and np.array_equal(c2, g2)


def _z_reference_formula(nside, pix):
    """decontamination/hp/__init__.py:120-131 (nest2xyf) and :376-419 (_xyf2loc, u = v = 0.5), restated."""
    f = pix // (nside * nside)
    v = pix & (nside * nside - 1)

    def compress(b):
        out = np.zeros_like(b)
        for bit in range(32):
            out |= ((b >> (2 * bit)) & 1) << bit
        return out

    x = (compress(v) + 0.5) / nside
    y = (compress(v >> 1) + 0.5) / nside
    r = 1.0 - np.floor_divide(f, 4.0)
    h = r - 1.0 + (x + y)
    mm = 2.0 - r * h
    z = np.where(mm > 1.0, h * 2.0 / 3.0, r * (1.0 - mm ** 2 / 3.0))
    return z


@pytest.mark.parametrize("nside", [1, 2, 4, 16, 64])
def test_healpix_pixel_centre_z(nside):
    from decontamination_amd.synth import nest_pixel_z
    pix = np.arange(12 * nside * nside, dtype=np.int64)
    z = nest_pixel_z(nside, pix)
    assert np.allclose(z, _z_reference_formula(nside, pix), rtol=0, atol=1e-12)
    # equal-area pixels: z is symmetric about the equator and uniformly spread
    assert abs(z.sum()) < 1e-9 * len(z)
    assert np.all(np.abs(z) < 1.0)


@pytest.mark.parametrize("mode", [REF32, FP64])
def test_top2_omp_matches_sequential_top2(mode):
    """orc_top2_omp (parallel over samples) against the sequential orc_top2 pinned above, duplicate
    neurons (lowest index wins) and a NaN row included; d1 against the distance formula of the mode."""
    rng = np.random.default_rng(21)
    W = rng.standard_normal((30, 5)).astype(np.float32)
    W[7] = W[3]
    X = rng.standard_normal((2000, 5)).astype(np.float32)
    X[5] = W[3]
    X[11, 2] = np.nan
    topo = R.topography(5, 6)
    c1, c2, d1 = R.top2_omp(W, X, mode)
    g1, g2 = R.top2(W, topo, X, None, mode)
    ok = ~np.isnan(X).any(1)
    assert np.array_equal(c1[ok], g1[ok]) and np.array_equal(c2[ok], g2[ok])
    assert c1[5] == 3 and c2[5] == 7
    if mode == FP64:
        ref = ((X[ok].astype(np.float64) - W[c1[ok]].astype(np.float64)) ** 2).sum(1)
    else:
        diff = (X[ok] - W[c1[ok]]).astype(np.float32)
        ref = (diff * diff).astype(np.float32).astype(np.float64).sum(1)
    assert np.allclose(d1[ok], ref, rtol=1e-12, atol=0)
    assert np.array_equal(R.top2_omp(R.init_rand(4, 4, 4, seed=0), golden_data(8), mode)[0],
                          np.array(GOLD["winners"]))


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_activation_map_omp_matches_sequential(threads):
    W0 = R.init_rand(4, 4, 4, seed=0)
    act = R.activation_map_omp(W0, golden_data(25_000), threads=threads)
    assert np.array_equal(act.reshape(4, 4), np.array(GOLD["activation_map"], dtype=np.float64))
    rng = np.random.default_rng(22)
    X = rng.standard_normal((3000, 6)).astype(np.float32)
    W = rng.standard_normal((12, 6)).astype(np.float32)
    rho = rng.random(3000)
    for mode in (REF32, FP64):
        assert np.allclose(R.activation_map_omp(W, X, rho, mode, threads), R.activation_map(W, X, rho, mode),
                           rtol=1e-12, atol=0)


def test_trajectory_fp64_golden_and_multi_epoch():
    """One epoch from init_rand(0) on the reference's 4x4 square test reproduces its golden centroids
    (test_som_batch.py:30-50, rtol 1e-4); three epochs equal the sequential FP64 restatement of
    SOM_Batch.train (last epoch's weights) to FP32 rounding."""
    W0 = R.init_rand(4, 4, 4, seed=0)
    X = golden_data(1000)
    W, c1, qe = R.trajectory_fp64(W0, R.topography(4, 4, "square"), X, 2.0, 1)
    assert W.dtype == np.float32
    assert np.allclose(W.reshape(4, 4, 4), np.array(GOLD["centroids_epoch"]), rtol=GOLD["centroids_epoch_rtol"])
    rng = np.random.default_rng(23)
    Xh = rng.standard_normal((1500, 3)).astype(np.float32)
    Wh0 = rng.standard_normal((20, 3)).astype(np.float32)
    Wt, _, qet = R.trajectory_fp64(Wh0, R.topography(4, 5), Xh, 2.5, 3)
    Ws, hist, qes, _ = R.train(Wh0, 4, 5, Xh, n_epochs=3, sigma=2.5, use_best_epoch=False, mode=FP64)
    assert np.allclose(Wt, Ws, rtol=1e-6, atol=1e-6)
    assert abs(qet - qes[-1]) <= 1e-5 * qes[-1]  # the QE of the final weights


# ---------------------------------------------------------------- float64 SOMs ---------------

def test_float64_oracle_equals_c_fp64_on_fp32_inputs():
    """On data, weights and topography that FP32 represents exactly (square grid: integer coordinates),
    the float64 restatements and the C FP64 kernels compute the same quantities: top-2 bit for bit,
    errors, one epoch and the vector-count mode to FP64 rounding."""
    rng = np.random.default_rng(31)
    m, n, dim = 6, 7, 8
    W = rng.standard_normal((m * n, dim)).astype(np.float32)
    X = rng.standard_normal((700, dim)).astype(np.float32)
    W[30] = W[3]
    X[:4] = W[3]
    X[7, 2] = np.nan
    X[9] = np.nan
    rho = rng.random(700)
    rho[::11] = 0.0
    rho[1::13] = -1.0
    topo = R.topography(m, n, "square")
    c1, c2, d1 = R.top2_f64(W, X)
    e1, e2, ed = R.top2_omp(W, X, FP64)
    assert np.array_equal(c1, e1) and np.array_equal(c2, e2) and np.array_equal(d1, ed, equal_nan=True)
    assert c1[0] == 3 and c2[0] == 30 and c1[9] == -1  # tie: lower index first; NaN row: none
    assert np.array_equal(R.find_bmus_f64(W, X), R.find_bmus(W, X, FP64))
    assert np.allclose(R.activation_map_f64(W, X, rho), R.activation_map(W, X, rho, FP64), rtol=1e-12, atol=0)
    assert np.allclose(R.errors_f64(W, topo, X, rho, "square"), R.compute_errors(W, topo, X, rho, "square", FP64),
                       rtol=1e-12, atol=0)
    sigma0 = np.float32(3.0)
    ok = (rho > 0) & (c1 >= 0)
    want, _, _, _ = R.factorised_epoch_fp64(W, topo, X[ok], rho[ok], R.sigma_at(sigma0, 1, 5), bmu=c1[ok],
                                            out_dtype=np.float64)
    assert np.allclose(R.epoch_f64(W, topo, X, rho, sigma0, 1, 5), want, rtol=1e-12, atol=1e-14)
    num, den = R.epoch_accumulate(W, topo, X, rho, 10, 5000, sigma0, FP64, True)
    want = np.where(den[:, None] > 1e-12, num / np.where(den[:, None] > 1e-12, den[:, None], 1.0), 0.0)
    assert np.allclose(R.iter_f64(W, topo, X, rho, sigma0, 10, 5000), want, rtol=1e-12, atol=1e-14)


def test_float64_oracle_brute_force_and_fp64_topography():
    """The float64 top-2 on values FP32 cannot hold equals a per-neuron Python loop of the reference rule;
    and the hexagonal threshold uses FP64 topography: the grid neighbours (0, 2) and (0, 3) lie at
    1.0000000000000004 > 1 in FP64 (a topographic error) but at 0.99999988 <= 1 with the FP32
    topography of float32 SOMs (none)."""
    rng = np.random.default_rng(32)
    W = rng.standard_normal((9, 5)) / 3.0
    X = rng.standard_normal((40, 5)) / 7.0
    assert not np.array_equal(X.astype(np.float32).astype(np.float64), X)
    c1, c2, d1 = R.top2_f64(W, X)
    for i in range(X.shape[0]):
        best = []
        for j in range(W.shape[0]):
            d = 0.0
            for k in range(W.shape[1]):
                t = X[i, k] - W[j, k]
                d += t * t
            best.append((d, j))
        best.sort()
        assert (c1[i], c2[i], d1[i]) == (best[0][1], best[1][1], best[0][0])
    t64 = R.topography(2, 4, "hexagonal", np.float64)
    assert R.sqd_f64(t64[2], t64[3]) == 1.0000000000000004
    Wn = np.full((8, 2), 100.0)
    Wn[2] = [0.0, 0.0]
    Wn[3] = [0.1, 0.0]
    Xs = np.array([[0.04, 0.0]])
    assert R.errors_f64(Wn, t64, Xs, None, "hexagonal") == (0.04, 1.0)
    _, te32 = R.compute_errors(Wn.astype(np.float32), R.topography(2, 4), Xs.astype(np.float32), None, "hexagonal",
                               FP64)
    assert te32 == 0.0


@pytest.mark.parametrize("topology", ["square", "hexagonal"])
def test_float64_factorised_epoch_equals_per_sample_form(topology):
    """factorised_epoch_f64 (H times the per-winner sums) equals the per-sample scatter of epoch_f64 to
    FP64 rounding, with and without densities, zero densities, a NaN row (no winner) and the schedule's
    sigma at a later epoch."""
    rng = np.random.default_rng(33)
    m, n, dim = 5, 6, 4
    W = rng.standard_normal((m * n, dim)) / 3.0
    X = rng.standard_normal((400, dim)) / 5.0
    X[7] = np.nan
    rho = rng.random(400) * 2.0
    rho[::9] = 0.0
    topo = R.topography(m, n, topology, np.float64)
    c1, _, _ = R.top2_f64(W, X)
    assert c1[7] == -1
    for r, e in ((None, 0), (rho, 0), (rho, 3)):
        want = R.epoch_f64(W, topo, X, r, 2.5, e, 5)
        got = R.factorised_epoch_f64(W, topo, X, r, 2.5, e, 5, c1)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-15), np.abs(got - want).max()
