"""Restated host driver of the reference SOM path on top of the C oracle kernels.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows, step by step:
  * topography        decontamination/algo/som_abstract.py:81-111
  * init_rand         decontamination/algo/som_abstract.py:115-132
  * compute_errors    decontamination/algo/som_abstract.py:528-601
  * activation map    decontamination/algo/som_abstract.py:605-666
  * get_winners       decontamination/algo/som_abstract.py:670-695
  * SOM_Batch.train   decontamination/algo/som_batch.py:214-505 (epoch and vector-count modes,
                      history, per-epoch errors, early stop, best epoch)
  * chunk protocol    decontamination/algo/__init__.py:21-27

``factorised_epoch_fp64`` / ``top2_fp64_np`` are a fast FP64 oracle for inputs too large for the
per-sample C loop; they are pinned against the C FP64 kernels in tests/test_oracle.py.
"""
import numpy as np

from . import FP64, REF32, f32, f64, lib, ptr

FLOAT_EPS = 1.0e-12


def topography(m, n, topology="hexagonal", dtype=np.float32):
    """som_abstract.py:81-111: index i*n+j; square [i, j]; hexagonal [i + 0.5*(j&1), j*sqrt(3)/2]."""
    if topology == "square":
        pts = [[i, j] for i in range(m) for j in range(n)]
    else:
        pts = [[i + (j & 1) * 0.5, j * 0.8660254037844386] for i in range(m) for j in range(n)]
    return np.array(pts, dtype=dtype).reshape(m * n, 2)


def init_rand(m, n, dim, seed=None, dtype=np.float32):
    """som_abstract.py:128-132."""
    return np.random.default_rng(seed=seed).random((m * n, dim), dtype=dtype)


def _chunks(dataset):
    """algo/__init__.py:21-27: ndarray -> one chunk; callable builder -> builder()()."""
    if dataset is None:
        return None
    if callable(dataset):
        return dataset()()
    return iter((dataset,))


def find_bmus(W, X, mode=REF32):
    W = f32(W); X = f32(X)
    out = np.empty(X.shape[0], np.int32)
    lib().orc_find_bmus(mode, ptr(W), W.shape[0], W.shape[1], ptr(X), X.shape[0], ptr(out))
    return out


def activation_map(W, X, rho=None, mode=REF32, m=None, n=None):
    W = f32(W); X = f32(X); rho = f64(rho)
    res = np.zeros(W.shape[0], np.float64)
    lib().orc_count_bmus(mode, ptr(W), W.shape[0], W.shape[1], ptr(X), ptr(rho), X.shape[0], ptr(res))
    return res if m is None else res.reshape(m, n)


def errors_raw(W, topo, X, rho=None, topology="hexagonal", mode=REF32):
    W = f32(W); X = f32(X); rho = f64(rho); topo = f32(topo)
    res = np.zeros(3, np.float64)
    lib().orc_compute_errors(mode, ptr(W), ptr(topo), W.shape[0], W.shape[1], ptr(X), ptr(rho), X.shape[0],
                             2.0 if topology == "square" else 1.0, ptr(res))
    return res


def compute_errors(W, topo, dataset, dataset_weights=None, topology="hexagonal", mode=REF32):
    """som_abstract.py:547-601 (NaN pair when mn < 2 or the weight sum is 0)."""
    if W.shape[0] < 2:
        return np.nan, np.nan
    res = np.zeros(3, np.float64)
    dch = _chunks(dataset)
    wch = _chunks(dataset_weights)
    pairs = zip(dch, wch) if wch is not None else ((c, None) for c in dch)
    for xc, rc in pairs:
        res += errors_raw(W, topo, xc, rc, topology, mode)
    return (res[0] / res[2], res[1] / res[2]) if res[2] > 0.0 else (np.nan, np.nan)


def top2(W, topo, X, rho=None, mode=REF32):
    W = f32(W); X = f32(X); rho = f64(rho); topo = f32(topo)
    c1 = np.empty(X.shape[0], np.int32); c2 = np.empty(X.shape[0], np.int32)
    lib().orc_top2(mode, ptr(W), ptr(topo), W.shape[0], W.shape[1], ptr(X), ptr(rho), X.shape[0], ptr(c1), ptr(c2))
    return c1, c2


def epoch_accumulate(W, topo, X, rho, cur, total, sigma0, mode=REF32, iter_mode=False, acc=None):
    """One launch of _train_step1_{epoch,iter}_kernel over a chunk; returns (num, den)."""
    W = f32(W); X = f32(X); rho = f64(rho); topo = f32(topo)
    mn, dim = W.shape
    dt = np.float64 if mode == FP64 else np.float32
    if acc is None:
        acc = (np.zeros((mn, dim), dt), np.zeros(mn, dt))
    num, den = acc
    fn = lib().orc_train_iter_accumulate if iter_mode else lib().orc_train_epoch_accumulate
    fn(mode, ptr(num), ptr(den), ptr(W), ptr(topo), ptr(X), ptr(rho), X.shape[0], cur, total,
       np.float32(sigma0), mn, dim)
    return acc


def divide(num, den, mode=REF32):
    mn, dim = num.shape
    out = np.empty((mn, dim), np.float32)
    lib().orc_divide(mode, ptr(num), ptr(den), mn, dim, ptr(out))
    return out


def train(W0, m, n, dataset, dataset_weights=None, n_epochs=None, n_vectors=None, sigma=None,
          topology="hexagonal", use_best_epoch=True, stop_quantization_error=None, stop_topographic_error=None,
          mode=REF32):
    """Restatement of SOM_Batch.train (som_batch.py:214-505). Returns (weights, history, qe, te)."""
    sigma = max(m, n) / 2.0 if sigma is None else float(sigma)
    sigma0 = np.float32(sigma)
    topo = topography(m, n, topology)
    W = f32(W0).copy()
    mn, dim = W.shape
    sq = -1.0e6 if stop_quantization_error is None else stop_quantization_error
    st = -1.0e6 if stop_topographic_error is None else stop_topographic_error

    def chunk_pairs():
        dch = _chunks(dataset)
        wch = _chunks(dataset_weights)
        if wch is None:
            for c in dch:
                yield c, None
        else:
            for c, r in zip(dch, wch):
                if c.shape[0] != r.shape[0]:
                    raise ValueError("`dataset` and `dataset_weights` chunks must be aligned")
                yield c, r

    if n_epochs is not None and n_vectors is None:
        history = np.empty((n_epochs, mn, dim), np.float32)
        qe = np.full(n_epochs, np.nan); te = np.full(n_epochs, np.nan)
        for e in range(n_epochs):
            acc = None
            for c, r in chunk_pairs():
                acc = epoch_accumulate(W, topo, c, r, e, n_epochs, sigma0, mode, False, acc)
            W = history[e] = divide(*acc, mode)
            qe[e], te[e] = compute_errors(W, topo, dataset, dataset_weights, topology, mode)
            if qe[e] <= sq and te[e] <= st:
                break
    elif n_epochs is None and n_vectors is not None:
        history = np.empty((1, mn, dim), np.float32)
        qe = np.zeros(1); te = np.zeros(1)
        acc = None
        cur = 0
        for c, r in chunk_pairs():
            count = min(c.shape[0], n_vectors - cur)
            acc = epoch_accumulate(W, topo, c[:count], None if r is None else r[:count], cur, n_vectors, sigma0,
                                   mode, True, acc)
            cur += count
            if cur >= n_vectors:
                break
        W = history[0] = divide(*acc, mode)
        qe[0], te[0] = compute_errors(W, topo, dataset, dataset_weights, topology, mode)
    else:
        raise ValueError("Invalid training method, specify either `n_epochs` or `n_vectors`")
    if use_best_epoch:
        W = history[np.nanargmin(qe)].copy()
    return W, history, qe, te


# ---------------------------------------------------------------------------------------------
# Fast factorised FP64 oracle (large inputs).
# ---------------------------------------------------------------------------------------------

def top2_fp64_np(W, X, block=8192, k=4):
    """FP64 top-2 BMUs with lowest-index tie-break: expanded-form FP64 screen, then the direct FP64
    formula on the k best candidates. Returns (c1, c2, d1) with d1 the direct FP64 squared distance.
    Pinned against orc_top2(FP64) in tests."""
    W64 = np.asarray(W, np.float32).astype(np.float64)
    nw = (W64 * W64).sum(1)
    N = X.shape[0]
    c1 = np.empty(N, np.int64); c2 = np.empty(N, np.int64); d1 = np.empty(N)
    for s in range(0, N, block):
        x = np.asarray(X[s:s + block], np.float32).astype(np.float64)
        d = nw[None, :] - 2.0 * (x @ W64.T)
        cand = np.argpartition(d, k, axis=1)[:, :k]
        diff = x[:, None, :] - W64[cand]
        dd = (diff * diff).sum(-1)
        # lexicographic (distance, index) per row: stable sort by index, then by distance
        key_idx = np.argsort(cand, axis=1, kind="stable")
        cand_s = np.take_along_axis(cand, key_idx, 1)
        dd_s = np.take_along_axis(dd, key_idx, 1)
        o = np.argsort(dd_s, axis=1, kind="stable")
        c1[s:s + block] = np.take_along_axis(cand_s, o[:, :1], 1)[:, 0]
        c2[s:s + block] = np.take_along_axis(cand_s, o[:, 1:2], 1)[:, 0]
        d1[s:s + block] = np.take_along_axis(dd_s, o[:, :1], 1)[:, 0]
    return c1, c2, d1


def top2_omp(W, X, mode=FP64):
    """Per-sample top-2 (c1, c2, d1) of the _compute_errors_xpu rule (som_abstract.py:832-853), exact
    distance formula in `mode` for every neuron, parallel over samples (orc_top2_omp)."""
    W = f32(W); X = f32(X)
    n = X.shape[0]
    c1 = np.empty(n, np.int32); c2 = np.empty(n, np.int32); d1 = np.empty(n, np.float64)
    lib().orc_top2_omp(mode, ptr(W), W.shape[0], W.shape[1], ptr(X), n, ptr(c1), ptr(c2), ptr(d1))
    return c1, c2, d1


def trajectory_fp64(W0, topo, X, sigma0, n_epochs):
    """The epoch-mode schedule (som_batch.py:214-375 with _train_step2_xpu in its factorised form) in FP64
    arithmetic from W0: exact FP64 BMUs of every sample, FP64 per-BMU sums, num = H.S, den = H.c, division;
    the weights are rounded to FP32 after every epoch (the reference keeps them in its dtype, FP32).
    Returns (final FP32 weights, BMUs of the last epoch, QE of the final weights)."""
    W = f32(W0)
    c1 = None
    sq = topo_sqdist(topo)
    for e in range(n_epochs):
        c1, _, _ = top2_omp(W, X, FP64)
        W, _, _, _ = factorised_epoch_fp64(W, topo, X, None, sigma_at(sigma0, e, n_epochs), bmu=c1,
                                           out_dtype=np.float32, sq=sq)
    _, _, d1 = top2_omp(W, X, FP64)
    return W, c1, float(np.mean(np.sqrt(d1)))


def topo_sqdist(topo):
    """|t_j - t_c|^2 with the reference's FP32 differences and squares, summed in FP64 (a5 on the grid)."""
    t = np.asarray(topo, np.float32)
    diff = (t[:, None, :] - t[None, :, :]).astype(np.float32)
    return (diff * diff).astype(np.float32).astype(np.float64).sum(-1)


def neighbourhood_fp64(topo, sigma, sq=None):
    """H[j, c] = exp(-|t_j - t_c|^2 / (2 sigma^2)) with the reference's FP32 topography differences."""
    if sq is None:
        sq = topo_sqdist(topo)
    return np.exp(-sq * (1.0 / (2.0 * sigma * sigma)))


def sigma_at(sigma0, e, etot):
    s = float(np.float32(sigma0)) * (1.0 / (1.0 + 2.0 * e / etot))
    return max(s, FLOAT_EPS)


def factorised_epoch_fp64(W, topo, X, rho, sigma, bmu=None, out_dtype=np.float32, sq=None):
    """One FP64 batch epoch in the factorised form num = H.S, den = H.c (algebraically identical to
    som_batch.py:509-566 in epoch mode because h depends only on (j, bmu(i))). Returns
    (W_new in out_dtype, bmu, c2, d1)."""
    mn, dim = np.asarray(W).shape
    c1, c2, d1 = top2_fp64_np(W, X) if bmu is None else (bmu, None, None)
    r = np.ones(X.shape[0]) if rho is None else np.asarray(rho, np.float64)
    keep = r > 0
    xk = np.asarray(X, np.float32)[keep].astype(np.float64)
    S = np.stack([np.bincount(c1[keep], weights=xk[:, k] * r[keep], minlength=mn) for k in range(dim)], 1)
    cnt = np.bincount(c1[keep], weights=r[keep], minlength=mn)
    H = neighbourhood_fp64(topo, sigma, sq)
    num = H @ S
    den = H @ cnt
    Wn = np.where(den[:, None] > FLOAT_EPS, num / np.where(den[:, None] > FLOAT_EPS, den[:, None], 1.0), 0.0)
    return Wn.astype(out_dtype), c1, c2, d1


# ---------------------------------------------------------------------------------------------------
# float64 SOMs (dtype=np.float64): the same rules with data, weights, topography, density and sums in
# FP64. Pure numpy over small inputs; pinned in tests/test_oracle.py against the C FP64 kernels on inputs
# that FP32 represents exactly, where both must agree.
# ---------------------------------------------------------------------------------------------------

def sqd_f64(A, B):
    """square_distance_xpu (algo/__init__.py:89-98) on float64 operands, broadcast over leading axes:
    sum over k, in k order, of fl(fl(a_k - b_k)^2)."""
    A = np.asarray(A, np.float64); B = np.asarray(B, np.float64)
    r = np.zeros(np.broadcast_shapes(A.shape[:-1], B.shape[:-1]))
    for k in range(A.shape[-1]):
        t = A[..., k] - B[..., k]
        r = r + t * t
    return r


def top2_f64(W, X, block=256):
    """(c1, c2, d1) of the _compute_errors_xpu top-2 rule (som_abstract.py:816-853) on float64 operands:
    NaN or >= 1e99 distances never enter, ties keep the lower index, -1 / NaN when absent."""
    W = np.asarray(W, np.float64); X = np.asarray(X, np.float64)
    N, mn = X.shape[0], W.shape[0]
    c1 = np.full(N, -1, np.int32); c2 = np.full(N, -1, np.int32); d1 = np.full(N, np.nan)
    for s in range(0, N, block):
        d = sqd_f64(X[s:s + block, None, :], W[None, :, :])
        d = np.where(np.isnan(d) | (d >= 1.0e99), np.inf, d)
        o = np.argsort(d, axis=1, kind="stable")  # (distance, index) order
        first = np.take_along_axis(d, o[:, :1], 1)[:, 0]
        ok1 = first < np.inf
        c1[s:s + block] = np.where(ok1, o[:, 0], -1)
        d1[s:s + block] = np.where(ok1, first, np.nan)
        if mn > 1:
            second = np.take_along_axis(d, o[:, 1:2], 1)[:, 0]
            c2[s:s + block] = np.where(second < np.inf, o[:, 1], -1)
    return c1, c2, d1


def find_bmus_f64(W, X):
    """_find_bmu_xpu (som_abstract.py:764-783) on float64 operands: first minimum, 0 when none."""
    c1, _, _ = top2_f64(W, X)
    return np.where(c1 >= 0, c1, 0).astype(np.int32)


def activation_map_f64(W, X, rho=None):
    """_count_bmus_kernel (som_abstract.py:699-732): result[bmu(x_i)] += rho_i for rho_i > 0."""
    b = find_bmus_f64(W, X)
    r = np.ones(X.shape[0]) if rho is None else np.asarray(rho, np.float64)
    keep = r > 0.0
    return np.bincount(b[keep], weights=r[keep], minlength=np.asarray(W).shape[0])


def errors_f64(W, topo, X, rho=None, topology="hexagonal"):
    """compute_errors (som_abstract.py:547-601, 816-878) of a float64 SOM: (QE, TE), the topography
    distance of the two BMUs in FP64 against r^2 = 2 (square) or 1 (hexagonal)."""
    c1, c2, d1 = top2_f64(W, X)
    r = np.ones(X.shape[0]) if rho is None else np.asarray(rho, np.float64)
    ok = (r > 0.0) & (c1 >= 0) & (c2 >= 0) & (c1 != c2)
    t = np.asarray(topo, np.float64)
    far = sqd_f64(t[c1[ok]], t[c2[ok]]) > (2.0 if topology == "square" else 1.0)
    s = r[ok].sum()
    if W.shape[0] < 2 or not s > 0.0:
        return np.nan, np.nan
    return float((r[ok] * np.sqrt(d1[ok])).sum() / s), float(r[ok][far].sum() / s)


def _accumulate_f64(W, topo, X, rho, inv_2s2):
    """_train_step2_xpu (som_batch.py:509-566) on float64 operands: every sample with rho > 0 and a BMU
    adds h_j = rho exp(-|t_j - t_bmu|^2 inv_2s2) x to num_j and h_j to den_j (inv_2s2 per sample)."""
    c1, _, _ = top2_f64(W, X)
    r = np.ones(X.shape[0]) if rho is None else np.asarray(rho, np.float64)
    keep = (r > 0.0) & (c1 >= 0)
    t = np.asarray(topo, np.float64)
    inv = np.broadcast_to(np.asarray(inv_2s2, np.float64), r.shape)[keep]
    h = r[keep][None, :] * np.exp(-sqd_f64(t[:, None, :], t[c1[keep]][None, :, :]) * inv[None, :])
    return h @ np.asarray(X, np.float64)[keep], h.sum(1)


def _divide_f64(num, den):
    """w = num / den where den > 1e-12, else 0 (som_batch.py:354-359), in FP64."""
    return np.where(den[:, None] > FLOAT_EPS, num / np.where(den[:, None] > FLOAT_EPS, den[:, None], 1.0), 0.0)


def epoch_f64(W, topo, X, rho, sigma0, e, n_epochs):
    """One epoch-mode epoch of a float64 SOM: sigma = sigma0 / (1 + 2 e / E) >= 1e-12 (som_batch.py:107-110)."""
    s = max(float(sigma0) * (1.0 / (1.0 + 2.0 * e / n_epochs)), FLOAT_EPS)
    return _divide_f64(*_accumulate_f64(W, topo, X, rho, 1.0 / (2.0 * s * s)))


def factorised_epoch_f64(W, topo, X, rho, sigma0, e, n_epochs, bmu):
    """epoch_f64 in the factorised form num = H.S, den = H.c for given winners bmu (-1: no winner): in epoch
    mode h depends only on (j, bmu(i)) (som_batch.py:217-228), so the per-sample scatter of
    _train_step2_xpu (som_batch.py:509-566) sums to H times the per-winner sums. FP64 throughout."""
    W = np.asarray(W, np.float64)
    mn, dim = W.shape
    s = max(float(sigma0) * (1.0 / (1.0 + 2.0 * e / n_epochs)), FLOAT_EPS)
    r = np.ones(X.shape[0]) if rho is None else np.asarray(rho, np.float64)
    bmu = np.asarray(bmu)
    keep = (r > 0.0) & (bmu >= 0)
    xk = np.asarray(X, np.float64)[keep]
    S = np.stack([np.bincount(bmu[keep], weights=xk[:, k] * r[keep], minlength=mn) for k in range(dim)], 1)
    cnt = np.bincount(bmu[keep], weights=r[keep], minlength=mn)
    t = np.asarray(topo, np.float64)
    num = np.empty((mn, dim))
    den = np.empty(mn)
    for j0 in range(0, mn, 1024):  # H by row blocks
        H = np.exp(-sqd_f64(t[j0:j0 + 1024, None, :], t[None, :, :]) * (1.0 / (2.0 * s * s)))
        num[j0:j0 + 1024] = H @ S
        den[j0:j0 + 1024] = H @ cnt
    return _divide_f64(num, den)


def iter_f64(W, topo, X, rho, sigma0, cur, n_vectors):
    """Vector-count mode of a float64 SOM over X (som_batch.py:154-210, 386-486): sigma_i =
    sigma0 / (1 + 2 (cur + i) / n_vectors) >= 1e-12 per sample, then the division."""
    i = np.arange(np.asarray(X).shape[0])
    s = np.maximum(float(sigma0) * (1.0 / (1.0 + 2.0 * (cur + i) / n_vectors)), FLOAT_EPS)
    return _divide_f64(*_accumulate_f64(W, topo, X, rho, 1.0 / (2.0 * s * s)))


# ---------------------------------------------------------------------------------------------------
# Covariance / SOM_PCA / galaxy-number-density restatements
# ---------------------------------------------------------------------------------------------------
def covariance(dim, dataset, dataset_weights=None, ddof=1):
    """Covariance.compute (covariance.py:124-212): sequential weighted Welford over the chunks
    (orc_welford_update, covariance.py:30-84), then _finalize_welford_cov (covariance.py:90-119)."""
    if ddof < 0:
        raise ValueError('ddof must be greater than or equal to zero.')
    sum_w = np.zeros(1, np.float64)
    mean = np.zeros(dim, np.float64)
    m2 = np.zeros((dim, dim), np.float64)
    delta = np.empty(dim, np.float64)
    dch = _chunks(dataset)
    wch = _chunks(dataset_weights)
    pairs = ((c, None) for c in dch) if wch is None else zip(dch, wch)
    for v, w in pairs:
        v = np.ascontiguousarray(v, np.float64)
        if w is not None:
            if v.shape[0] != np.shape(w)[0]:
                raise ValueError('`dataset` and `dataset_weights` chunks must be aligned')
            w = np.ascontiguousarray(w, np.float64)
        lib().orc_welford_update(ptr(sum_w), ptr(mean), ptr(m2), ptr(delta), ptr(v), ptr(w), v.shape[0], dim)
    norm = sum_w[0] - ddof
    if norm <= 0.0:
        raise ValueError('Empty dataset or degrees of freedom <= 0.')
    cov = np.empty((dim, dim), np.float64)
    for j in range(dim):
        cov[j, j] = m2[j, j] / norm
        for k in range(j + 1, dim):
            cov[j, k] = cov[k, j] = m2[j, k] / norm
    return cov


def pca_weights(cov, min_weight, max_weight, m, n, dim, scale_by_variance=False, apply_cdf=False, cdf_gain=1.0):
    """SOM_PCA._diag_cov_matrix (som_pca.py:116-177) with Covariance.diagonalize (covariance.py:217-250):
    eigh, flipped to decreasing eigenvalues; the (m, n) grid spans the two leading directions."""
    import math
    ev, evec = np.linalg.eigh(cov)
    ev, evec = np.flip(ev, axis=0), np.flip(evec, axis=1)
    v0, v1 = evec[:, 0], evec[:, 1]
    if scale_by_variance:
        if ev[0] > 0.0:
            v0 = v0 * np.sqrt(ev[0])
        if ev[1] > 0.0:
            v1 = v1 * np.sqrt(ev[1])
    sigma = np.sqrt(v0 * v0 + v1 * v1)
    sigma[sigma == 0.0] = 1.0
    cdf_scale = cdf_gain / (math.sqrt(2.0) * sigma)
    lx = np.linspace(-1.0, 1.0, m) if apply_cdf else np.linspace(min_weight, max_weight, m)
    ly = np.linspace(-1.0, 1.0, n) if apply_cdf else np.linspace(min_weight, max_weight, n)
    W = np.empty((m, n, dim), np.float64)
    for i in range(m):
        for j in range(n):
            for d in range(dim):
                w = v0[d] * lx[i] + v1[d] * ly[j]
                W[i, j, d] = (min_weight + (max_weight - min_weight) * 0.5 * (1.0 + math.erf(w * cdf_scale[d]))
                              if apply_cdf else w)
    return ev, evec, W


def gndm(catalog_activation_map, footprint_activation_map, winners, m, n, dtype=np.float32):
    """Decontamination_SOM._compute_gndm (decontamination_som.py:600-628): n_gal, n_pix, gnd, gndm."""
    n_gal = np.sum(catalog_activation_map)
    n_pix = np.sum(footprint_activation_map)
    gnd = np.divide(catalog_activation_map, footprint_activation_map,
                    out=np.full(catalog_activation_map.shape, np.nan, dtype=dtype),
                    where=footprint_activation_map > 0.0)
    return n_gal, n_pix, gnd, gnd.reshape(m * n)[winners]


def activation_map_omp(W, X, rho=None, mode=REF32, threads=None):
    """activation_map, parallel over samples (orc_count_bmus_omp): per-thread FP64 maps summed at the end.
    For unit weights the sums are integer counts, equal to the sequential loop's."""
    import os
    W = f32(W); X = f32(X); rho = f64(rho)
    t = int(threads or os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    res = np.zeros(W.shape[0], np.float64)
    tm = np.zeros((t, W.shape[0]), np.float64)
    lib().orc_count_bmus_omp(mode, ptr(W), W.shape[0], W.shape[1], ptr(X), ptr(rho), X.shape[0], ptr(res), ptr(tm), t)
    return res
