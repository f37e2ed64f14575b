"""Covariance with constant memory usage, computed on the GPU.

Mirrors ``decontamination.algo.covariance.Covariance`` (decontamination/algo/covariance.py:20-280):
``compute`` (:124-212), ``diagonalize`` (:217-254) and ``project_pca`` (:259-280), with the same
arguments, row rule (a row counts when its weight is finite and > 0 and all its values are finite),
``ddof`` normalisation and errors.

The reference updates Welford's sums sample by sample on one CPU thread. Here every chunk goes
through the uploader ring and two HBM-bound passes on the device (``somhip_covariance_chunk``: the
chunk's total weight, mean and co-moment matrix from the corrected two-pass sums, reduced in a
fixed order), and the chunks are merged on the host with Chan's update. Both are exact algorithms
for the same quantity; they agree to FP64 rounding.
"""
import ctypes
import typing

import numpy as np

from . import _lib
from .runtime import Uploader, get_context
from .som_abstract import iterate_chunks


def _merge(acc, wc, mc, m2c):
    """Chan et al.'s pairwise update of (W, mean, M2) with one chunk's (W, mean, M2)."""
    w, mean, m2 = acc
    if wc <= 0.0:
        return acc
    if w <= 0.0:
        return wc, mc.copy(), m2c.copy()
    wn = w + wc
    delta = mc - mean
    return wn, mean + delta * (wc / wn), m2 + m2c + np.outer(delta, delta) * (w * wc / wn)


class Covariance(object):
    """Covariance calculation with constant memory usage (GPU)."""

    @staticmethod
    def chunk_moments(ctx, uploader, vectors, weights, dim):
        """(W, mean, M2) of one chunk, computed on the device."""
        v = np.asarray(vectors)
        if v.dtype != np.float32:
            v = v.astype(np.float64, copy=False)
        v = np.ascontiguousarray(v)
        if v.ndim != 2 or v.shape[1] != dim:
            raise ValueError(f'dataset chunks must have shape (N, {dim})')
        r = None
        if weights is not None:
            r = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
            if r.shape[0] != v.shape[0]:
                raise ValueError('`dataset` and `dataset_weights` chunks must be aligned')
        xp, rp = uploader.put(v, r)
        out = np.empty(1 + dim + dim * dim, np.float64)
        _lib.call("somhip_covariance_chunk", ctx.handle, ctypes.c_void_p(xp), 1 if v.dtype == np.float64 else 0,
                  None if rp is None else ctypes.c_void_p(rp), v.shape[0], dim, _lib.vp(out))
        return out[0], out[1:1 + dim], out[1 + dim:].reshape(dim, dim)

    @staticmethod
    def compute(dim: int, dataset: typing.Union[np.ndarray, typing.Callable],
                dataset_weights: typing.Optional[typing.Union[np.ndarray, typing.Callable]] = None, ddof: int = 1,
                show_progress_bar: bool = False, *, device: int = 0) -> np.ndarray:
        """Computes the (dim, dim) FP64 covariance matrix of the given dataset (array or generator
        builder), weighted by ``dataset_weights``, normalised by sum(weights) - ddof."""
        if ddof < 0:
            raise ValueError('ddof must be greater than or equal to zero.')
        if not 1 <= dim <= 64:
            raise NotImplementedError('the GPU covariance supports 1 <= dim <= 64')
        ctx = get_context(device)
        up = Uploader(ctx, slots=2)
        acc = (0.0, np.zeros(dim, np.float64), np.zeros((dim, dim), np.float64))
        chunks = iterate_chunks(dataset, dataset_weights)
        if show_progress_bar:
            import tqdm
            chunks = tqdm.tqdm(chunks)
        try:
            for v, w in chunks:
                acc = _merge(acc, *Covariance.chunk_moments(ctx, up, v, w, dim))
        finally:
            up.close()
        sum_w, _, m2 = acc
        norm = sum_w - ddof
        if norm <= 0.0:
            raise ValueError('Empty dataset or degrees of freedom <= 0.')
        return m2 / norm

    @staticmethod
    def diagonalize(cov_matrix: np.ndarray, sort: bool = True) -> typing.Tuple[np.ndarray, np.ndarray]:
        """Eigenvalues and eigenvectors of a covariance matrix, by decreasing eigenvalue when ``sort``."""
        if cov_matrix.shape[0] != cov_matrix.shape[1]:
            raise ValueError('Input covariance matrix is not square.')
        eigenvalues, eigenvectors = np.linalg.eigh(cov_matrix)
        if sort:
            eigenvalues = np.flip(eigenvalues, axis=0)
            eigenvectors = np.flip(eigenvectors, axis=1)
        return eigenvalues, eigenvectors

    @staticmethod
    def project_pca(dataset: np.ndarray, eigenvectors: np.ndarray) -> np.ndarray:
        """Projects a dataset onto the PCA basis (dataset . eigenvectors)."""
        return np.dot(dataset, eigenvectors)
