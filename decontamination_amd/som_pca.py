"""SOM spanning the first two principal components: the initialisation of the decontamination SOM.

Mirrors ``decontamination.algo.som_pca.SOM_PCA`` (decontamination/algo/som_pca.py:23-281): same
constructor, properties, HDF5 header extras, ``train`` (covariance of the dataset, then the grid)
and ``train_from_cov_matrix``. The covariance pass over the samples runs on the GPU
(``Covariance.compute``); the eigen-decomposition of the (dim, dim) matrix and the (m, n, dim) grid
are host work of a few microseconds, written as the reference writes them.
"""
import math
import typing

import numpy as np

from .covariance import Covariance
from .som_abstract import SOM_Abstract


# noinspection PyPep8Naming
class SOM_PCA(SOM_Abstract):
    """Self Organizing Maps that span the first two principal components with constant memory usage."""

    __MODE__ = 'pca'

    def __init__(self, m: int, n: int, dim: int, dtype=np.float32, topology: typing.Optional[str] = None, **kwargs):
        if dim < 2:
            raise ValueError('SOM_PCA requires dim >= 2.')
        super().__init__(m, n, dim, dtype, topology, **kwargs)
        self._scale_by_variance = False
        self._apply_cdf = False
        self._cdf_gain = 1.0
        self._cov_matrix = np.zeros((dim, dim), dtype=np.float64)
        self._eigenvalues = np.zeros((dim,), dtype=np.float64)
        self._eigenvectors = np.zeros((dim, dim), dtype=np.float64)
        self._header_extra = {
            'mode': '__MODE__',
            'scale_by_variance': '_scale_by_variance',
            'apply_cdf': '_apply_cdf',
            'cdf_gain': '_cdf_gain',
        }

    @property
    def cov_matrix(self) -> np.ndarray:
        """Covariance matrix of the training dataset."""
        return self._cov_matrix

    @property
    def eigenvalues(self) -> np.ndarray:
        """Eigenvalues of the covariance matrix."""
        return self._eigenvalues

    @property
    def eigenvectors(self) -> np.ndarray:
        """Eigenvectors of the covariance matrix."""
        return self._eigenvectors

    @staticmethod
    def _diag_cov_matrix(cov_matrix: np.ndarray, min_weight: float, max_weight: float, m: int, n: int, dim: int,
                         scale_by_variance: bool, apply_cdf: bool, cdf_gain: float):
        """Eigen-decomposition and the (m, n, dim) FP64 grid on the two leading directions
        (som_pca.py:116-177): a linear grid over [min_weight, max_weight], or a per-component
        Gaussian CDF of the grid over [-1, 1] when ``apply_cdf``."""
        eigenvalues, eigenvectors = Covariance.diagonalize(cov_matrix, sort=True)
        v0 = eigenvectors[:, 0]
        v1 = eigenvectors[:, 1]
        if scale_by_variance:
            if eigenvalues[0] > 0.0:
                v0 = v0 * np.sqrt(eigenvalues[0])
            if eigenvalues[1] > 0.0:
                v1 = v1 * np.sqrt(eigenvalues[1])
        sigma = np.sqrt(v0 * v0 + v1 * v1)
        sigma[sigma == 0.0] = 1.0
        cdf_scale = cdf_gain / (math.sqrt(2.0) * sigma)
        lo, hi = (-1.0, 1.0) if apply_cdf else (min_weight, max_weight)
        grid_x = np.linspace(lo, hi, m)
        grid_y = np.linspace(lo, hi, n)
        # w[i, j, d] = v0[d] x_i + v1[d] y_j, each product and sum rounded as in the scalar loop
        lin = v0[None, None, :] * grid_x[:, None, None] + v1[None, None, :] * grid_y[None, :, None]
        if apply_cdf:
            erf = np.vectorize(math.erf, otypes=[np.float64])
            weights = min_weight + (max_weight - min_weight) * 0.5 * (1.0 + erf(lin * cdf_scale[None, None, :]))
        else:
            weights = lin
        return eigenvalues, eigenvectors, weights

    def train(self, dataset: typing.Union[np.ndarray, typing.Callable],
              dataset_weights: typing.Optional[typing.Union[np.ndarray, typing.Callable]] = None,
              min_weight: float = 0.0, max_weight: float = 1.0, scale_by_variance: bool = False,
              apply_cdf: bool = False, cdf_gain: float = 1.0, show_progress_bar: bool = False) -> None:
        """Trains the neural network from the given dataset (array or generator builder)."""
        cov_matrix = Covariance.compute(self._dim, dataset, dataset_weights, ddof=1,
                                        show_progress_bar=show_progress_bar, device=self._device)
        self.train_from_cov_matrix(cov_matrix, min_weight=min_weight, max_weight=max_weight,
                                   scale_by_variance=scale_by_variance, apply_cdf=apply_cdf, cdf_gain=cdf_gain,
                                   show_progress_bar=show_progress_bar)

    def train_from_cov_matrix(self, cov_matrix: np.ndarray, min_weight: float = 0.0, max_weight: float = 1.0,
                              scale_by_variance: bool = False, apply_cdf: bool = False, cdf_gain: float = 1.0,
                              show_progress_bar: bool = False) -> None:
        """Trains the neural network from a covariance matrix precomputed with ``Covariance``."""
        if max_weight <= min_weight:
            raise ValueError('max_weight must be > min_weight.')
        if apply_cdf and cdf_gain <= 0.0:
            raise ValueError('cdf_gain must be > 0 when apply_cdf is True.')
        eigenvalues64, eigenvectors64, weights64 = SOM_PCA._diag_cov_matrix(
            cov_matrix.astype(np.float64, copy=False), min_weight, max_weight, self._m, self._n, self._dim,
            scale_by_variance=scale_by_variance, apply_cdf=apply_cdf, cdf_gain=cdf_gain)
        self._cov_matrix[:] = cov_matrix.astype(self._dtype, copy=False)
        self._eigenvalues[:] = eigenvalues64.astype(self._dtype, copy=False)
        self._eigenvectors[:] = eigenvectors64.astype(self._dtype, copy=False)
        self._weights[:] = weights64.astype(self._dtype, copy=False).reshape(self._weights.shape)
