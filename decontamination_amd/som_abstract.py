"""SOM base class with the reference's API, backed by the HIP kernels of libsomhip.

Mirrors ``decontamination.algo.som_abstract.SOM_Abstract``
(decontamination/algo/som_abstract.py:23-695): same constructor, topography, ``init_rand`` /
``init_from``, HDF5 ``save`` / ``load`` layout, properties, ``get_distance_map``,
``compute_errors``, ``get_activation_map`` and ``get_winners`` with the same arguments, return
types and dtypes. Additive keyword-only options select the BMU search precision:

  precision  'bf16_rerank' (default) split-BF16 MFMA distance GEMM + exact rerank: the BMUs equal the
             reference formula's on every sample; 'fp32' plain FP32 MFMA GEMM; 'exact' direct scan.
  arithmetic 'ref32' (default) the reference's distance arithmetic (FP32 difference and square, FP64
             sum); 'fp64' FP64 throughout.

``dtype=np.float64`` keeps weights, topography, data and sums in FP64 on the device, as the
reference's float64 SOMs do; their search is the exact FP64 scan (precision 'exact', arithmetic
'fp64', the only values such a SOM takes). float32 and float64 are the dtypes the GPU path supports.

This package is the GPU path only. ``enable_gpu=False`` raises: the CPU path stays the reference's
own. ``threads_per_blocks`` is accepted for signature compatibility and ignored (launch shapes are
fixed by the kernels).
"""
import typing

import numpy as np

from . import _lib
from .runtime import DeviceArray, Model, Uploader, get_context


def dataset_to_generator_builder(dataset):
    """decontamination/algo/__init__.py:21-27: ndarray -> one chunk, callable builder unchanged."""
    if dataset is None:
        return None
    return dataset if callable(dataset) else (lambda: lambda: (dataset,))


def iterate_chunks(dataset, dataset_weights):
    """Yields aligned (vectors, weights-or-None) chunks following the reference's chunk protocol
    (som_batch.py:295-342, som_abstract.py:562-593)."""
    dgb = dataset_to_generator_builder(dataset)
    wgb = dataset_to_generator_builder(dataset_weights)
    if wgb is None:
        for v in dgb()():
            yield v, None
    else:
        for v, w in zip(dgb()(), wgb()()):
            if v.shape[0] != w.shape[0]:
                raise ValueError('`dataset` and `dataset_weights` chunks must be aligned')
            yield v, w


# noinspection PyPep8Naming
class SOM_Abstract(object):
    """Self Organizing Maps (abstract class), GPU implementation.

    Parameters are those of the reference (som_abstract.py:46): m, n, dim, dtype (np.float32),
    topology ('square' or 'hexagonal', None means hexagonal).
    """

    __MODE__ = 'abstract'

    def __init__(self, m: int, n: int, dim: int, dtype=np.float32, topology: typing.Optional[str] = None, *,
                 precision: typing.Optional[str] = None, arithmetic: typing.Optional[str] = None, device: int = 0):
        self._m = int(m)
        self._n = int(n)
        self._dim = int(dim)
        self._dtype = dtype.type if hasattr(dtype, 'type') else dtype
        self._topology = topology or 'hexagonal'
        self._rebuild_topography()
        self._weights = np.empty((self._m * self._n, self._dim), dtype=self._dtype)
        self._history = np.empty((0, self._m * self._n, self._dim), dtype=self._dtype)
        self._quantization_errors = np.empty(0, dtype=np.float64)
        self._topographic_errors = np.empty(0, dtype=np.float64)
        self._header_extra = {'mode': '__MODE__'}
        self._dataset_extra = {}
        f64 = self._f64()
        self._precision = self._check_precision(precision or ('exact' if f64 else 'bf16_rerank'))
        self._arithmetic = self._check_arithmetic(arithmetic or ('fp64' if f64 else 'ref32'))
        self._check_f64_search(self._precision, self._arithmetic)
        self._device = int(device)
        self._dev_model = None
        self._staging = None

    # -- options --------------------------------------------------------------------------------
    @staticmethod
    def _check_precision(p):
        if p not in _lib.PRECISIONS:
            raise ValueError(f'Invalid precision `{p}`, expected one of {sorted(_lib.PRECISIONS)}')
        return p

    @staticmethod
    def _check_arithmetic(a):
        if a not in _lib.ARITHMETICS:
            raise ValueError(f'Invalid arithmetic `{a}`, expected one of {sorted(_lib.ARITHMETICS)}')
        return a

    def _f64(self):
        """True for a float64 SOM."""
        try:
            return np.dtype(self._dtype) == np.float64
        except TypeError:
            return False

    def _check_f64_search(self, p, a):
        if self._f64() and (p != 'exact' or a != 'fp64'):
            raise ValueError('float64 SOMs search with the exact FP64 distance (precision `exact`, arithmetic '
                             f'`fp64`), not precision `{p}` with arithmetic `{a}`')

    def _prec_codes(self, precision=None, arithmetic=None):
        p = self._check_precision(precision or self._precision)
        a = self._check_arithmetic(arithmetic or self._arithmetic)
        self._check_f64_search(p, a)
        return _lib.PRECISIONS[p], _lib.ARITHMETICS[a]

    @property
    def precision(self) -> str:
        """BMU search precision ('bf16_rerank', 'fp32' or 'exact')."""
        return self._precision

    @property
    def arithmetic(self) -> str:
        """Exact distance arithmetic ('ref32' or 'fp64')."""
        return self._arithmetic

    # -- topography (som_abstract.py:81-111) ---------------------------------------------------
    @staticmethod
    def _neuron_locations_square(m: int, n: int):
        for i in range(m):
            for j in range(n):
                yield [i, j]

    @staticmethod
    def _neuron_locations_hexagonal(m: int, n: int):
        for i in range(m):
            for j in range(n):
                yield [i + (j & 1) * 0.5, j * 0.8660254037844386]

    def _rebuild_topography(self):
        if self._topology == 'square':
            pts = list(SOM_Abstract._neuron_locations_square(self._m, self._n))
        else:
            pts = list(SOM_Abstract._neuron_locations_hexagonal(self._m, self._n))
        self._topography = np.array(pts, dtype=self._dtype).reshape(self._m * self._n, 2)

    # -- initialisation (som_abstract.py:115-159) -----------------------------------------------
    def init_rand(self, seed: typing.Optional[int] = None) -> None:
        """Initializes the neural network randomly (uniform [0, 1), numpy default_rng(seed))."""
        rng = np.random.default_rng(seed=seed)
        self._weights[...] = rng.random((self._m * self._n, self._dim), dtype=self._dtype)

    def init_from(self, other: 'SOM_Abstract') -> None:
        """Initializes the neural network from another one (any object with the reference's fields)."""
        if self._m != other._m or self._n != other._n or self._dim != other._dim or self._dtype != other._dtype \
                or self._topology != other._topology:
            raise TypeError('Incompatible shapes, dtypes or topologies')
        self._weights[...] = other._weights

    # -- HDF5 model format (som_abstract.py:163-319) ---------------------------------------------
    def _init_hdf5_extra(self):
        header_extra = dict(self._header_extra)
        dataset_extra = dict(self._dataset_extra)
        header_extra['m'] = '_m'
        header_extra['n'] = '_n'
        header_extra['dim'] = '_dim'
        header_extra['topology'] = '_topology'
        dataset_extra['weights'] = '_weights'
        dataset_extra['quantization_errors'] = '_quantization_errors'
        dataset_extra['topographic_errors'] = '_topographic_errors'
        return header_extra, dataset_extra

    def save(self, filename: str) -> None:
        """Saves the trained neural network to a HDF5 file (group `model`, same attributes/datasets)."""
        import h5py
        header_extra, dataset_extra = self._init_hdf5_extra()
        with h5py.File(filename, mode='w') as file:
            model_group = file.create_group('model', track_order=True)
            for name, field in header_extra.items():
                data = getattr(self, field)
                if data is not None:
                    model_group.attrs[name] = data
            for name, field in dataset_extra.items():
                data = getattr(self, field)
                if data is not None:
                    model_group.create_dataset(name, data=data, shape=data.shape, dtype=data.dtype)
        self._rebuild_topography()

    def load(self, filename: str) -> None:
        """Loads the trained neural network from a HDF5 file."""
        import h5py
        header_extra, dataset_extra = self._init_hdf5_extra()
        with h5py.File(filename, mode='r') as file:
            model_group = file['model']
            for name, field in header_extra.items():
                try:
                    value = model_group.attrs[name]
                    if isinstance(value, (np.int32, np.int64)):
                        setattr(self, field, int(value))
                    elif isinstance(value, (np.float32, np.float64)):
                        setattr(self, field, float(value))
                    else:
                        setattr(self, field, value)
                except KeyError:
                    pass
            for name, field in dataset_extra.items():
                try:
                    setattr(self, field, np.array(model_group[name]).astype(self._dtype))
                except KeyError:
                    pass
        self._rebuild_topography()
        self._dev_model = None

    # -- properties (som_abstract.py:323-423) ---------------------------------------------------
    @property
    def m(self) -> int:
        """Number of neuron rows."""
        return self._m

    @property
    def n(self) -> int:
        """Number of neuron columns."""
        return self._n

    @property
    def dim(self) -> int:
        """Dimensionality of the input data."""
        return self._dim

    @property
    def dtype(self):
        """Neural network data type."""
        return self._dtype

    @property
    def topology(self) -> str:
        """Neural network topology, either 'square' or 'hexagonal'."""
        return self._topology

    @property
    def weights(self) -> np.ndarray:
        """Neural network weights with the shape [m*n, dim]."""
        return self._weights.reshape((self._m * self._n, self._dim))

    @property
    def centroids(self) -> np.ndarray:
        """Neural network weights with the shape [m, n, dim]."""
        return self._weights.reshape((self._m, self._n, self._dim))

    @property
    def quantization_errors(self) -> np.ndarray:
        """Quantization errors per epoch."""
        return self._quantization_errors

    @property
    def topographic_errors(self) -> np.ndarray:
        """Topographic errors per epoch."""
        return self._topographic_errors

    # -- distance map (som_abstract.py:427-524), host-side numpy -------------------------------
    _X_HEX_STENCIL = np.array([+1, +1, +1, +0, -1, +0, +0, +1, +0, -1, -1, -1], dtype=int)
    _Y_HEX_STENCIL = np.array([+1, +0, -1, -1, +0, +1, +1, +0, -1, -1, +0, +1], dtype=int)
    _X_SQU_STENCIL = np.array([+0, -1, -1, -1, +0, +1, +1, +1, +0, -1, -1, -1, +0, +1, +1, +1], dtype=int)
    _Y_SQU_STENCIL = np.array([-1, -1, +0, +1, +1, +1, +0, -1, -1, -1, +0, +1, +1, +1, +0, -1], dtype=int)

    @staticmethod
    def _distance_map(result, centroids, x_stencil, y_stencil, m, n, l):
        for x in range(m):
            for y in range(n):
                offset = (y & 1) * l
                w1 = centroids[x, y]
                for z in range(l):
                    i = x + x_stencil[z + offset]
                    j = y + y_stencil[z + offset]
                    if 0 <= i < m and 0 <= j < n:
                        w2 = centroids[i, j]
                        result[x, y, z] = np.sqrt(np.sum((w1 - w2) ** 2))

    def get_distance_map(self, normalization: typing.Optional[str] = None) -> np.ndarray:
        """Returns the distance map of the weights ('sum' or 'mean' normalization)."""
        normalization = normalization or 'sum'
        if self._topology == 'square':
            result = np.full((self._m, self._n, 8), np.nan, dtype=self._dtype)
            SOM_Abstract._distance_map(result, self.centroids, SOM_Abstract._X_SQU_STENCIL,
                                       SOM_Abstract._Y_SQU_STENCIL, self._m, self._n, 8)
        else:
            result = np.full((self._m, self._n, 6), np.nan, dtype=self._dtype)
            SOM_Abstract._distance_map(result, self.centroids, SOM_Abstract._X_HEX_STENCIL,
                                       SOM_Abstract._Y_HEX_STENCIL, self._m, self._n, 6)
        if normalization == 'sum':
            result = np.nansum(result, axis=2)
        elif normalization == 'mean':
            result = np.nanmean(result, axis=2)
        else:
            raise ValueError(f'Invalid normalization method `{normalization}`')
        result_max = np.nanmax(result)
        if np.isfinite(result_max) and result_max > 0.0:
            return (result / result_max).astype(self._dtype)
        raise ValueError('Distance map has no finite positive values.')

    # -- device plumbing ------------------------------------------------------------------------
    def _require_gpu(self, enable_gpu):
        if not enable_gpu:
            raise NotImplementedError('decontamination_amd implements the GPU path only; the CPU path is the '
                                      "reference's own (decontamination.SOM_*)")
        try:
            dt = np.dtype(self._dtype)
        except TypeError:
            dt = None
        if dt not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise NotImplementedError('the GPU path supports dtype=np.float32 and dtype=np.float64')

    def _ctx(self):
        return get_context(self._device)

    def _device_dtype(self):
        """Type of the device weights and history: float64 for float64 SOMs, else float32."""
        return np.float64 if self._f64() else np.float32

    def _device_model(self, upload=True):
        """Device model holding the current host weights."""
        if self._dev_model is None or self._dev_model.handle is None:
            self._dev_model = Model(self._ctx(), self._m, self._n, self._dim, self._topology, self._topography,
                                    dtype=self._device_dtype())
        if upload:
            self._dev_model.set_weights(self._weights)
        return self._dev_model

    def _stage(self):
        if self._staging is None:
            self._staging = Uploader(self._ctx(), slots=2)
        return self._staging

    def _upload_chunk(self, vectors, weights):
        """Device pointers of one chunk; its copy overlaps the kernels of the previous chunk."""
        x = np.ascontiguousarray(vectors, dtype=self._dtype)
        if x.ndim != 2 or x.shape[1] != self._dim:
            raise ValueError(f'dataset chunks must have shape (N, {self._dim})')
        r = None
        if weights is not None:
            r = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
            if r.shape[0] != x.shape[0]:
                raise ValueError('`dataset` and `dataset_weights` chunks must be aligned')
        xp, rp = self._stage().put(x, r)
        return xp, rp, x.shape[0]

    # -- errors (som_abstract.py:528-601) -------------------------------------------------------
    def compute_errors(self, dataset, dataset_weights=None, show_progress_bar: bool = False,
                       enable_gpu: bool = True, threads_per_blocks: typing.Optional[int] = None, *,
                       precision: typing.Optional[str] = None,
                       arithmetic: typing.Optional[str] = None) -> typing.Tuple[float, float]:
        """For the given input, computes the quantization and topographic errors."""
        if self._m * self._n < 2:
            return np.nan, np.nan
        self._require_gpu(enable_gpu)
        prec, arith = self._prec_codes(precision, arithmetic)
        model = self._device_model()
        ctx = self._ctx()
        result = DeviceArray.zeros(ctx, (3,), np.float64)
        chunks = iterate_chunks(dataset, dataset_weights)
        if show_progress_bar:
            import tqdm
            chunks = tqdm.tqdm(chunks)
        for v, w in chunks:
            xp, rp, n = self._upload_chunk(v, w)
            model.compute_errors(xp, rp, n, result.ptr, prec, arith)
        out = result.copy_to_host()
        return (out[0] / out[2], out[1] / out[2]) if out[2] > 0.0 else (np.nan, np.nan)

    # -- activation map (som_abstract.py:605-666) -----------------------------------------------
    def get_activation_map(self, dataset, dataset_weights=None, show_progress_bar: bool = False,
                           enable_gpu: bool = True, threads_per_blocks: typing.Optional[int] = None, *,
                           precision: typing.Optional[str] = None,
                           arithmetic: typing.Optional[str] = None) -> np.ndarray:
        """For the given input, returns the (m, n) matrix of (weighted) BMU counts."""
        self._require_gpu(enable_gpu)
        prec, arith = self._prec_codes(precision, arithmetic)
        model = self._device_model()
        ctx = self._ctx()
        result = DeviceArray.zeros(ctx, (self._m * self._n,), np.float64)
        chunks = iterate_chunks(dataset, dataset_weights)
        if show_progress_bar:
            import tqdm
            chunks = tqdm.tqdm(chunks)
        for v, w in chunks:
            xp, rp, n = self._upload_chunk(v, w)
            model.count_bmus(xp, rp, n, result.ptr, prec, arith)
        return result.copy_to_host().astype(self._dtype).reshape((self._m, self._n))

    # -- winners (som_abstract.py:670-695) ------------------------------------------------------
    def get_winners(self, dataset: np.ndarray, enable_gpu: bool = True,
                    threads_per_blocks: typing.Optional[int] = None, *, precision: typing.Optional[str] = None,
                    arithmetic: typing.Optional[str] = None) -> np.ndarray:
        """For the given input, returns the int32 vector of best matching unit indices."""
        self._require_gpu(enable_gpu)
        prec, arith = self._prec_codes(precision, arithmetic)
        model = self._device_model()
        xp, _, n = self._upload_chunk(dataset, None)
        out = DeviceArray.empty(self._ctx(), (n,), np.int32)
        model.find_bmus(xp, n, out.ptr, prec, arith)
        return out.copy_to_host()
