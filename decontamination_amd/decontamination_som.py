"""Systematics decontamination with a batch SOM: the consumer of the SOM path.

Mirrors the batch-SOM part of ``decontamination.decontamination.Decontamination_SOM``
(decontamination/decontamination/decontamination_som.py:20-667 with the dataset plumbing of
decontamination_abstract.py:19-95): PCA initialisation, batch training, the footprint winners, and
``compute_gndm`` with the galaxy-number-density products (gnd, gndc, gndm, gndcm, visibility).

``compute_gndm`` is one device call (``somhip_gndm``): the reference searches the footprint's BMUs
twice more for its two activation maps, although they are the winners it already holds; here the
maps are exact fixed-point scatters of the winners, followed by the FP32 division and the gather
gnd[winners]. The iterative SOM (``batch=False``), clustering and catalog generation are not part of
this package.
"""
import ctypes
import typing

import numpy as np

from . import _lib
from .runtime import DeviceArray, get_context
from .som_abstract import iterate_chunks
from .som_batch import SOM_Batch
from .som_pca import SOM_PCA


def ds_divide(a, b):
    """dataset_ops.ds_divide (decontamination/decontamination/dataset_ops.py): elementwise a / b of two
    arrays, or a generator builder of the chunkwise quotients of two builders."""
    if isinstance(a, np.ndarray) and isinstance(b, np.ndarray):
        if a.shape[0] != b.shape[0]:
            raise ValueError('`a` and `b` chunks must be aligned')
        return np.divide(a, b)
    if callable(a) and callable(b):
        def builder():
            for ac, bc in zip(a(), b()):
                if ac.shape[0] != bc.shape[0]:
                    raise ValueError('`a` and `b` chunks must be aligned')
                yield np.divide(ac, bc)
        return builder
    raise ValueError('`a` and `b` must both be arrays or both be callable')


def _materialise(dataset):
    if dataset is None or isinstance(dataset, np.ndarray):
        return dataset
    return np.concatenate([np.asarray(c) for c, _ in iterate_chunks(dataset, None)])


def compute_gndm(ctx, winners: np.ndarray, corrected_density: np.ndarray, m: int, n: int, dtype=np.float32):
    """(catalog_activation_map, footprint_activation_map, n_gal, n_pix, gnd, gndm) from the footprint
    winners and the coverage-corrected galaxy number density, as Decontamination_SOM.compute_gndm
    produces them (decontamination_som.py:600-667)."""
    if dtype is not np.float32:
        raise NotImplementedError('the fused gndm kernel writes float32 maps; float64 SOMs take '
                                  'Decontamination_SOM.compute_gndm, which builds the two activation maps')
    mn = m * n
    w = np.ascontiguousarray(winners, dtype=np.int32).reshape(-1)
    rho = np.ascontiguousarray(corrected_density, dtype=np.float64).reshape(-1)
    if w.shape[0] != rho.shape[0]:
        raise ValueError('winners and galaxy number density must be aligned')
    N = w.shape[0]
    d_w = DeviceArray.from_host(ctx, w)
    d_rho = DeviceArray.from_host(ctx, rho)
    d_maps = DeviceArray.empty(ctx, (3, mn), np.float32)
    d_gndm = DeviceArray.empty(ctx, (max(N, 1),), np.float32)
    _lib.call("somhip_gndm", ctx.handle, ctypes.c_void_p(d_w.ptr), ctypes.c_void_p(d_rho.ptr), N, mn,
              ctypes.c_void_p(d_maps.view_ptr(0)), ctypes.c_void_p(d_maps.view_ptr(mn)),
              ctypes.c_void_p(d_maps.view_ptr(2 * mn)), ctypes.c_void_p(d_gndm.ptr))
    maps = d_maps.copy_to_host()
    gndm = d_gndm.copy_to_host()[:N]
    for a in (d_w, d_rho, d_maps, d_gndm):
        a.free()
    cat = maps[0].reshape(m, n)
    foot = maps[1].reshape(m, n)
    # np.sum of the FP32 maps, as the reference sums them on the host (decontamination_som.py:606-607)
    return cat, foot, np.sum(cat), np.sum(foot), maps[2].reshape(m, n), gndm


# noinspection PyPep8Naming
class Decontamination_SOM(object):
    """Systematics decontamination using the batch Self Organizing Map method (GPU)."""

    def __init__(self, nside: int, footprint: np.ndarray, coverage: np.ndarray,
                 footprint_systematics: typing.Union[np.ndarray, typing.Callable],
                 galaxy_number_density: typing.Union[np.ndarray, typing.Callable], m: int, n: int, dim: int,
                 batch: bool, dtype=np.float32, topology: typing.Optional[str] = 'hexagonal',
                 alpha: float = None, sigma: float = None, **som_options):
        if not batch:
            raise NotImplementedError('decontamination_amd implements the batch SOM (batch=True) only')
        self._nside = nside
        self._footprint = footprint
        self._coverage = coverage
        self._footprint_systematics = footprint_systematics
        self._galaxy_number_density = galaxy_number_density
        self._corrected_galaxy_number_density = ds_divide(galaxy_number_density, coverage)
        self._pca = SOM_PCA(m, n, dim, dtype=dtype, topology=topology, **som_options)
        self._batch = batch
        self._som = SOM_Batch(m, n, dim, dtype=dtype, topology=topology, sigma=sigma, **som_options)
        self._winners = None
        self._cluster_ids = None
        self._catalog_activation_map = None
        self._footprint_activation_map = None
        self._n_gal = None
        self._n_pix = None
        self._gnd = None
        self._gndm = None

    # -- dataset (decontamination_abstract.py:19-95) --------------------------------------------
    @property
    def nside(self) -> int:
        return self._nside

    @property
    def footprint(self):
        return self._footprint

    @property
    def coverage(self):
        return self._coverage

    @property
    def footprint_systematics(self):
        return self._footprint_systematics

    @property
    def galaxy_number_density(self):
        return self._galaxy_number_density

    @property
    def corrected_galaxy_number_density(self):
        return self._corrected_galaxy_number_density

    # -- models (decontamination_som.py:112-240) ------------------------------------------------
    @property
    def pca(self) -> SOM_PCA:
        return self._pca

    @property
    def som(self) -> SOM_Batch:
        return self._som

    @property
    def m(self) -> int:
        return self._som.m

    @property
    def n(self) -> int:
        return self._som.n

    @property
    def dim(self) -> int:
        return self._som.dim

    @property
    def batch(self) -> bool:
        return self._batch

    @property
    def dtype(self):
        return self._som.dtype

    @property
    def topology(self) -> str:
        return self._som.topology

    @property
    def sigma(self) -> float:
        return self._som.sigma

    @property
    def quantization_errors(self) -> np.ndarray:
        return self._som.quantization_errors

    @property
    def topographic_errors(self) -> np.ndarray:
        return self._som.topographic_errors

    @property
    def weights(self) -> np.ndarray:
        return self._som.weights

    @property
    def centroids(self) -> np.ndarray:
        return self._som.centroids

    @property
    def winners(self) -> np.ndarray:
        """Winners (best matching units) of the footprint pixels."""
        return self._winners

    @property
    def cluster_ids(self):
        return self._cluster_ids

    # -- galaxy number density (decontamination_som.py:258-420) --------------------------------
    @property
    def catalog_activation_map(self) -> np.ndarray:
        return self._catalog_activation_map

    @property
    def footprint_activation_map(self) -> np.ndarray:
        return self._footprint_activation_map

    @property
    def n_gal(self):
        return self._n_gal

    @property
    def n_pix(self):
        return self._n_pix

    @property
    def gnd(self) -> np.ndarray:
        """Galaxy number density per neuron."""
        return self._gnd

    @property
    def gndc(self) -> np.ndarray:
        """Galaxy number density contrast per neuron, (gnd - n_gal / n_pix) / (n_gal / n_pix)."""
        return (self._gnd - (self._n_gal / self._n_pix)) / (self._n_gal / self._n_pix)

    @property
    def gndm(self) -> np.ndarray:
        """Galaxy number density per footprint pixel, gnd[winners]."""
        return self._gndm

    @property
    def gndcm(self) -> np.ndarray:
        """Galaxy number density contrast per footprint pixel."""
        return (self._gndm - (self._n_gal / self._n_pix)) / (self._n_gal / self._n_pix)

    @property
    def visibility(self) -> np.ndarray:
        """(1 + gndcm) x coverage."""
        return (1.0 + self.gndcm) * self._coverage

    # -- training (decontamination_som.py:542-596) ----------------------------------------------
    def train(self, cov_matrix: typing.Optional[np.ndarray] = None, n_epochs: typing.Optional[int] = None,
              n_vectors: typing.Optional[int] = None, use_best_epoch: bool = True,
              stop_quantization_error: typing.Optional[float] = None,
              stop_topographic_error: typing.Optional[float] = None, show_progress_bar: bool = True,
              enable_gpu: bool = True, threads_per_blocks: typing.Optional[int] = None) -> None:
        """PCA initialisation (from ``cov_matrix`` or from the footprint weighted by the corrected
        density), batch SOM training, then the winners of the footprint."""
        if cov_matrix is not None:
            self._pca.train_from_cov_matrix(cov_matrix, min_weight=0.0, max_weight=1.0)
        else:
            self._pca.train(self._footprint_systematics, dataset_weights=self._corrected_galaxy_number_density,
                            min_weight=0.0, max_weight=1.0)
        self._som.init_from(self._pca)
        self._som.train(self._footprint_systematics, dataset_weights=self._corrected_galaxy_number_density,
                        n_epochs=n_epochs, n_vectors=n_vectors, use_best_epoch=use_best_epoch,
                        stop_quantization_error=stop_quantization_error,
                        stop_topographic_error=stop_topographic_error, show_progress_bar=show_progress_bar,
                        enable_gpu=enable_gpu, threads_per_blocks=threads_per_blocks)
        self._winners = self._som.get_winners(self._footprint_systematics, enable_gpu=enable_gpu,
                                              threads_per_blocks=threads_per_blocks)

    # -- galaxy number density map (decontamination_som.py:600-667) ------------------------------
    def _compute_gndm(self, catalog_activation_map: np.ndarray, footprint_activation_map: np.ndarray):
        """gnd = catalog / footprint where footprint > 0 (NaN elsewhere) and gndm = gnd[winners], from
        given activation maps (host, as the reference)."""
        self._n_gal = np.sum(catalog_activation_map)
        self._n_pix = np.sum(footprint_activation_map)
        gnd = np.divide(catalog_activation_map, footprint_activation_map,
                        out=np.full(catalog_activation_map.shape, np.nan, dtype=self._som.dtype),
                        where=footprint_activation_map > 0.0)
        gndm = gnd.reshape(self._som.m * self._som.n)[self._winners]
        return gnd, gndm

    def compute_gndm(self, show_progress_bar: bool = True, enable_gpu: bool = True,
                     threads_per_blocks: typing.Optional[int] = None) -> None:
        """Both activation maps of the footprint, gnd and gndm, in one device call from the winners."""
        if self._footprint_systematics is None or self._galaxy_number_density is None:
            raise ValueError('Underlying SOM network not trained')
        self._som._require_gpu(enable_gpu)
        if self._winners is None:
            self._winners = self._som.get_winners(_materialise(self._footprint_systematics), enable_gpu=enable_gpu)
        if self._som._f64():
            # float64 maps: the reference's own sequence, two activation maps on the device, then gnd and
            # gndm from them (decontamination_som.py:657-667)
            self._catalog_activation_map = self._som.get_activation_map(
                self._footprint_systematics, dataset_weights=self._corrected_galaxy_number_density,
                enable_gpu=enable_gpu)
            self._footprint_activation_map = self._som.get_activation_map(self._footprint_systematics,
                                                                          enable_gpu=enable_gpu)
            self._gnd, self._gndm = self._compute_gndm(self._catalog_activation_map, self._footprint_activation_map)
            return
        (self._catalog_activation_map, self._footprint_activation_map, self._n_gal, self._n_pix, self._gnd,
         self._gndm) = compute_gndm(get_context(self._som._device), self._winners,
                                    _materialise(self._corrected_galaxy_number_density), self._som.m, self._som.n,
                                    self._som.dtype)

    # -- HDF5 model format (decontamination_som.py:451-537) --------------------------------------
    def save(self, filename: str) -> None:
        """Saves the trained decontamination model: the PCA to ``<name>_pca.hdf5``, the SOM's ``model``
        group to ``filename``, then a ``dataset`` group with nside, footprint, coverage,
        footprint_systematics, galaxy_number_density and the footprint winners (``winner``). Datasets
        given as generator builders are written as the arrays their chunks concatenate to."""
        import h5py
        self._pca.save(filename.replace('.hdf5', '_pca.hdf5'))
        self._som.save(filename)
        fields = (('nside', self._nside), ('footprint', self._footprint), ('coverage', self._coverage),
                  ('footprint_systematics', _materialise(self._footprint_systematics)),
                  ('galaxy_number_density', _materialise(self._galaxy_number_density)),
                  ('winner', self._winners))
        with h5py.File(filename, 'r+') as file:
            group = file.create_group('dataset')
            for name, data in fields:
                group.create_dataset(name, data=data)
        self._corrected_galaxy_number_density = ds_divide(_materialise(self._galaxy_number_density),
                                                          np.asarray(self._coverage))

    def load(self, filename: str) -> None:
        """Loads a model written by ``save``: the PCA, the SOM and the ``dataset`` group. nside is read
        as the scalar it was written as (the arrays with ``[:]``)."""
        import h5py
        self._pca.load(filename.replace('.hdf5', '_pca.hdf5'))
        self._som.load(filename)
        with h5py.File(filename, 'r') as file:
            group = file['dataset']
            self._nside = group['nside'][()]
            self._footprint = group['footprint'][:]
            self._coverage = group['coverage'][:]
            self._footprint_systematics = group['footprint_systematics'][:]
            self._galaxy_number_density = group['galaxy_number_density'][:]
            self._winners = group['winner'][:]
        self._corrected_galaxy_number_density = ds_divide(self._galaxy_number_density, self._coverage)

    def compute_clustered_gndm(self, *args, **kwargs):
        raise NotImplementedError('neuron clustering is not part of decontamination_amd')
