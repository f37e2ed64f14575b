"""CPU: the HDF5 model format's host logic (SOM_Abstract.save/load, som_abstract.py:163-319).

h5py is not installed in this image, so these tests put an in-memory stand-in module (tests/h5stub.py)
under the name ``h5py`` for the duration of each test. It keeps what ``File``/``create_group``/
``attrs``/``create_dataset`` receive and returns numbers as NumPy scalars, as h5py reads attributes back. What
they check is this package's own logic: which header attributes and datasets are written, the
int/float conversion of attributes on load, the cast of datasets to the model's dtype, missing keys
left untouched, and the topography rebuilt from the loaded topology. The on-disk HDF5 layout itself
is not pinned here (no h5py to write or read a real file).
"""
import numpy as np
import pytest

import decontamination_amd as dam
from h5stub import fake_h5py  # noqa: F401  (pytest fixture)


def test_som_batch_round_trip(fake_h5py):
    som = dam.SOM_Batch(5, 6, 3, topology='square', sigma=2.5)
    som.init_rand(seed=1)
    som._n_epochs = 7
    som._quantization_errors = np.array([0.5, 0.25])
    som._topographic_errors = np.array([0.1, 0.05])
    som.save('model.h5')
    g = fake_h5py['model.h5']['model']
    assert set(g.attrs) == {'mode', 'sigma', 'n_epochs', 'n_vectors', 'm', 'n', 'dim', 'topology'}
    assert g.attrs['mode'] == 'batch' and g.attrs['topology'] == 'square'
    assert set(g.data) == {'weights', 'quantization_errors', 'topographic_errors'}

    other = dam.SOM_Batch(5, 6, 3)  # hexagonal, default sigma: load overrides both
    other.load('model.h5')
    assert other.topology == 'square' and other.sigma == 2.5 and isinstance(other.sigma, float)
    assert other._n_epochs == 7 and isinstance(other._n_epochs, int)
    assert other.m == 5 and other.n == 6 and other.dim == 3 and isinstance(other.m, int)
    assert np.array_equal(other.weights, som.weights) and other.weights.dtype == np.float32
    # every dataset is cast to the model's dtype on load, the error arrays too (som_abstract.py:300-310)
    assert np.array_equal(other.quantization_errors, np.float32([0.5, 0.25]))
    assert np.array_equal(other.topographic_errors, np.float32([0.1, 0.05]))
    assert np.array_equal(other._topography, dam.SOM_Batch(5, 6, 3, topology='square')._topography)


def test_som_pca_extras_and_dtype_cast(fake_h5py):
    som = dam.SOM_PCA(4, 4, 3, topology='hexagonal')
    som._scale_by_variance = True
    som._apply_cdf = True
    som._cdf_gain = 2.5
    som.init_rand(seed=2)
    som.save('pca.h5')
    # a file written with FP64 weights loads into the model's dtype
    fake_h5py['pca.h5']['model'].data['weights'] = som.weights.astype(np.float64) + 1e-12
    other = dam.SOM_PCA(4, 4, 3)
    other.load('pca.h5')
    assert bool(other._scale_by_variance) and bool(other._apply_cdf) and other._cdf_gain == 2.5
    assert other.weights.dtype == np.float32
    assert np.array_equal(other.weights, (som.weights.astype(np.float64) + 1e-12).astype(np.float32))


def test_missing_keys_are_left_untouched(fake_h5py):
    som = dam.SOM_Batch(3, 3, 2)
    som.init_rand(seed=3)
    som.save('partial.h5')
    g = fake_h5py['partial.h5']['model']
    del g.data['quantization_errors']
    del g.attrs['sigma']
    other = dam.SOM_Batch(3, 3, 2, sigma=9.0)
    other._quantization_errors = np.array([42.0])
    other.load('partial.h5')
    assert other.sigma == 9.0
    assert np.array_equal(other.quantization_errors, [42.0])
    assert np.array_equal(other.weights, som.weights)


def test_load_of_a_missing_file_raises(fake_h5py):
    with pytest.raises(FileNotFoundError):
        dam.SOM_Batch(3, 3, 2).load('absent.h5')


def test_decontamination_som_files_and_dataset_group(fake_h5py):
    """Decontamination_SOM.save writes the PCA beside the SOM (`_pca.hdf5`) and a `dataset` group with the
    reference's names (decontamination_som.py:451-489); generator builders are written as their arrays;
    load reads nside back as a scalar and re-derives the corrected density."""
    N, dim = 50, 3
    rng = np.random.default_rng(4)
    X = rng.standard_normal((N, dim)).astype(np.float32)
    gal = rng.random(N) * 4
    cov = rng.random(N) * 0.5 + 0.5
    chunks = lambda: lambda: (X[a:a + 20] for a in range(0, N, 20))  # noqa: E731
    d = dam.Decontamination_SOM(64, np.arange(N), cov, chunks, gal, 3, 4, dim, True)
    d.som.init_rand(seed=5)
    d.pca.init_rand(seed=6)
    d._winners = rng.integers(0, 12, N).astype(np.int32)
    d.save('run.hdf5')
    assert set(fake_h5py) == {'run.hdf5', 'run_pca.hdf5'}
    assert set(fake_h5py['run.hdf5']) == {'model', 'dataset'}
    g = fake_h5py['run.hdf5']['dataset']
    assert set(g.data) == {'nside', 'footprint', 'coverage', 'footprint_systematics', 'galaxy_number_density',
                           'winner'}
    assert np.array_equal(g.data['footprint_systematics'], X)
    assert fake_h5py['run_pca.hdf5']['model'].attrs['mode'] == 'pca'
    e = dam.Decontamination_SOM(1, np.zeros(1), np.ones(1), np.zeros((1, dim), np.float32), np.zeros(1), 3, 4, dim,
                                True)
    e.load('run.hdf5')
    assert e.nside == 64 and np.ndim(e.nside) == 0
    assert np.array_equal(e.footprint_systematics, X) and np.array_equal(e.winners, d.winners)
    assert np.array_equal(e.weights, d.weights) and np.array_equal(e.pca.weights, d.pca.weights)
    assert np.array_equal(e.corrected_galaxy_number_density, gal / cov)
    with pytest.raises(ValueError):
        fake_h5py['run.hdf5']['dataset']['nside'][:]  # a scalar dataset does not slice, as in h5py
