"""CPU: the HDF5 model format's host logic (SOM_Abstract.save/load, som_abstract.py:163-319).

h5py is not installed in this image, so these tests put an in-memory stand-in module under the name
``h5py`` for the duration of each test. It keeps what ``File``/``create_group``/``attrs``/
``create_dataset`` receive and returns numbers as NumPy scalars, as h5py reads attributes back. What
they check is this package's own logic: which header attributes and datasets are written, the
int/float conversion of attributes on load, the cast of datasets to the model's dtype, missing keys
left untouched, and the topography rebuilt from the loaded topology. The on-disk HDF5 layout itself
is not pinned here (no h5py to write or read a real file).
"""
import sys
import types

import numpy as np
import pytest

import decontamination_amd as dam


class _Group:
    def __init__(self):
        self.attrs = _Attrs()
        self.data = {}

    def create_dataset(self, name, data, shape, dtype):
        self.data[name] = np.array(data, dtype=dtype).reshape(shape)

    def __getitem__(self, name):
        return self.data[name]


class _Attrs(dict):
    def __setitem__(self, key, value):
        # h5py stores Python scalars as NumPy scalars and returns them as such; strings stay str
        super().__setitem__(key, value if isinstance(value, str) else np.asarray(value)[()])


class _File:
    files = {}

    def __init__(self, filename, mode):
        self.filename = filename
        if mode == 'w':
            _File.files[filename] = {}
        elif filename not in _File.files:
            raise FileNotFoundError(filename)

    def create_group(self, name, track_order=False):
        g = _Group()
        _File.files[self.filename][name] = g
        return g

    def __getitem__(self, name):
        return _File.files[self.filename][name]

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


@pytest.fixture
def fake_h5py(monkeypatch):
    mod = types.ModuleType('h5py')
    mod.File = _File
    _File.files = {}
    monkeypatch.setitem(sys.modules, 'h5py', mod)
    return _File.files


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
