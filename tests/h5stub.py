"""In-memory stand-in for the parts of h5py the HDF5 model format uses (h5py is not installed here).

It keeps what ``File``/``create_group``/``attrs``/``create_dataset`` receive, returns attributes as
NumPy scalars as h5py reads them back, and hands datasets out as objects that slice like h5py datasets:
``[:]`` on an array, ``[()]`` on a scalar (``[:]`` on a scalar dataset raises, as h5py does). Files
opened with mode 'r' or 'r+' must exist. Install it with the ``fake_h5py`` fixture of a test module::

    from h5stub import fake_h5py  # noqa: F401  (pytest fixture)
"""
import sys
import types

import numpy as np
import pytest


class Dataset:
    def __init__(self, arr):
        self.arr = arr

    @property
    def shape(self):
        return self.arr.shape

    @property
    def dtype(self):
        return self.arr.dtype

    def __getitem__(self, key):
        if self.arr.ndim == 0 and not (key == () or key is Ellipsis):
            raise ValueError('Illegal slicing argument for scalar dataspace')
        return np.array(self.arr[key])

    def __array__(self, dtype=None, copy=None):
        return np.array(self.arr, dtype=dtype)


class Group:
    def __init__(self):
        self.attrs = Attrs()
        self.data = {}

    def create_dataset(self, name, data, shape=None, dtype=None):
        if name in self.data:
            raise ValueError(f'Unable to create dataset (name already exists): {name}')
        arr = np.array(data, dtype=dtype)
        self.data[name] = arr.reshape(shape) if shape is not None else arr

    def __getitem__(self, name):
        return Dataset(self.data[name])


class Attrs(dict):
    def __setitem__(self, key, value):
        # h5py stores Python scalars as NumPy scalars and returns them as such; strings stay str
        super().__setitem__(key, value if isinstance(value, str) else np.asarray(value)[()])


class File:
    files = {}

    def __init__(self, filename, mode):
        self.filename = filename
        if mode == 'w':
            File.files[filename] = {}
        elif mode in ('r', 'r+'):
            if filename not in File.files:
                raise FileNotFoundError(filename)
        else:
            raise ValueError(f'unsupported mode {mode}')
        self.mode = mode

    def create_group(self, name, track_order=False):
        if self.mode == 'r':
            raise ValueError('file opened read-only')
        if name in File.files[self.filename]:
            raise ValueError(f'Unable to create group (name already exists): {name}')
        g = Group()
        File.files[self.filename][name] = g
        return g

    def __getitem__(self, name):
        return File.files[self.filename][name]

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


@pytest.fixture
def fake_h5py(monkeypatch):
    mod = types.ModuleType('h5py')
    mod.File = File
    File.files = {}
    monkeypatch.setitem(sys.modules, 'h5py', mod)
    return File.files
