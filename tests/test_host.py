"""Host-side logic of the product classes (no GPU): constructor state, topography, initialisation,
distance map, argument validation and the chunk protocol."""
import json
import os

import numpy as np
import pytest

from decontamination_amd import SOM_Abstract, SOM_Batch
from decontamination_amd.som_abstract import iterate_chunks, dataset_to_generator_builder

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


def test_constructor_defaults_follow_reference():
    som = SOM_Batch(6, 8, 5)
    assert (som.m, som.n, som.dim) == (6, 8, 5)
    assert som.dtype is np.float32
    assert som.topology == "hexagonal"
    assert som.sigma == 4.0
    assert som.weights.shape == (48, 5) and som.centroids.shape == (6, 8, 5)
    assert som.quantization_errors.shape == (0,)
    assert SOM_Batch(3, 3, 2, sigma=1.5).sigma == 1.5
    assert som.precision == "bf16_rerank" and som.arithmetic == "ref32"


def test_float64_soms_search_exactly_in_fp64():
    """dtype=np.float64 (reference som_abstract.py:46): FP64 weights and topography on the host, and the
    exact FP64 search as the only precision; other dtypes are refused before any device call."""
    som = SOM_Batch(3, 4, 2, dtype=np.float64)
    assert som.dtype is np.float64 and som.weights.dtype == np.float64 and som._topography.dtype == np.float64
    assert som.precision == "exact" and som.arithmetic == "fp64"
    assert som._topography[3, 1] == 3 * 0.8660254037844386
    assert SOM_Batch(3, 4, 2, dtype=np.float64, precision="exact", arithmetic="fp64").precision == "exact"
    with pytest.raises(ValueError):
        SOM_Batch(3, 4, 2, dtype=np.float64, precision="bf16_rerank")
    with pytest.raises(ValueError):
        SOM_Batch(3, 4, 2, dtype=np.float64, arithmetic="ref32")
    with pytest.raises(ValueError):
        som.get_winners(np.zeros((2, 2)), precision="fp32")
    with pytest.raises(NotImplementedError):
        SOM_Batch(2, 2, 2, dtype=np.int32).get_winners(np.zeros((2, 2), np.int32))


def test_topography_closed_forms():
    sq = SOM_Batch(3, 4, 2, topology="square")._topography
    assert np.array_equal(sq, np.array([[i, j] for i in range(3) for j in range(4)], np.float32))
    hx = SOM_Batch(3, 4, 2)._topography
    for i in range(3):
        for j in range(4):
            assert hx[i * 4 + j, 0] == np.float32(i + 0.5 * (j % 2))
            assert hx[i * 4 + j, 1] == np.float32(j * np.sqrt(3.0) / 2.0)


def test_init_rand_and_init_from():
    a = SOM_Batch(4, 4, 4, topology="square")
    a.init_rand(seed=0)
    assert np.array_equal(a.weights, np.random.default_rng(0).random((1000, 4), np.float32)[:16])
    b = SOM_Batch(4, 4, 4, topology="square")
    b.init_from(a)
    assert np.array_equal(a.weights, b.weights)
    with pytest.raises(TypeError):
        SOM_Batch(4, 4, 4).init_from(a)  # topology differs


def test_golden_distance_map():
    som = SOM_Abstract(4, 4, 4, topology="square")
    som.init_rand(seed=0)
    assert np.allclose(som.get_distance_map(), np.array(GOLD["distance_map"], np.float32))
    with pytest.raises(ValueError):
        som.get_distance_map("median")


def test_cpu_path_is_not_provided():
    som = SOM_Batch(2, 2, 3)
    som.init_rand(0)
    with pytest.raises(NotImplementedError):
        som.get_winners(np.zeros((3, 3), np.float32), enable_gpu=False)
    with pytest.raises(NotImplementedError):
        som.train(np.zeros((3, 3), np.float32), n_epochs=1, enable_gpu=False)


def test_invalid_training_method_and_options():
    som = SOM_Batch(2, 2, 3)
    with pytest.raises(ValueError):
        som.train(np.zeros((3, 3), np.float32))
    with pytest.raises(ValueError):
        som.train(np.zeros((3, 3), np.float32), n_epochs=1, n_vectors=3)
    with pytest.raises(ValueError):
        SOM_Batch(2, 2, 3, precision="fp16")
    with pytest.raises(ValueError):
        SOM_Batch(2, 2, 3, arithmetic="fp128")


def test_chunk_protocol():
    x = np.arange(12, dtype=np.float32).reshape(6, 2)
    assert dataset_to_generator_builder(None) is None
    chunks = list(iterate_chunks(x, None))
    assert len(chunks) == 1 and chunks[0][1] is None
    gen = lambda: lambda: (x[i:i + 2] for i in range(0, 6, 2))
    wgen = lambda: lambda: (np.ones(2) for _ in range(3))
    assert len(list(iterate_chunks(gen, wgen))) == 3
    bad = lambda: lambda: (np.ones(3) for _ in range(3))
    with pytest.raises(ValueError):
        list(iterate_chunks(gen, bad))


def test_save_load_requires_h5py_like_reference(tmp_path):
    try:
        import h5py  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError):
            SOM_Batch(2, 2, 2).save(str(tmp_path / "m.h5"))
        return
    som = SOM_Batch(2, 3, 2, topology="square", sigma=2.5)
    som.init_rand(1)
    som.save(str(tmp_path / "m.h5"))
    other = SOM_Batch(2, 3, 2)
    other.load(str(tmp_path / "m.h5"))
    assert other.topology == "square" and other.sigma == 2.5
    assert np.array_equal(other.weights, som.weights)
