"""libsomhip loads and exports every symbol include/somhip.h declares (no compute calls; no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from decontamination_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "somhip.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(somhip_[a-z0-9_]+)\s*\(", text)))


def test_library_built_and_loads():
    assert os.path.exists(_lib.LIB_PATH), "run python -m decontamination_amd.build"
    _lib.load()


def test_every_declared_symbol_is_exported_and_bound():
    L = ctypes.CDLL(_lib.LIB_PATH)
    declared = header_functions()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), f"{name} declared in somhip.h but not exported"
    assert sorted(_lib.exported_symbols()) == declared


def test_header_constants_match_binding():
    text = open(HEADER).read()
    consts = dict(re.findall(r"#define\s+(SOMHIP_[A-Z0-9_]+)\s+\(?(-?\d+)\)?", text))
    assert int(consts["SOMHIP_OK"]) == _lib.OK
    assert int(consts["SOMHIP_ERR_ARG"]) == _lib.ERR_ARG
    assert int(consts["SOMHIP_ERR_NODEVICE"]) == _lib.ERR_NODEVICE
    shift = re.search(r"#define\s+SOMHIP_NOALLOC_MAX_ROWS\s+\(1 << (\d+)\)", text)
    assert shift and (1 << int(shift.group(1))) == _lib.NOALLOC_MAX_ROWS
    assert int(consts["SOMHIP_PREC_EXACT"]) == _lib.PREC_EXACT
    assert int(consts["SOMHIP_PREC_FP32"]) == _lib.PREC_FP32
    assert int(consts["SOMHIP_PREC_BF16_RERANK"]) == _lib.PREC_BF16_RERANK
    assert int(consts["SOMHIP_RERANK_REF32"]) == _lib.RERANK_REF32
    assert int(consts["SOMHIP_RERANK_FP64"]) == _lib.RERANK_FP64
    assert int(consts["SOMHIP_TOPOLOGY_SQUARE"]) == _lib.TOPOLOGY_SQUARE
    assert int(consts["SOMHIP_TOPOLOGY_HEXAGONAL"]) == _lib.TOPOLOGY_HEXAGONAL


def test_version_string():
    assert b"gfx950" in _lib.load().somhip_version()


def test_null_arguments_are_rejected_without_a_device():
    L = _lib.load()
    assert L.somhip_ctx_create(0, None) == _lib.ERR_ARG
    assert L.somhip_device_count(None) == _lib.ERR_ARG
    assert L.somhip_find_bmus(None, None, 0, 0, 0, None) == _lib.ERR_ARG
    assert "null" in _lib.last_error()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device error path")
def test_no_device_fails_loudly():
    L = _lib.load()
    h = ctypes.c_void_p()
    rc = L.somhip_ctx_create(0, ctypes.byref(h))
    assert rc == _lib.ERR_NODEVICE
    with pytest.raises(_lib.SomHipError):
        from decontamination_amd import SOM_Batch
        som = SOM_Batch(2, 2, 3)
        som.init_rand(0)
        som.get_winners(np.zeros((4, 3), np.float32))
