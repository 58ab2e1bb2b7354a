"""CPU-side checks of the C ABI: the library loads and exports every symbol
include/prnu_api.h declares (no compute call needs a GPU here)."""
import ctypes
import os

import pytest

from rocket_amd import _lib


def test_header_declares_symbols():
    names = _lib.header_symbols()
    assert "prnu_ctx_create" in names and "prnu_gram_ncc" in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libprnu_hip.so is not built (python -m rocket_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _lib.header_symbols() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_ctypes_signatures_cover_header():
    assert set(_lib.header_symbols()) == set(_lib._SIGNATURES)


def test_error_paths_without_gpu():
    lib = _lib.load()
    assert lib.prnu_api_version() == 1
    h = ctypes.c_void_p()
    # invalid sizes are rejected before any device call
    assert lib.prnu_ctx_create(1, 5, 0, ctypes.byref(h)) == _lib.EINVAL
    assert b"at least 2x2" in lib.prnu_last_error()
    assert lib.prnu_ctx_create(8, 8, 7, ctypes.byref(h)) == _lib.EINVAL
    assert lib.prnu_max_input_size(None) == -1
