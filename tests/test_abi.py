"""CPU checks of the C-ABI library: it loads and exports every entry point
declared in include/axsq.h (no compute calls: there is no GPU here)."""
import ctypes
import os

import pytest

from axom_amd import _lib


def test_library_exports_every_declared_symbol():
    syms = _lib.header_symbols()
    assert len(syms) >= 20
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_lib._SIGS) == _lib.header_symbols()


def test_error_path_without_gpu_work():
    L = _lib.lib()
    assert L.axsq_version() >= 100
    h = ctypes.c_void_p()
    # invalid fp_mode is rejected before any device work
    rc = L.axsq_bvh_build(ctypes.byref(h), None, 10, 1.0, 7, None)
    assert rc == -1
    assert b"fp_mode" in L.axsq_last_error()
    # NULL handles are rejected
    assert L.axsq_sdf_eval(None, None, None, None, 1, 0, 0, None, None, None, None, None) == -5
