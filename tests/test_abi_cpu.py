"""CPU checks of the C ABI library: it loads, and it exports every function
declared in include/microhh_amd.h.  No compute call is made without a GPU."""
import ctypes as C
import os

import pytest

from microhh_amd import abi


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    declared = abi.header_symbols()
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert sorted(abi.SIGNATURES) == abi.header_symbols()


def test_model_header_exports():
    from microhh_amd import model
    declared = abi.header_symbols("microhh_amd_model.h")
    assert sorted(model.MODEL_SIGNATURES) == declared
    L = model.lib()
    assert all(hasattr(L, n) for n in declared)


def test_model_create_rejects_null():
    from microhh_amd import model
    assert model.lib().mhh_model_create(None, None, None, None, None, None, None) == 1


def test_error_paths_without_gpu():
    L = abi.lib()
    # null arguments are rejected before any HIP call
    assert L.mhh_ctx_create(None, 0, None, None) == 1
    assert b"null" in L.mhh_last_error()
    assert L.mhh_advec_2i5_exec(None) == 1
    assert L.mhh_advec_2i5_set_fluxlimit(None, None, 0) == 1
    gd = abi.GridDesc(8, 8, 4, 3, 3, 1, 1.0, 1.0, 1.0)
    h = C.c_void_p()
    assert L.mhh_ctx_create(C.byref(gd), 0, None, C.byref(h)) == 1   # ktot < 6
    assert b"ktot" in L.mhh_last_error()
