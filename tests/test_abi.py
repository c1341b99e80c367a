"""CPU checks of the drop-in boundary: the HIP library loads and exports every
function include/thermofem_hip.h declares (no compute calls without a GPU)."""
import ctypes as C
import os

from thermofem_amd import _lib


def test_library_exports_every_header_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    names = _lib.header_functions()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    names = set(_lib.header_functions())
    assert names <= set(_lib._SIGS), sorted(names - set(_lib._SIGS))


def test_library_is_gfx950_code_object():
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_binding_loads_without_gpu():
    L = _lib.lib()
    assert L.tf_last_error() is not None
