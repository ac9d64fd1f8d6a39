"""The C-ABI library loads and exports every symbol include/fv3d.h declares.
No compute call is made, so this runs without a GPU."""
import ctypes
import os
import re

from fv3d import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "fv3d.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(fv3d_[a-z_0-9]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_all_symbols():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_binding_declares_signatures():
    L = _lib.load()
    for s in declared_symbols():
        assert getattr(L, s).restype is not None or s in ("fv3d_ctx_destroy",)
