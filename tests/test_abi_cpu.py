"""CPU checks of the C ABI: the library loads and exports every entry point
that include/gempic_hip.h declares; no compute call is made here."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

import gempic_amd
from gempic_amd import _lib

HEADER = os.path.join(ROOT, "include", "gempic_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gempic_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_python_surface():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    L = _lib.load()
    assert L.gempic_version() >= 100


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_argument_errors_without_gpu_work():
    """Argument validation happens before any device call."""
    L = _lib.load()
    h = ctypes.c_void_p()
    rc = L.gempic_splitting_create(32, 0.0, 1.0, 4, 3, 0, 2, 0, 2, 100, 1, 1.0, 1.0, 0.0,
                                   ctypes.byref(h))
    assert rc == _lib.ERR_ARGUMENT
    assert b"degree" in L.gempic_last_error()
    rc = L.gempic_splitting_create(31, 0.0, 1.0, 3, 3, 0, 2, 0, 2, 100, 1, 1.0, 1.0, 0.0,
                                   ctypes.byref(h))
    assert rc == _lib.ERR_ARGUMENT
    with pytest.raises(ValueError):
        gempic_amd.ParticleMeshCoupling1D(gempic_amd.OneDGrid(0.0, 1.0, 8), 10, 3, "bogus")


def test_product_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "gempic.jl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                src = open(os.path.join(dirpath, f)).read()
                assert "pyoracle" not in src and "liboracle" not in src, f
                assert "gempic_oracle" not in src, f
