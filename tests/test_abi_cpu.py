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


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_2d_argument_errors_without_gpu_work():
    """The 2D entry points validate their arguments before any device call."""
    L = _lib.load()
    m = ctypes.c_void_p()
    assert L.gempic_maxwell2d_create(16, 0.0, 1.0, 16, 0.0, 1.0, 4, ctypes.byref(m)) == \
        _lib.ERR_ARGUMENT
    assert L.gempic_maxwell2d_create(6, 0.0, 1.0, 16, 0.0, 1.0, 3, ctypes.byref(m)) == \
        _lib.ERR_ARGUMENT
    assert L.gempic_maxwell2d_create(300, 0.0, 1.0, 16, 0.0, 1.0, 3, ctypes.byref(m)) == \
        _lib.ERR_ARGUMENT
    import numpy as np
    x = np.zeros(4)
    rho = np.zeros(100 * 100)
    rc = L.gempic_pmc2d_add_charge(100, 100, 0.0, 1.0, 0.0, 1.0, 3, 0, 4, _lib.ptr(x),
                                   _lib.ptr(x), _lib.ptr(x), _lib.ptr(rho))
    assert rc == _lib.ERR_ARGUMENT and b"4096" in L.gempic_last_error()
    rc = L.gempic_pmc2d_evaluate(8, 8, 0.0, 1.0, 0.0, 1.0, 4, 0, 4, _lib.ptr(x), _lib.ptr(x),
                                 _lib.ptr(rho), _lib.ptr(x))
    assert rc == _lib.ERR_ARGUMENT
    with pytest.raises(ValueError):
        gempic_amd.ParticleMeshCoupling2D(gempic_amd.ParticleGroup(2, 2, 4),
                                          gempic_amd.TwoDGrid(0.0, 1.0, 8, 0.0, 1.0, 8), 3, "bogus")
    with pytest.raises(TypeError):
        gempic_amd.ParticleMeshCoupling2D(gempic_amd.ParticleGroup(1, 2, 4),
                                          gempic_amd.TwoDGrid(0.0, 1.0, 8, 0.0, 1.0, 8), 3,
                                          "galerkin")


def test_2d_host_setup_matches_the_oracle():
    """The host set-up of the 2D solver (mass lines, eigenvalues, quadrature
    right-hand sides of user functions) against the oracle, bit for bit."""
    import math
    import types
    import numpy as np
    import pyoracle as orc
    from gempic_amd import core2d
    for deg in (0, 1, 2, 3):
        np.testing.assert_array_equal(core2d.spline_fem_mass_line(deg),
                                      orc.spline_fem_mass_line(deg))
    line = core2d.spline_fem_mass_line(3) * 0.37
    np.testing.assert_array_equal(core2d.spline_fem_compute_mass_eig(13, 3, line),
                                  orc.spline_fem_compute_mass_eig(13, 3, line))
    grid = gempic_amd.TwoDGrid(0.0, 2 * math.pi, 6, 0.0, 3.0, 5)
    o = orc.Maxwell2D(0.0, 2 * math.pi, 6, 0.0, 3.0, 5, 3)
    stub = types.SimpleNamespace(s_deg_0=3, s_deg_1=2, mesh=grid)
    f = lambda x, y: math.cos(x) * math.sin(2.0 * y) + 0.25 * x  # noqa: E731
    for form, comp in ((0, 1), (1, 1), (1, 2), (1, 3), (2, 1), (2, 2), (2, 3), (3, 1)):
        got = core2d.compute_fem_rhs_(np.zeros(30), stub, f, comp, form)
        np.testing.assert_array_equal(got, o.fem_rhs(f, comp, form))
    g = gempic_amd.TwoDGrid(2.0, 4, 3.0, 6)
    assert (g.xmin, g.xmax, g.ymax, g.dx, g.dy) == (0.0, 2.0, 3.0, 0.5, 0.5)
