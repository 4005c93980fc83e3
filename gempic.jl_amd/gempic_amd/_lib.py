"""Loader and ctypes signatures of libgempic_hip.so (include/gempic_hip.h).

The library is the product: every compute call below runs HIP kernels on the
GPU. There is no CPU fallback; a missing library or GPU raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libgempic_hip.so")

OK = 0
ERR_ARGUMENT = 1
ERR_ASSERTION = 2
ERR_DEVICE = 3
ERR_INEXACT = 4
ERR_STATE = 5

GALERKIN = 0
COLLOCATION = 1
STRANG_FUSED = 0
STRANG_SUBSTEP = 1
ASYNC = 16
GRAPH = 64          # GEMPIC_GRAPH: an ensemble's regular steps as a HIP graph
DETERMINISTIC = 32

_D = ctypes.c_double
_I64 = ctypes.c_int64
_I = ctypes.c_int
_P = ctypes.POINTER(ctypes.c_double)
_PI64 = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); the exported surface of include/gempic_hip.h
SIGNATURES = {
    "gempic_version": (_I, []),
    "gempic_last_error": (ctypes.c_char_p, []),
    "gempic_device_count": (_I, []),
    "gempic_splitting_create": (_I, [_I64, _D, _D, _I, _I, _I, _I, _I, _I, _I64, _I, _D, _D, _D, _PP]),
    "gempic_splitting_destroy": (_I, [_VP]),
    "gempic_set_particles": (_I, [_VP, _P, _I64, _I64]),
    "gempic_get_particles": (_I, [_VP, _P, _I64, _I64]),
    "gempic_set_fields": (_I, [_VP, _P, _P, _P]),
    "gempic_get_fields": (_I, [_VP, _P, _P, _P]),
    "gempic_get_j": (_I, [_VP, _P, _P]),
    "gempic_operator_hb": (_I, [_VP, _D]),
    "gempic_operator_he": (_I, [_VP, _D]),
    "gempic_operator_hp1": (_I, [_VP, _D]),
    "gempic_operator_hp2": (_I, [_VP, _D]),
    "gempic_strang_splitting": (_I, [_VP, _D, _I64, _I]),
    "gempic_synchronize": (_I, [_VP]),
    "gempic_strang_splitting_adjoint": (_I, [_VP, _D, _I64, _I64, _D, _P, _I64, _P, _P, _P, _I]),
    "gempic_get_adjoint": (_I, [_VP, _P, _P, _I64, _I64, _P, _P, _P]),
    "gempic_solve_poisson_adjoint": (_I, [_VP, _P, _P]),
    "gempic_get_adjoint_domain": (_I, [_VP, _P, _P]),
    "gempic_ensemble_create": (_I, [_I, ctypes.POINTER(_VP), ctypes.POINTER(_VP)]),
    "gempic_ensemble_destroy": (_I, [_VP]),
    "gempic_ensemble_strang_splitting": (_I, [_VP, _P, _I64, _I]),
    "gempic_ensemble_autotune": (_I, [_VP, _P, _I64, ctypes.POINTER(_I), _I,
                                      ctypes.POINTER(_I), _P]),
    "gempic_ensemble_strang_splitting_adjoint": (_I, [_VP, _P, _I64, _I64, _P, _I]),
    "gempic_set_sort_interval": (_I, [_VP, _I64]),
    "gempic_sort_particles": (_I, [_VP]),
    "gempic_solve_poisson": (_I, [_VP, _P, _P]),
    "gempic_write_step": (_I, [_VP, _D, _I, _P, _P, _P, _P]),
    "gempic_profile_enable": (_I, [_VP, _I]),
    "gempic_profile_read": (_I, [_VP, _I, _P, _PI64]),
    "gempic_profile_reset": (_I, [_VP]),
    "gempic_set_launch": (_I, [_VP, _I]),
    "gempic_autotune": (_I, [_VP, _D, _I64, ctypes.POINTER(_I), _I, _PI64, _I,
                             ctypes.POINTER(_I), _PI64, _P]),
    "gempic_set_fusion": (_I, [_VP, _I]),
    "gempic_sample_landau": (_I, [_I, _I64, _D, _D, _D, _D, _P, _I64]),
    "gempic_sample_landau_particles": (_I, [_VP, _D, _D, _D]),
    "gempic_boris_staggering": (_I, [_VP, _D]),
    "gempic_boris_strang_splitting": (_I, [_VP, _D, _I64, _I]),
    "gempic_boris_push_v_epart": (_I, [_VP, _D]),
    "gempic_boris_push_v_bpart": (_I, [_VP, _D]),
    "gempic_boris_push_x_accumulate_j": (_I, [_VP, _D]),
    "gempic_boris_set_mid": (_I, [_VP, _P, _P, _P]),
    "gempic_boris_get_mid": (_I, [_VP, _P, _P, _P]),
    "gempic_pmc1d_add_charge": (_I, [_I64, _D, _D, _I, _I, _I64, _P, _P, _P]),
    "gempic_pmc1d_add_charge_pp": (_I, [_I64, _D, _D, _I, _I, _I64, _P, _P, _P]),
    "gempic_pmc1d_evaluate": (_I, [_I64, _D, _D, _I, _I, _I64, _P, _P, _P]),
    "gempic_pmc1d_evaluate_pp": (_I, [_I64, _D, _D, _I, _I, _I64, _P, _P, _P]),
    "gempic_pmc1d_add_current_update_v": (_I, [_I64, _D, _D, _I, _I, _I64, _P, _P, _P, _D, _P, _P, _P]),
    "gempic_pmc1d_add_current_update_v_pp": (_I, [_I64, _D, _D, _I, _I, _I64, _P, _P, _P, _D, _P, _P, _P]),
    "gempic_b_to_pp": (_I, [_I, _I64, _P, _P]),
    "gempic_maxwell1d_create": (_I, [_I64, _D, _D, _I, _PP]),
    "gempic_maxwell1d_destroy": (_I, [_VP]),
    "gempic_maxwell1d_tables": (_I, [_VP, _P, _P, _P, _P]),
    "gempic_maxwell1d_compute_e_from_rho": (_I, [_VP, _P, _P]),
    "gempic_maxwell1d_compute_e_from_j": (_I, [_VP, _P, _P, _I]),
    "gempic_maxwell1d_compute_e_from_b": (_I, [_VP, _P, _D, _P]),
    "gempic_maxwell1d_compute_b_from_e": (_I, [_VP, _P, _D, _P]),
    "gempic_maxwell1d_inner_product": (_I, [_VP, _P, _P, _I, _P]),
    "gempic_maxwell1d_l2norm_squared": (_I, [_VP, _P, _I, _P]),
    "gempic_pmc2d_add_charge": (_I, [_I64, _I64, _D, _D, _D, _D, _I, _I, _I64, _P, _P, _P, _P]),
    "gempic_pmc2d_evaluate": (_I, [_I64, _I64, _D, _D, _D, _D, _I, _I, _I64, _P, _P, _P, _P]),
    "gempic_pmc2d_evaluate_multiple": (_I, [_I64, _I64, _D, _D, _D, _D, _I, _I, _I64, _P, _P,
                                            _P, _P, _P, _P]),
    "gempic_pmc2d_add_charge_pp": (_I, [_I64, _I64, _D, _D, _D, _D, _I, _I, _I64, _P, _P, _P, _P]),
    "gempic_pmc2d_evaluate_pp": (_I, [_I64, _I64, _D, _D, _D, _D, _I, _I, _I64, _P, _P, _P, _P]),
    "gempic_b_to_pp_2d": (_I, [_I, _I64, _I64, _P, _P]),
    "gempic_pmc2d_add_charge_dev": (_I, [_I64, _I64, _D, _D, _D, _D, _I, _I, _I64, _VP, _VP, _VP,
                                         _VP, _I, _VP]),
    "gempic_pmc2d_evaluate_dev": (_I, [_I64, _I64, _D, _D, _D, _D, _I, _I, _I64, _VP, _VP, _VP,
                                       _VP, _I, _VP]),
    "gempic_pmc2d_evaluate_multiple_dev": (_I, [_I64, _I64, _D, _D, _D, _D, _I, _I, _I64, _VP, _VP,
                                                _VP, _VP, _VP, _VP, _I, _VP]),
    "gempic_set_arithmetic": (_I, [_VP, _I]),
    "gempic_device_info": (_I, [ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, _I]),
    "gempic_splitting2d_create": (_I, [_I64, _I64, _D, _D, _D, _D, _I, _I64, _D, _D, _D, _PP]),
    "gempic_splitting2d_destroy": (_I, [_VP]),
    "gempic_splitting2d_set_particles": (_I, [_VP, _P, _I64, _I64]),
    "gempic_splitting2d_get_particles": (_I, [_VP, _P, _I64, _I64]),
    "gempic_splitting2d_set_fields": (_I, [_VP, _P, _P, _P, _P, _P, _P]),
    "gempic_splitting2d_get_fields": (_I, [_VP, _P, _P, _P, _P, _P, _P]),
    "gempic_splitting2d_get_j": (_I, [_VP, _P, _P, _P]),
    "gempic_splitting2d_operator_hb": (_I, [_VP, _D]),
    "gempic_splitting2d_operator_he": (_I, [_VP, _D]),
    "gempic_splitting2d_operator_hp1": (_I, [_VP, _D]),
    "gempic_splitting2d_operator_hp2": (_I, [_VP, _D]),
    "gempic_splitting2d_operator_hp3": (_I, [_VP, _D]),
    "gempic_splitting2d_strang_splitting": (_I, [_VP, _D, _I64, _I]),
    "gempic_splitting2d_synchronize": (_I, [_VP]),
    "gempic_splitting2d_set_sort_interval": (_I, [_VP, _I64]),
    "gempic_splitting2d_rho": (_I, [_VP, _P]),
    "gempic_splitting2d_rho_from_e": (_I, [_VP, _P]),
    "gempic_maxwell1d_compute_rhs_from_function": (_I, [_VP, _VP, _VP, _I, _P]),
    "gempic_maxwell1d_l2projection": (_I, [_VP, _VP, _VP, _I, _P]),
    "gempic_pmc2d_sync": (_I, [_VP]),
    "gempic_pmc2d_add_charge_dev_noalloc": (_I, [_I64, _I64, _D, _D, _D, _D, _I, _I, _I64, _VP,
                                                 _VP, _VP, _VP, _I, _VP]),
    "gempic_maxwell2d_create": (_I, [_I64, _D, _D, _I64, _D, _D, _I, _PP]),
    "gempic_maxwell2d_destroy": (_I, [_VP]),
    "gempic_maxwell2d_tables": (_I, [_VP, _P, _P, _P, _P]),
    "gempic_maxwell2d_solve": (_I, [_VP, _I, _I, _P, _P]),
    "gempic_maxwell2d_multiply_mass": (_I, [_VP, _I, _I, _P, _P]),
    "gempic_maxwell2d_inner_product": (_I, [_VP, _P, _P, _I, _I, _P]),
    "gempic_maxwell2d_compute_e_from_b": (_I, [_VP, _D, _P, _P, _P, _P, _P, _P]),
    "gempic_maxwell2d_compute_b_from_e": (_I, [_VP, _D, _P, _P, _P, _P, _P, _P]),
    "gempic_maxwell2d_compute_e_from_j": (_I, [_VP, _P, _I, _P]),
    "gempic_maxwell2d_compute_rho_from_e": (_I, [_VP, _P, _P, _P, _P]),
    "gempic_maxwell2d_compute_e_from_rho": (_I, [_VP, _P, _P, _P]),
}


class GempicError(RuntimeError):
    """A device or state error reported by the library."""


class InexactError(ArithmeticError):
    """trunc(Int, x) of a non-finite or out-of-range position (Julia's InexactError)."""


_lib = None


def load():
    """Load the in-tree library; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GempicError(
                f"{LIB_PATH} is not built: run `make -C {PKG_ROOT}` (hipcc, gfx950). "
                "The particle-field path has no CPU fallback.")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    """Map a C-ABI status onto the reference's exception types."""
    if rc == OK:
        return
    msg = load().gempic_last_error().decode(errors="replace")
    if rc == ERR_ARGUMENT:
        raise ValueError(msg)          # Julia ArgumentError
    if rc == ERR_ASSERTION:
        raise AssertionError(msg)
    if rc == ERR_INEXACT:
        raise InexactError(msg)
    raise GempicError(msg)


def ptr(a):
    """double* of a float64 numpy array (contiguous in C or Fortran order)."""
    import numpy as np
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.dtype == np.float64, "float64 array expected"
    assert a.flags.c_contiguous or a.flags.f_contiguous, "contiguous array expected"
    return a.ctypes.data_as(_P)
