"""ctypes view of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE: this module is the checker. Only tests/, the smoke()
check in __graft_entry__.py and bench.py's cpu_baseline leg import it; the
product package (gempic.jl_amd/gempic_amd) never does.

Each wrapper names the reference function it restates (file:line in the
GEMPIC.jl tree); the C file carries the line-by-line restatement.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_D = ctypes.c_double
_I64 = ctypes.c_int64
_I = ctypes.c_int
_P = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.POINTER(ctypes.c_double))
_FN2 = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p)

GALERKIN = 0
COLLOCATION = 1


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "or_jl_mod_f64": (_D, [_D, _D]),
            "or_jl_mod_i64": (_I64, [_I64, _I64]),
            "or_bspline_basis": (None, [_I, _D, _P]),
            "or_eval_uniform_periodic_spline_curve": (None, [_I, _I64, _P, _P]),
            "or_gausslegendre": (_I, [_I, _P, _P]),
            "or_splinepp_tables": (_I, [_I, _P, _P]),
            "or_b_to_pp": (None, [_I, _I64, _P, _P]),
            "or_pmc1d_init": (_I, [_VP, _D, _D, _I64, _I, _I]),
            "or_add_charge": (None, [_P, _VP, _D, _D]),
            "or_add_charge_pp": (None, [_P, _VP, _D, _D]),
            "or_evaluate": (_D, [_VP, _D, _P]),
            "or_evaluate_pp": (_D, [_VP, _D, _P]),
            "or_add_current_update_v": (_D, [_P, _VP, _D, _D, _D, _D, _P, _D]),
            "or_add_current_update_v_nob": (_D, [_P, _VP, _D, _D, _D, _D, _D]),
            "or_add_current_update_v_pp": (_D, [_P, _VP, _D, _D, _D, _D, _P, _D]),
            "or_add_charge_n": (None, [_P, _VP, _I64, _P, _P]),
            "or_evaluate_n": (None, [_VP, _I64, _P, _P, _P]),
            "or_maxwell1d_create": (_VP, [_D, _D, _I64, _I]),
            "or_maxwell1d_destroy": (None, [_VP]),
            "or_maxwell1d_tables": (None, [_VP, _P, _P, _P, _P]),
            "or_maxwell1d_work": (_P, [_VP]),
            "or_r2hc": (None, [_VP, _P, _P]),
            "or_hc2r": (None, [_VP, _P, _P]),
            "or_solve_circulant": (None, [_VP, _P, _P]),
            "or_compute_e_from_rho": (None, [_P, _VP, _P]),
            "or_compute_e_from_j": (None, [_P, _VP, _P, _I]),
            "or_compute_e_from_b": (None, [_P, _VP, _D, _P]),
            "or_compute_b_from_e": (None, [_P, _VP, _D, _P]),
            "or_compute_rhs_from_function": (None, [_P, _VP, _D, _D, _D, _D, _I]),
            "or_l2projection": (_I, [_P, _VP, _D, _D, _D, _D, _I]),
            "or_l2norm_squared": (_D, [_VP, _P, _I]),
            "or_inner_product": (_D, [_VP, _P, _P, _I]),
            "or_split_create": (_VP, [_I64, _D, _D, _I, _I, _I, _I, _I, _I, _I64, _I, _D, _D, _D,
                                      _P, _P, _P, _P]),
            "or_split_destroy": (None, [_VP]),
            "or_split_j": (_P, [_VP, _I]),
            "or_operator_hb": (None, [_VP, _D]),
            "or_operator_he": (None, [_VP, _D]),
            "or_operator_hp1": (None, [_VP, _D]),
            "or_operator_hp2": (None, [_VP, _D]),
            "or_strang_splitting": (None, [_VP, _D, _I64]),
            "or_sobol2": (None, [ctypes.c_uint64, _P, _P]),
            "or_landau_newton": (_I, [_D, _D, _D, _P]),
            "or_sample_landau": (_I, [_I, _I64, _D, _D, _D, _D, _P, _I64]),
            "or_boris_mid": (_P, [_VP, _I]),
            "or_boris_set_mid": (None, [_VP, _P, _P, _P]),
            "or_boris_push_v_epart": (None, [_VP, _D]),
            "or_boris_push_v_bpart": (None, [_VP, _D]),
            "or_boris_push_x_accumulate_j": (None, [_VP, _D]),
            "or_boris_staggering": (None, [_VP, _D]),
            "or_boris_strang": (None, [_VP, _D, _I64]),
            "or_split_set_particles": (_I, [_VP, _P, _I64]),
            "or_split_set_tasks": (_I, [_VP, _I]),
            "or_bspline_deriv": (None, [_I, _D, _P]),
            "or_strang_loss": (_D, [_VP, _D, _I64, _D]),
            "or_strang_adjoint": (_I, [_VP, _D, _I64, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
            "or_solve_poisson_adjoint": (None, [_VP, _P, _P]),
            "or_strang_adjoint_domain": (None, [_VP, _P, _P]),
            "or_solve_poisson": (None, [_P, _VP, _P]),
            "or_write_step": (None, [_VP, _D, _I, _P, _P, _P, _P, _P, _P, _P]),
            "or_pmc2d_init": (_I, [_VP, _D, _D, _I64, _D, _D, _I64, _I, _I]),
            "or_pmc2d_add_charge": (None, [_P, _VP, _D, _D, _D]),
            "or_pmc2d_add_charge_pp": (None, [_P, _VP, _D, _D, _D]),
            "or_pmc2d_evaluate": (_D, [_VP, _D, _D, _P]),
            "or_pmc2d_evaluate_multiple": (None, [_VP, _D, _D, _P, _P, _P]),
            "or_pmc2d_add_charge_n": (None, [_P, _VP, _I64, _P, _P, _P]),
            "or_pmc2d_evaluate_n": (None, [_VP, _I64, _P, _P, _P, _P]),
            "or_pmc2d_evaluate_multiple_n": (None, [_VP, _I64, _P, _P, _P, _P, _P, _P]),
            "or_pmc2d_b_to_pp": (None, [_VP, _P, _P]),
            "or_pmc2d_evaluate_pp": (_D, [_VP, _D, _D, _P, ctypes.POINTER(_I)]),
            "or_spline_fem_mass_line": (None, [_I, _P]),
            "or_spline_fem_compute_mass_eig": (None, [_I64, _I, _P, _P]),
            "or_maxwell2d_create": (_VP, [_D, _D, _I64, _D, _D, _I64, _I]),
            "or_maxwell2d_destroy": (None, [_VP]),
            "or_maxwell2d_table": (_P, [_VP, _I]),
            "or_maxwell2d_solve": (None, [_VP, _I, _I, _P, _P]),
            "or_maxwell2d_multiply_mass": (None, [_VP, _P, _I, _P, _I, _P, _P]),
            "or_maxwell2d_inner_product": (_D, [_VP, _P, _P, _I, _I]),
            "or_maxwell2d_compute_e_from_b": (None, [_VP, _D, _PP, _PP]),
            "or_maxwell2d_compute_b_from_e": (None, [_VP, _D, _PP, _PP]),
            "or_maxwell2d_compute_e_from_j": (None, [_VP, _P, _I, _P]),
            "or_maxwell2d_compute_rho_from_e": (None, [_VP, _PP, _P]),
            "or_maxwell2d_compute_e_from_rho": (None, [_VP, _P, _PP]),
            "or_maxwell2d_fem_rhs": (None, [_VP, _FN2, _VP, _I, _I, _P]),
            "or_maxwell2d_l2projection": (None, [_VP, _FN2, _VP, _I, _I, _P]),
            "or_split2d_create": (_VP, [_I64, _I64, _D, _D, _D, _D, _I, _I64, _D, _D, _D, _P,
                                        _PP, _PP]),
            "or_split2d_destroy": (None, [_VP]),
            "or_split2d_j": (_P, [_VP, _I]),
            "or_split2d_operator_hb": (None, [_VP, _D]),
            "or_split2d_operator_he": (None, [_VP, _D]),
            "or_split2d_operator_hp1": (None, [_VP, _D]),
            "or_split2d_operator_hp2": (None, [_VP, _D]),
            "or_split2d_operator_hp3": (None, [_VP, _D]),
            "or_split2d_strang": (None, [_VP, _D, _I64]),
            "or_split2d_rho": (None, [_VP, _P]),
            "or_split2d_rho_from_e": (None, [_VP, _P]),
            "or_split2d_energy": (None, [_VP, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.c_contiguous or a.flags.f_contiguous)
    return a.ctypes.data_as(_P)


# --------------------------------------------------------------------------
# low-level helpers

def jl_mod(x: float, y: float) -> float:
    return lib().or_jl_mod_f64(x, y)


def bspline_deriv(degree: int, offset: float) -> np.ndarray:
    out = np.zeros(degree + 1)
    lib().or_bspline_deriv(degree, offset, _p(out))
    return out


def bspline_basis(degree: int, offset: float) -> np.ndarray:
    """uniform_bsplines_eval_basis (low_level_bsplines.jl:38-61)."""
    out = np.zeros(degree + 1)
    lib().or_bspline_basis(degree, offset, _p(out))
    return out


def eval_uniform_periodic_spline_curve(degree: int, scoef: np.ndarray) -> np.ndarray:
    """low_level_bsplines.jl:121-137."""
    scoef = np.ascontiguousarray(scoef, dtype=np.float64)
    out = np.zeros_like(scoef)
    lib().or_eval_uniform_periodic_spline_curve(degree, scoef.size, _p(scoef), _p(out))
    return out


def gausslegendre(n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    assert lib().or_gausslegendre(n, _p(x), _p(w)) == 0
    return x, w


def splinepp_tables(degree: int):
    """poly_coeffs / poly_coeffs_fp of SplinePP (splinepp.jl:30-233) as
    (deg+1, deg+1) arrays indexed [power_row, piece]."""
    d1 = degree + 1
    a = np.zeros(d1 * d1)
    b = np.zeros(d1 * d1)
    assert lib().or_splinepp_tables(degree, _p(a), _p(b)) == 0
    return a.reshape(d1, d1, order="F"), b.reshape(d1, d1, order="F")


def b_to_pp(degree: int, b: np.ndarray) -> np.ndarray:
    """b_to_pp (splinepp.jl:241-261): returns (deg+1, ncells) Fortran array."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros((degree + 1, b.size), order="F")
    lib().or_b_to_pp(degree, b.size, _p(b), _p(out))
    return out


# --------------------------------------------------------------------------
# ParticleMeshCoupling1D

class PMC1D:
    """ParticleMeshCoupling1D (particle_mesh_coupling_1d.jl:26-95)."""

    _SIZE = 8 * 1 + 8 * 4 + 4 * 3 + 4 + 16 * 8 + 2 * 36 * 8 + 64

    def __init__(self, xmin, xmax, nx, degree, smoothing=GALERKIN):
        self._buf = ctypes.create_string_buffer(1024)
        rc = lib().or_pmc1d_init(ctypes.addressof(self._buf), xmin, xmax, nx, degree, smoothing)
        if rc != 0:
            raise ValueError(f"invalid ParticleMeshCoupling1D arguments (rc={rc})")
        self.nx = nx
        self.degree = degree

    @property
    def ptr(self):
        return ctypes.addressof(self._buf)

    def add_charge(self, rho, x, q):
        lib().or_add_charge(_p(rho), self.ptr, x, q)

    def add_charge_pp(self, rho, x, q):
        lib().or_add_charge_pp(_p(rho), self.ptr, x, q)

    def evaluate(self, x, field):
        return lib().or_evaluate(self.ptr, x, _p(field))

    def evaluate_pp(self, x, field_pp):
        return lib().or_evaluate_pp(self.ptr, x, _p(field_pp))

    def add_current_update_v(self, j, x_old, x_new, q, qm, b, v):
        return lib().or_add_current_update_v(_p(j), self.ptr, x_old, x_new, q, qm, _p(b), v)

    def add_current_update_v_nob(self, j, x_old, x_new, q, qm, v):
        return lib().or_add_current_update_v_nob(_p(j), self.ptr, x_old, x_new, q, qm, v)

    def add_current_update_v_pp(self, j, x_old, x_new, q, qm, b, v):
        return lib().or_add_current_update_v_pp(_p(j), self.ptr, x_old, x_new, q, qm, _p(b), v)

    def add_charge_n(self, rho, x, q):
        x = np.ascontiguousarray(x, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        lib().or_add_charge_n(_p(rho), self.ptr, x.size, _p(x), _p(q))

    def evaluate_n(self, x, field):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        lib().or_evaluate_n(self.ptr, x.size, _p(x), _p(field), _p(out))
        return out


# --------------------------------------------------------------------------
# Maxwell1DFEM

class Maxwell1D:
    """Maxwell1DFEM (maxwell_1d_fem.jl:29-177) and its verbs."""

    def __init__(self, xmin, xmax, n, degree):
        self.h = lib().or_maxwell1d_create(xmin, xmax, n, degree)
        if not self.h:
            raise ValueError("invalid Maxwell1DFEM arguments")
        self.n = n
        self.degree = degree
        self.delta_x = (xmax - xmin) / n

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_maxwell1d_destroy(self.h)
            self.h = None

    def tables(self):
        t = [np.zeros(self.n) for _ in range(4)]
        lib().or_maxwell1d_tables(self.h, *[_p(a) for a in t])
        return dict(eig_mass0=t[0], eig_mass1=t[1], eig_weak_ampere=t[2], eig_weak_poisson=t[3])

    def work(self):
        return np.ctypeslib.as_array(lib().or_maxwell1d_work(self.h), shape=(self.n,)).copy()

    def r2hc(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        lib().or_r2hc(self.h, _p(x), _p(out))
        return out

    def hc2r(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        lib().or_hc2r(self.h, _p(x), _p(out))
        return out

    def solve_circulant(self, eig, rhs):
        lib().or_solve_circulant(self.h, _p(np.ascontiguousarray(eig)), _p(np.ascontiguousarray(rhs)))
        return self.work()

    def compute_e_from_rho(self, e, rho):
        lib().or_compute_e_from_rho(_p(e), self.h, _p(rho))

    def compute_e_from_j(self, e, j, component):
        lib().or_compute_e_from_j(_p(e), self.h, _p(j), component)

    def compute_e_from_b(self, e, dt, b):
        lib().or_compute_e_from_b(_p(e), self.h, dt, _p(b))

    def compute_b_from_e(self, b, dt, e):
        lib().or_compute_b_from_e(_p(b), self.h, dt, _p(e))

    def compute_rhs_from_function(self, coefs, degree, a_cos=0.0, a_sin=0.0, k=0.0, c=0.0):
        lib().or_compute_rhs_from_function(_p(coefs), self.h, a_cos, a_sin, k, c, degree)

    def l2projection(self, coefs, degree, a_cos=0.0, a_sin=0.0, k=0.0, c=0.0):
        if lib().or_l2projection(_p(coefs), self.h, a_cos, a_sin, k, c, degree) != 0:
            raise ValueError(f"degree {degree} not available")

    def l2norm_squared(self, coefs, degree):
        return lib().or_l2norm_squared(self.h, _p(coefs), degree)

    def inner_product(self, c1, c2, degree):
        return lib().or_inner_product(self.h, _p(c1), _p(c2), degree)


# --------------------------------------------------------------------------
# HamiltonianSplitting

class Splitting:
    """HamiltonianSplitting{1,V} over caller-owned (aliased) arrays
    (hamiltonian_splitting.jl:20-86). `particles` is the (1+V+nw, N)
    column-major particle array (a Fortran-ordered numpy array)."""

    def __init__(self, nx, xmin, xmax, maxwell_degree, degree0, degree1, particles, e1, e2, b,
                 charge=1.0, mass=1.0, common_weight=0.0, smoothing0=GALERKIN,
                 smoothing1=GALERKIN, n_weights=1):
        assert particles.flags.f_contiguous and particles.dtype == np.float64
        dim_v = particles.shape[0] - 1 - n_weights
        self.particles, self.e1, self.e2, self.b = particles, e1, e2, b
        self.nx = nx
        self.h = lib().or_split_create(nx, xmin, xmax, maxwell_degree, degree0, smoothing0,
                                       degree1, smoothing1, dim_v, particles.shape[1], n_weights,
                                       charge, mass, common_weight, _p(particles), _p(e1),
                                       _p(e2), _p(b))
        if not self.h:
            raise ValueError("invalid HamiltonianSplitting arguments")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_split_destroy(self.h)
            self.h = None

    def j(self, component):
        return np.ctypeslib.as_array(lib().or_split_j(self.h, component), shape=(self.nx,)).copy()

    def set_tasks(self, ntasks):
        """The reference's nthreads() task split of the 1d2v operators
        (hamiltonian_splitting.jl:61-66)."""
        if lib().or_split_set_tasks(self.h, int(ntasks)) != 0:
            raise ValueError("n_particles % ntasks != 0")

    def operatorHB(self, dt):
        lib().or_operator_hb(self.h, dt)

    def operatorHE(self, dt):
        lib().or_operator_he(self.h, dt)

    def operatorHp1(self, dt):
        lib().or_operator_hp1(self.h, dt)

    def operatorHp2(self, dt):
        lib().or_operator_hp2(self.h, dt)

    def strang_splitting(self, dt, nsteps):
        lib().or_strang_splitting(self.h, dt, nsteps)

    def solve_poisson(self, efield, rho):
        lib().or_solve_poisson(_p(efield), self.h, _p(rho))

    # reverse mode (no reference counterpart): loss = wloss * sum over steps
    # of 1/2 l2norm_squared(B_s, p-1), plus <ct, final state>
    def strang_loss(self, dt, nsteps, wloss):
        return lib().or_strang_loss(self.h, dt, nsteps, wloss)

    def strang_adjoint(self, dt, nsteps, wloss, ct_part=None, ct_e1=None, ct_e2=None, ct_b=None):
        """Runs nsteps forward (like strang_splitting) and returns
        (loss, g_part, g_e1, g_e2, g_b): the cotangents of the initial state
        (g_part has the particle array's shape; the weight row is zero)."""
        loss = np.zeros(1)
        g = np.zeros_like(self.particles, order="F")
        ge1, ge2, gb = np.zeros(self.nx), np.zeros(self.nx), np.zeros(self.nx)
        cp = None if ct_part is None else np.asfortranarray(ct_part, dtype=np.float64)
        rc = lib().or_strang_adjoint(self.h, dt, nsteps, wloss, _p(cp), _p(ct_e1), _p(ct_e2),
                                     _p(ct_b), _p(loss), _p(g), _p(ge1), _p(ge2), _p(gb))
        if rc:
            raise ValueError(f"or_strang_adjoint: {rc}")
        return float(loss[0]), g, ge1, ge2, gb

    def strang_adjoint_domain(self):
        """(d loss / d Lx with the state held fixed, sum_i d loss / d w_i) of
        the last strang_adjoint."""
        gl, gw = np.zeros(1), np.zeros(1)
        lib().or_strang_adjoint_domain(self.h, _p(gl), _p(gw))
        return float(gl[0]), float(gw[0])

    def solve_poisson_adjoint(self, ct_e):
        g = np.zeros(self.particles.shape[1])
        lib().or_solve_poisson_adjoint(self.h, _p(np.ascontiguousarray(ct_e, dtype=np.float64)),
                                       _p(g))
        return g

    # HamiltonianSplittingBoris (hamiltonian_splitting_boris.jl) on this state
    def boris_mid(self):
        """(e_dofs_mid[1], e_dofs_mid[2], b_dofs_mid)"""
        return tuple(np.ctypeslib.as_array(lib().or_boris_mid(self.h, k), shape=(self.nx,)).copy()
                     for k in range(3))

    def boris_set_mid(self, e1m, e2m, bm):
        lib().or_boris_set_mid(self.h, _p(e1m), _p(e2m), _p(bm))

    def boris_push_v_epart(self, dt):
        lib().or_boris_push_v_epart(self.h, dt)

    def boris_push_v_bpart(self, dt):
        lib().or_boris_push_v_bpart(self.h, dt)

    def boris_push_x_accumulate_j(self, dt):
        lib().or_boris_push_x_accumulate_j(self.h, dt)

    def boris_staggering(self, dt):
        lib().or_boris_staggering(self.h, dt)

    def boris_strang(self, dt, nsteps):
        lib().or_boris_strang(self.h, dt, nsteps)

    def write_step(self, time, degree, e1, e2, b, e1_n, e2_n, efield_poisson):
        out = np.zeros(11)
        lib().or_write_step(self.h, time, degree, _p(e1), _p(e2), _p(b), _p(e1_n), _p(e2_n),
                            _p(efield_poisson), _p(out))
        return out


# --------------------------------------------------------------------------
# 2D: ParticleMeshCoupling2D, TwoDMaxwell (gempic_oracle2d.c)

class _Pmc2dStruct(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("xmin", _D), ("ymin", _D),
                ("dx", _D), ("dy", _D), ("scaling", _D), ("degree", _I), ("n_span", _I),
                ("pp", _D * 36), ("pp_fp", _D * 36)]


class PMC2D:
    """ParticleMeshCoupling2D (particle_mesh_coupling_2d.jl:12-45) on a
    TwoDGrid (mesh.jl:17-42). Grid arrays are nx*ny, x fastest."""

    def __init__(self, xmin, xmax, nx, ymin, ymax, ny, degree, smoothing=GALERKIN):
        self.s = _Pmc2dStruct()
        rc = lib().or_pmc2d_init(ctypes.byref(self.s), xmin, xmax, nx, ymin, ymax, ny, degree,
                                 smoothing)
        if rc != 0:
            raise ValueError("invalid ParticleMeshCoupling2D arguments")
        self.nx, self.ny, self.degree = nx, ny, degree

    def ptr(self):
        return ctypes.byref(self.s)

    def add_charge(self, rho, x, y, w):
        lib().or_pmc2d_add_charge(_p(rho), self.ptr(), x, y, w)

    def add_charge_pp(self, rho, x, y, w):
        lib().or_pmc2d_add_charge_pp(_p(rho), self.ptr(), x, y, w)

    def evaluate(self, x, y, field):
        return lib().or_pmc2d_evaluate(self.ptr(), x, y, _p(field))

    def evaluate_multiple(self, x, y, f1, f2):
        out = np.zeros(2)
        lib().or_pmc2d_evaluate_multiple(self.ptr(), x, y, _p(f1), _p(f2), _p(out))
        return out[0], out[1]

    def add_charge_n(self, rho, x, y, w):
        """The reference's per-particle loop over arrays (in order)."""
        x, y, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, w))
        lib().or_pmc2d_add_charge_n(_p(rho), self.ptr(), x.size, _p(x), _p(y), _p(w))

    def evaluate_n(self, x, y, field):
        x, y = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y))
        out = np.zeros(x.size)
        lib().or_pmc2d_evaluate_n(self.ptr(), x.size, _p(x), _p(y), _p(field), _p(out))
        return out

    def evaluate_multiple_n(self, x, y, f1, f2):
        x, y = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y))
        o1, o2 = np.zeros(x.size), np.zeros(x.size)
        lib().or_pmc2d_evaluate_multiple_n(self.ptr(), x.size, _p(x), _p(y), _p(f1), _p(f2),
                                           _p(o1), _p(o2))
        return o1, o2

    def b_to_pp(self, b):
        d = self.degree + 1
        pp = np.zeros((d * d, self.nx * self.ny), order="F")
        lib().or_pmc2d_b_to_pp(self.ptr(), _p(b), _p(pp))
        return pp

    def evaluate_pp(self, x, y, pp):
        err = ctypes.c_int(0)
        v = lib().or_pmc2d_evaluate_pp(self.ptr(), x, y, _p(pp), ctypes.byref(err))
        if err.value:
            raise IndexError("BoundsError: position outside the pp table")
        return v


def spline_fem_mass_line(degree):
    out = np.zeros(degree + 1)
    lib().or_spline_fem_mass_line(degree, _p(out))
    return out


def spline_fem_compute_mass_eig(n, degree, mass_line):
    out = np.zeros(n)
    ml = np.ascontiguousarray(mass_line, dtype=np.float64)
    lib().or_spline_fem_compute_mass_eig(n, degree, _p(ml), _p(out))
    return out


def _ptrs3(arrs):
    return (ctypes.POINTER(_D) * 3)(*[_p(a) for a in arrs])


class Maxwell2D:
    """TwoDMaxwell (maxwell_2d_fem.jl:11-88) with its TwoDPoisson and
    TwoDLinearSolverSplineMass solvers."""

    def __init__(self, xmin, xmax, nx, ymin, ymax, ny, degree):
        self.h = lib().or_maxwell2d_create(xmin, xmax, nx, ymin, ymax, ny, degree)
        if not self.h:
            raise ValueError("invalid TwoDMaxwell arguments")
        self.nx, self.ny, self.degree = nx, ny, degree

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_maxwell2d_destroy(self.h)
            self.h = None

    def table(self, which, n):
        return np.ctypeslib.as_array(lib().or_maxwell2d_table(self.h, which), shape=(n,)).copy()

    def solve(self, form, component, rhs):
        out = np.zeros(self.nx * self.ny)
        lib().or_maxwell2d_solve(self.h, form, component, _p(rhs), _p(out))
        return out

    def multiply_mass(self, line1, line2, cin):
        out = np.zeros(self.nx * self.ny)
        l1 = np.ascontiguousarray(line1, dtype=np.float64)
        l2 = np.ascontiguousarray(line2, dtype=np.float64)
        lib().or_maxwell2d_multiply_mass(self.h, _p(l1), len(l1) - 1, _p(l2), len(l2) - 1,
                                         _p(cin), _p(out))
        return out

    def inner_product(self, c1, c2, component, form):
        return lib().or_maxwell2d_inner_product(self.h, _p(c1), _p(c2), component, form)

    def compute_e_from_b(self, e, dt, b):
        lib().or_maxwell2d_compute_e_from_b(self.h, dt, _ptrs3(b), _ptrs3(e))

    def compute_b_from_e(self, b, dt, e):
        lib().or_maxwell2d_compute_b_from_e(self.h, dt, _ptrs3(e), _ptrs3(b))

    def compute_e_from_j(self, e, current, component):
        lib().or_maxwell2d_compute_e_from_j(self.h, _p(current), component, _p(e))

    def compute_rho_from_e(self, rho, e):
        lib().or_maxwell2d_compute_rho_from_e(self.h, _ptrs3(e), _p(rho))

    def compute_e_from_rho(self, e, rho):
        lib().or_maxwell2d_compute_e_from_rho(self.h, _p(rho), _ptrs3(e))

    def fem_rhs(self, f, component, form):
        cb = _FN2(lambda x, y, ctx: float(f(x, y)))
        out = np.zeros(self.nx * self.ny)
        lib().or_maxwell2d_fem_rhs(self.h, cb, None, component, form, _p(out))
        return out

    def l2projection(self, f, component, form):
        cb = _FN2(lambda x, y, ctx: float(f(x, y)))
        out = np.zeros(self.nx * self.ny)
        lib().or_maxwell2d_l2projection(self.h, cb, None, component, form, _p(out))
        return out


class Splitting2D3V:
    """The 2D3V Hamiltonian splitting (gempic_oracle2d3v.c; the reference's
    hamiltonian_splitting_2d3v.jl is empty). particles (6 x n, rows x1, x2,
    v1, v2, v3, w, Fortran order), e and b (3 arrays of nx*ny each) are
    aliased and updated in place."""

    def __init__(self, nx, ny, xmin, xmax, ymin, ymax, degree, particles, e, b, charge=1.0,
                 mass=1.0, common_weight=0.0):
        assert particles.flags.f_contiguous and particles.shape[0] == 6
        self._keep = (particles, e, b)
        self._pe, self._pb = _ptrs3(e), _ptrs3(b)
        self.h = lib().or_split2d_create(nx, ny, xmin, xmax, ymin, ymax, degree,
                                         particles.shape[1], charge, mass, common_weight,
                                         _p(particles), self._pe, self._pb)
        if not self.h:
            raise ValueError("invalid 2D3V splitting arguments")
        self.nx, self.ny = nx, ny

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_split2d_destroy(self.h)
            self.h = None

    def j(self, component):
        return np.ctypeslib.as_array(lib().or_split2d_j(self.h, component),
                                     shape=(self.nx * self.ny,)).copy()

    def operatorHB(self, dt):
        lib().or_split2d_operator_hb(self.h, dt)

    def operatorHE(self, dt):
        lib().or_split2d_operator_he(self.h, dt)

    def operatorHp1(self, dt):
        lib().or_split2d_operator_hp1(self.h, dt)

    def operatorHp2(self, dt):
        lib().or_split2d_operator_hp2(self.h, dt)

    def operatorHp3(self, dt):
        lib().or_split2d_operator_hp3(self.h, dt)

    def strang_splitting(self, dt, nsteps):
        lib().or_split2d_strang(self.h, dt, nsteps)

    def rho(self):
        out = np.zeros(self.nx * self.ny)
        lib().or_split2d_rho(self.h, _p(out))
        return out

    def rho_from_e(self):
        out = np.zeros(self.nx * self.ny)
        lib().or_split2d_rho_from_e(self.h, _p(out))
        return out

    def energy(self):
        """kinetic, 1/2 (E_c, M1 E_c) for c = 1..3, 1/2 (B_c, M2 B_c) for c = 1..3"""
        out = np.zeros(7)
        lib().or_split2d_energy(self.h, _p(out))
        return out


# --------------------------------------------------------------------------
# sample!(pg, alpha, k, sigma, mesh) (particle_sampling.jl:236-311)

def sobol2(k):
    """The k-th point (k >= 1) of Sobol.jl's SobolSeq(2)."""
    r = np.zeros(2)
    lib().or_sobol2(k, _p(r[:1]), _p(r[1:]))
    return float(r[0]), float(r[1])


def landau_newton(r, alpha, k):
    out = np.zeros(1)
    it = lib().or_landau_newton(r, alpha, k, _p(out))
    return float(out[0]), it


def sample_landau(dim_v, n, alpha, k, sigma, dimx):
    """The (1 + dim_v + 1, n) column-major particle array of sample!."""
    a = np.zeros((2 + dim_v, n), order="F")
    it = lib().or_sample_landau(dim_v, n, alpha, k, sigma, dimx, _p(a), 2 + dim_v)
    return a, it
