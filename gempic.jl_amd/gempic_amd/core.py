"""Host-side mirror of GEMPIC.jl's particle-field API over the HIP library.

Names, argument meaning and error behaviour follow the reference (Julia);
Julia's mutating `f!` functions are spelled `f_` here. Compute runs on the
GPU through include/gempic_hip.h. Set-up arithmetic that the reference does
once on the host (meshes, quadrature-based right-hand sides) stays on the
host, as in the reference.
"""
from __future__ import annotations

import ctypes
import math
import weakref

import numpy as np

from . import _lib
from ._lib import check, ptr

# particle group -> (weak reference to) the GPU splitting that holds it, so
# that the diagnostics called with the group itself (solve_poisson!,
# write_step!) run on that splitting's GPU copy
_ON_GPU = weakref.WeakKeyDictionary()

__all__ = [
    "OneDGrid", "ParticleGroup", "ParticleMeshCoupling1D", "Maxwell1DFEM",
    "HamiltonianSplitting", "add_charge_", "add_charge_pp_", "evaluate", "evaluate_pp",
    "add_current_update_v_", "add_current_update_v_pp_", "b_to_pp",
    "compute_e_from_rho_", "compute_e_from_j_", "compute_e_from_b_", "compute_b_from_e_",
    "compute_rhs_from_function_", "l2projection_", "inner_product", "l2norm_squared",
    "operatorHB", "operatorHE", "operatorHp1", "operatorHp2", "strang_splitting_",
    "solve_poisson_", "TimeHistoryDiagnostics", "write_step_", "get_charge", "save", "load",
    "HamiltonianSplittingBoris", "staggering_", "push_v_epart_", "push_v_bpart_",
    "push_x_accumulate_j_", "sample_", "sample_particles_", "strang_splitting_adjoint_",
    "get_adjoint", "solve_poisson_adjoint", "AdjointResult", "Ensemble",
    "ensemble_strang_splitting_", "ensemble_strang_splitting_adjoint_", "ensemble_autotune",
]


# ---------------------------------------------------------------------------
# mesh.jl

class OneDGrid:
    """OneDGrid(xmin, xmax, nx) (mesh.jl:52-67)."""

    def __init__(self, xmin: float, xmax: float, nx: int):
        self.nx = int(nx)
        self.xmin = float(xmin)
        self.xmax = float(xmax)
        self.dimx = self.xmax - self.xmin
        self.dx = self.dimx / (self.nx - 1)        # as in the reference (unused by the path)
        self.x = np.linspace(self.xmin, self.xmax, self.nx + 1)


# ---------------------------------------------------------------------------
# particle_group.jl

class ParticleGroup:
    """ParticleGroup{D,V}(n; charge, mass, n_weights, common_weight)
    (particle_group.jl:15-46). `array` is the (D+V+n_weights, N) matrix in
    column-major (Fortran) order, exactly the reference's memory layout."""

    def __init__(self, D: int, V: int, n_particles: int, charge=1.0, mass=1.0, n_weights=1,
                 common_weight=0.0):
        self.dims = (int(D), int(V))
        self.n_particles = int(n_particles)
        self.array = np.zeros((D + V + n_weights, n_particles), order="F")
        if common_weight == 0.0:
            common_weight = 1.0 / n_particles
        self.common_weight = float(common_weight)
        self.charge = float(charge)
        self.mass = float(mass)
        self.n_weights = int(n_weights)
        self.q_over_m = self.charge / self.mass


def get_charge(p: ParticleGroup, i: int, i_wi: int = 0) -> float:
    """get_charge (particle_group.jl:67-69); 0-based particle index."""
    D, V = p.dims
    return p.charge * p.array[D + V + i_wi, i] * p.common_weight


def save(file: str, step: int, p: ParticleGroup) -> str:
    """save(file, step, p) (particle_group.jl:152-165): the particle state as
    "x" (D x N), "v" (V x N) and "w" (n_weights x N) in
    "<file>-<step:06d>.npz" (the reference writes the same three arrays to a
    JLD2 file of that name; NumPy's archive is the format at hand here, and
    Julia's NPZ package reads it). Call it after a synchronising call (or
    to_host) so that the array holds the GPU state. Returns the path."""
    D, V = p.dims
    path = "%s-%06d.npz" % (file, int(step))
    np.savez(path, x=np.asfortranarray(p.array[:D]), v=np.asfortranarray(p.array[D:D + V]),
             w=np.asfortranarray(p.array[D + V:]))
    return path


def load(path: str, p: ParticleGroup) -> ParticleGroup:
    """The state written by save, back into p.array (shapes must match); a
    splitting then takes it with to_device (resume)."""
    D, V = p.dims
    with np.load(path, allow_pickle=False) as z:
        x, v, w = z["x"], z["v"], z["w"]
    if x.shape != (D, p.n_particles) or v.shape != (V, p.n_particles) or \
            w.shape != (p.array.shape[0] - D - V, p.n_particles):
        raise ValueError("the saved state does not match the particle group's shape")
    p.array[:D], p.array[D:D + V], p.array[D + V:] = x, v, w
    return p


# ---------------------------------------------------------------------------
# particle_mesh_coupling_1d.jl

_SMOOTHING = {"galerkin": _lib.GALERKIN, "collocation": _lib.COLLOCATION}


class ParticleMeshCoupling1D:
    """ParticleMeshCoupling1D(mesh, no_particles, spline_degree, smoothing_type)
    (particle_mesh_coupling_1d.jl:26-95)."""

    def __init__(self, mesh: OneDGrid, no_particles: int, spline_degree: int,
                 smoothing_type: str):
        if smoothing_type not in _SMOOTHING:
            raise ValueError(f"Smoothing Type {smoothing_type} not implemented for "
                             "kernel_smoother_spline_1d.")
        self.dims = 1
        self.n_grid = mesh.nx
        self.n_dofs = mesh.nx
        self.no_particles = int(no_particles)
        self.spline_degree = int(spline_degree)
        self.n_span = self.spline_degree + 1
        self.xmin = mesh.xmin
        self.xmax = mesh.xmax
        self.Lx = mesh.xmax - mesh.xmin
        self.delta_x = self.Lx / self.n_grid
        self.smoothing_type = smoothing_type
        self.smoothing = _SMOOTHING[smoothing_type]
        self.scaling = 1.0 / self.delta_x if smoothing_type == "collocation" else 1.0
        self.n_quad_points = (self.spline_degree + 2) // 2

    def _args(self):
        return (self.n_grid, self.xmin, self.xmax, self.spline_degree, self.smoothing)


def _arr(x):
    return np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64)))


def add_charge_(rho: np.ndarray, p: ParticleMeshCoupling1D, position, marker_charge, *wp):
    """add_charge! (particle_mesh_coupling_1d.jl:261-280), batched over arrays
    of positions/charges (a scalar pair is one particle). With a
    ParticleMeshCoupling2D: add_charge_(rho, pm, xp, yp, wp)
    (particle_mesh_coupling_2d.jl:94-105)."""
    if isinstance(p, ParticleMeshCoupling2D):
        return add_charge_2d_(rho, p, position, marker_charge, *wp)
    x = _arr(position)
    q = _arr(marker_charge) * np.ones_like(x)
    check(_lib.load().gempic_pmc1d_add_charge(*p._args(), x.size, ptr(x), ptr(q), ptr(rho)))


def add_charge_pp_(rho: np.ndarray, p: ParticleMeshCoupling1D, position, marker_charge, *wp):
    """add_charge_pp! (particle_mesh_coupling_1d.jl:106-122); with a
    ParticleMeshCoupling2D: add_charge_pp_(rho, pm, xp, yp, wp)
    (particle_mesh_coupling_2d.jl:127-148)."""
    if isinstance(p, ParticleMeshCoupling2D):
        return add_charge_pp_2d_(rho, p, position, marker_charge, *wp)
    x = _arr(position)
    q = _arr(marker_charge) * np.ones_like(x)
    check(_lib.load().gempic_pmc1d_add_charge_pp(*p._args(), x.size, ptr(x), ptr(q), ptr(rho)))


def evaluate(p: ParticleMeshCoupling1D, position, field_dofs: np.ndarray, *field2d):
    """evaluate (particle_mesh_coupling_1d.jl:438-453); array positions give
    the batched form of diagnostics.jl:260-268. With a ParticleMeshCoupling2D:
    evaluate(pm, xp, yp, field_dofs) (particle_mesh_coupling_2d.jl:189-203)."""
    if isinstance(p, ParticleMeshCoupling2D):
        return evaluate_2d(p, position, field_dofs, *field2d)
    x = _arr(position)
    out = np.zeros_like(x)
    f = np.ascontiguousarray(field_dofs, dtype=np.float64)
    check(_lib.load().gempic_pmc1d_evaluate(*p._args(), x.size, ptr(x), ptr(f), ptr(out)))
    return float(out[0]) if np.ndim(position) == 0 else out


def evaluate_pp(p: ParticleMeshCoupling1D, position, field_dofs_pp: np.ndarray, *pp2d):
    """evaluate_pp (particle_mesh_coupling_1d.jl:242-250); field_dofs_pp is
    the (degree+1, n_cells) matrix of b_to_pp. With a ParticleMeshCoupling2D:
    evaluate_pp(pm, xp, yp, pp) (particle_mesh_coupling_2d.jl:162-178)."""
    if isinstance(p, ParticleMeshCoupling2D):
        return evaluate_pp_2d(p, position, field_dofs_pp, *pp2d)
    x = _arr(position)
    out = np.zeros_like(x)
    f = np.asfortranarray(field_dofs_pp, dtype=np.float64)
    check(_lib.load().gempic_pmc1d_evaluate_pp(*p._args(), x.size, ptr(x), ptr(f), ptr(out)))
    return float(out[0]) if np.ndim(position) == 0 else out


def add_current_update_v_(j_dofs: np.ndarray, p: ParticleMeshCoupling1D, position_old,
                          position_new, marker_charge, qoverm: float, bfield_dofs, vi):
    """add_current_update_v! (particle_mesh_coupling_1d.jl:296-376); with
    bfield_dofs=None the 1d1v form without B (:471-529). Returns the updated
    v (scalar in, scalar out)."""
    xo = _arr(position_old)
    xn = _arr(position_new) * np.ones_like(xo)
    q = _arr(marker_charge) * np.ones_like(xo)
    v = _arr(vi) * np.ones_like(xo)
    b = None if bfield_dofs is None else np.ascontiguousarray(bfield_dofs, dtype=np.float64)
    check(_lib.load().gempic_pmc1d_add_current_update_v(*p._args(), xo.size, ptr(xo), ptr(xn),
                                                        ptr(q), float(qoverm), ptr(b), ptr(v),
                                                        ptr(j_dofs)))
    return float(v[0]) if np.ndim(vi) == 0 else v


def add_current_update_v_pp_(j_dofs: np.ndarray, p: ParticleMeshCoupling1D, position_old,
                             position_new, marker_charge, qoverm: float, bfield_dofs, vi):
    """add_current_update_v_pp! (particle_mesh_coupling_1d.jl:131-190)."""
    xo = _arr(position_old)
    xn = _arr(position_new) * np.ones_like(xo)
    q = _arr(marker_charge) * np.ones_like(xo)
    v = _arr(vi) * np.ones_like(xo)
    b = np.ascontiguousarray(bfield_dofs, dtype=np.float64)
    check(_lib.load().gempic_pmc1d_add_current_update_v_pp(*p._args(), xo.size, ptr(xo),
                                                           ptr(xn), ptr(q), float(qoverm),
                                                           ptr(b), ptr(v), ptr(j_dofs)))
    return float(v[0]) if np.ndim(vi) == 0 else v


def b_to_pp(p: ParticleMeshCoupling1D, ncells: int, b_coeffs: np.ndarray) -> np.ndarray:
    """b_to_pp(spline_pp, ncells, b_coeffs) (splinepp.jl:241-261)."""
    b = np.ascontiguousarray(b_coeffs, dtype=np.float64)
    out = np.zeros((p.spline_degree + 1, ncells), order="F")
    check(_lib.load().gempic_b_to_pp(p.spline_degree, ncells, ptr(b), ptr(out)))
    return out


# ---------------------------------------------------------------------------
# maxwell_1d_fem.jl

class Maxwell1DFEM:
    """Maxwell1DFEM(mesh, degree) (maxwell_1d_fem.jl:29-177); the circulant
    solves run on the GPU."""

    def __init__(self, mesh: OneDGrid, degree: int):
        if degree not in (1, 2, 3):
            raise ValueError(f"Wrong value of degree = {degree}  (1,2 or 3)")
        self.xmin = mesh.xmin
        self.Lx = mesh.xmax - mesh.xmin
        self.n_dofs = mesh.nx
        self.delta_x = self.Lx / self.n_dofs
        self.s_deg_0 = degree
        self.s_deg_1 = degree - 1
        self._mesh = mesh
        h = ctypes.c_void_p()
        check(_lib.load().gempic_maxwell1d_create(mesh.nx, mesh.xmin, mesh.xmax, degree,
                                                  ctypes.byref(h)))
        self._h = h
        t = [np.zeros(self.n_dofs) for _ in range(4)]
        check(_lib.load().gempic_maxwell1d_tables(self._h, *[ptr(a) for a in t]))
        self.eig_mass0, self.eig_mass1, self.eig_weak_ampere, self.eig_weak_poisson = t

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().gempic_maxwell1d_destroy(h)
            self._h = None


def compute_e_from_rho_(e, self: Maxwell1DFEM, rho):
    """compute_e_from_rho! (maxwell_1d_fem.jl:244-255); with a TwoDMaxwell,
    e is the list of field components (maxwell_2d_fem.jl:199-201)."""
    if isinstance(self, TwoDMaxwell):
        return compute_e_from_rho_2d_(e, self, rho)
    check(_lib.load().gempic_maxwell1d_compute_e_from_rho(self._h, ptr(e), ptr(_arr(rho))))


def compute_e_from_j_(e, self: Maxwell1DFEM, current, component: int):
    """compute_e_from_j! (maxwell_1d_fem.jl:263-289; TwoDMaxwell:
    maxwell_2d_fem.jl:455-459)."""
    if isinstance(self, TwoDMaxwell):
        return compute_e_from_j_2d_(e, self, current, component)
    check(_lib.load().gempic_maxwell1d_compute_e_from_j(self._h, ptr(e), ptr(_arr(current)),
                                                        int(component)))


def compute_e_from_b_(field_out, self: Maxwell1DFEM, delta_t: float, field_in):
    """compute_e_from_b! (maxwell_1d_fem.jl:384-396); with a TwoDMaxwell,
    compute_e_from_b_(e, solver, dt, b) on three components each
    (maxwell_2d_fem.jl:370-412)."""
    if isinstance(self, TwoDMaxwell):
        return compute_e_from_b_2d_(field_out, self, delta_t, field_in)
    check(_lib.load().gempic_maxwell1d_compute_e_from_b(self._h, ptr(field_out), float(delta_t),
                                                        ptr(_arr(field_in))))


def compute_b_from_e_(field_out, self: Maxwell1DFEM, delta_t: float, field_in):
    """compute_b_from_e! (maxwell_1d_fem.jl:407-420); with a TwoDMaxwell,
    compute_b_from_e_(b, solver, dt, e) (maxwell_2d_fem.jl:423-444)."""
    if isinstance(self, TwoDMaxwell):
        return compute_b_from_e_2d_(field_out, self, delta_t, field_in)
    check(_lib.load().gempic_maxwell1d_compute_b_from_e(self._h, ptr(field_out), float(delta_t),
                                                        ptr(_arr(field_in))))


def inner_product(self: Maxwell1DFEM, coefs1_dofs, coefs2_dofs, degree: int, *form) -> float:
    """inner_product (maxwell_1d_fem.jl:461-475); with a TwoDMaxwell,
    inner_product(solver, c1, c2, component, form) (maxwell_2d_fem.jl:511-575)."""
    if isinstance(self, TwoDMaxwell):
        return inner_product_2d(self, coefs1_dofs, coefs2_dofs, degree, *form)
    out = np.zeros(1)
    check(_lib.load().gempic_maxwell1d_inner_product(self._h, ptr(_arr(coefs1_dofs)),
                                                     ptr(_arr(coefs2_dofs)), int(degree),
                                                     ptr(out)))
    return float(out[0])


def l2norm_squared(self: Maxwell1DFEM, coefs_dofs, degree: int) -> float:
    """l2norm_squared (maxwell_1d_fem.jl:299-313)."""
    out = np.zeros(1)
    check(_lib.load().gempic_maxwell1d_l2norm_squared(self._h, ptr(_arr(coefs_dofs)),
                                                      int(degree), ptr(out)))
    return float(out[0])


def _bspline_basis(degree: int, x: float):
    """uniform_bsplines_eval_basis (low_level_bsplines.jl:38-61), host set-up copy."""
    b = [0.0] * (degree + 1)
    b[0] = 1.0
    for j in range(1, degree + 1):
        xx = -x
        jr = float(j)
        inv_j = 1.0 / jr
        saved = 0.0
        for r in range(j):
            xx = xx + 1.0
            temp = b[r] * inv_j
            b[r] = saved + xx * temp
            saved = (jr - xx) * temp
        b[j] = saved
    return b


_FN1D = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_void_p)


def compute_rhs_from_function_(coefs_dofs, self: Maxwell1DFEM, func, degree: int):
    """compute_rhs_from_function! (maxwell_1d_fem.jl:188-220) by the library
    (gempic_maxwell1d_compute_rhs_from_function), which calls func on the
    host at the Gauss-Legendre points."""
    cb = _FN1D(lambda x, ctx: float(func(x)))
    check(_lib.load().gempic_maxwell1d_compute_rhs_from_function(self._h, cb, None, int(degree),
                                                                 ptr(coefs_dofs)))
    return coefs_dofs


def l2projection_(coefs_dofs, self: Maxwell1DFEM, func, degree: int):
    """l2projection! (maxwell_1d_fem.jl:347-374) by the library. The
    reference solves into its workspace and then returns coefs_dofs ./
    delta_x, i.e. the rhs divided by dx; that result is kept."""
    cb = _FN1D(lambda x, ctx: float(func(x)))
    check(_lib.load().gempic_maxwell1d_l2projection(self._h, cb, None, int(degree),
                                                    ptr(coefs_dofs)))
    return coefs_dofs


# ---------------------------------------------------------------------------
# launch autotuning (HamiltonianSplitting.autotune)

# default candidates: workgroups per particle pass (0 = automatic; MI355X has
# 256 CUs and the reduction scratch holds 8 workgroups per CU) and Strang
# steps between sorts
AUTOTUNE_BLOCKS = (0, 512, 1024, 2048)
AUTOTUNE_SORT_INTERVALS = (10, 20, 40, 80)


def autotune_key(V, n, nx, deg0, deg1, dt, n_steps, blocks, sort_intervals,
                 kernels_per_step=1, arithmetic="reference", device="") -> str:
    """Cache key of one autotuning problem: everything the timing depends on
    (the problem, the candidates, the fused step's kernel split and
    arithmetic, and the GPU: its architecture and compute-unit count)."""
    return (f"V={int(V)},n={int(n)},nx={int(nx)},deg={int(deg0)}/{int(deg1)},dt={float(dt)!r},"
            f"steps={int(n_steps)},blocks={list(map(int, blocks))},"
            f"intervals={list(map(int, sort_intervals))},kernels={int(kernels_per_step)},"
            f"arith={arithmetic},device={device}")


def device_info() -> str:
    """The current GPU as "<arch>/<CUs>cu" (the architecture name HIP reports)."""
    arch = ctypes.create_string_buffer(64)
    cus = ctypes.c_int()
    check(_lib.load().gempic_device_info(ctypes.byref(cus), arch, 64))
    return f"{arch.value.decode()}/{cus.value}cu"


def autotune_cache_lookup(path: str, key: str):
    """The stored result for `key` in the JSON cache file, or None."""
    import json
    import os
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    hit = table.get(key)
    if hit is None or not isinstance(hit.get("blocks"), int) or \
            not isinstance(hit.get("sort_interval"), int):
        return None
    return hit


def autotune_cache_store(path: str, key: str, result: dict):
    """Add `result` under `key` to the JSON cache file (written atomically)."""
    import json
    import os
    table = {}
    if os.path.exists(path):
        with open(path) as f:
            table = json.load(f)
    table[key] = {k: v for k, v in result.items() if k != "cached"}
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


# ---------------------------------------------------------------------------
# hamiltonian_splitting.jl, _1d2v.jl, _1d1v.jl

class HamiltonianSplitting:
    """HamiltonianSplitting{D,V}(maxwell_solver, kernel_smoother_0,
    kernel_smoother_1, particle_group, e_dofs, b_dofs)
    (hamiltonian_splitting.jl:20-86).

    e_dofs / b_dofs and particle_group.array are aliased like in the
    reference. With sync=True (default) every operator uploads them before
    it runs and writes them back after it, so the caller sees the
    reference's in-place semantics. With sync=False the state stays in GPU
    memory between calls; call to_host() / to_device() explicitly.
    """

    def __init__(self, D: int, V: int, maxwell_solver: Maxwell1DFEM,
                 kernel_smoother_0: ParticleMeshCoupling1D,
                 kernel_smoother_1: ParticleMeshCoupling1D, particle_group: ParticleGroup,
                 e_dofs, b_dofs, sync: bool = True):
        if (D, V) != particle_group.dims:
            raise AssertionError("dims == particle_group.dims")
        if kernel_smoother_0.n_dofs != kernel_smoother_1.n_dofs:
            raise AssertionError("kernel_smoother_0.n_dofs == kernel_smoother_1.n_dofs")
        if D != 1:
            raise ValueError("HamiltonianSplitting: only D = 1 has a reference implementation")
        self.dims = (D, V)
        self.maxwell_solver = maxwell_solver
        self.kernel_smoother_0 = kernel_smoother_0
        self.kernel_smoother_1 = kernel_smoother_1
        self.particle_group = particle_group
        self.spline_degree = 3
        self.Lx = maxwell_solver.Lx
        self.x_min = maxwell_solver.xmin
        self.delta_x = self.Lx / kernel_smoother_1.n_dofs
        self.e_dofs = e_dofs
        self.b_dofs = b_dofs
        nx = kernel_smoother_0.n_dofs
        self.j_dofs = [np.zeros(nx), np.zeros(nx)]
        self.sync = sync
        _ON_GPU[particle_group] = weakref.ref(self)
        self.kernels_per_step = 1
        self.arithmetic = "reference"
        pg = particle_group
        mesh = maxwell_solver._mesh
        h = ctypes.c_void_p()
        check(_lib.load().gempic_splitting_create(
            nx, mesh.xmin, mesh.xmax, maxwell_solver.s_deg_0, kernel_smoother_0.spline_degree,
            kernel_smoother_0.smoothing, kernel_smoother_1.spline_degree,
            kernel_smoother_1.smoothing, V, pg.n_particles, pg.n_weights, pg.charge, pg.mass,
            pg.common_weight, ctypes.byref(h)))
        self._h = h
        self.to_device()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().gempic_splitting_destroy(h)
            self._h = None

    # -- host <-> device
    def to_device(self, particles: bool = True, fields: bool = True):
        L = _lib.load()
        if particles:
            a = self.particle_group.array
            assert a.flags.f_contiguous
            check(L.gempic_set_particles(self._h, ptr(a), a.shape[0], a.shape[1]))
        if fields:
            e2 = self.e_dofs[1] if self.dims[1] == 2 else None
            check(L.gempic_set_fields(self._h, ptr(self.e_dofs[0]), ptr(e2), ptr(self.b_dofs)))

    def to_host(self, particles: bool = True, fields: bool = True):
        L = _lib.load()
        if particles:
            a = self.particle_group.array
            check(L.gempic_get_particles(self._h, ptr(a), a.shape[0], a.shape[1]))
        if fields:
            e2 = self.e_dofs[1] if self.dims[1] == 2 else None
            check(L.gempic_get_fields(self._h, ptr(self.e_dofs[0]), ptr(e2), ptr(self.b_dofs)))
            check(L.gempic_get_j(self._h, ptr(self.j_dofs[0]), ptr(self.j_dofs[1])))

    def _run(self, fn, *args):
        if self.sync:
            self.to_device()
        check(fn(self._h, *args))
        if self.sync:
            self.to_host()

    def profile(self, enable: bool = True):
        check(_lib.load().gempic_profile_enable(self._h, int(enable)))

    def profile_read(self, kernel_id: int):
        ms = np.zeros(1)
        cnt = ctypes.c_int64()
        check(_lib.load().gempic_profile_read(self._h, kernel_id, ptr(ms), ctypes.byref(cnt)))
        return float(ms[0]), int(cnt.value)

    def set_sort_interval(self, k: int):
        check(_lib.load().gempic_set_sort_interval(self._h, int(k)))

    def set_launch(self, blocks: int):
        """Workgroups per particle pass (0 = automatic)."""
        check(_lib.load().gempic_set_launch(self._h, int(blocks)))

    def set_fusion(self, kernels_per_step: int):
        """The fused Strang step in one particle kernel (default) or two."""
        check(_lib.load().gempic_set_fusion(self._h, int(kernels_per_step)))
        self.kernels_per_step = int(kernels_per_step)

    def set_arithmetic(self, arithmetic: str):
        """"reference" (default): every product and sum rounded as in the
        reference's Julia code, bitwise its per-particle results. "fused":
        the 1d2v passes evaluate the same quantities with closed-form spline
        bases and fused multiply-adds (equal to rounding, not bitwise)."""
        code = {"reference": 0, "fused": 1}.get(arithmetic)
        if code is None:
            raise ValueError('arithmetic must be "reference" or "fused"')
        check(_lib.load().gempic_set_arithmetic(self._h, code))
        self.arithmetic = arithmetic

    def autotune(self, dt: float, n_steps: int | None = None,
                 blocks=AUTOTUNE_BLOCKS, sort_intervals=AUTOTUNE_SORT_INTERVALS,
                 cache: str | None = None) -> dict:
        """Pick the workgroups per particle pass and the sort interval by timing.

        Stands in for the reference's only launch knob, the Julia thread count
        (hamiltonian_splitting.jl:62). Every (blocks, sort interval) pair runs
        n_steps fused Strang steps of dt on this splitting's own state
        (gempic_autotune), which is restored after each trial; the fastest
        pair stays set. n_steps defaults to twice the largest sort interval
        (at least 20). With `cache` (a JSON file path) a result found under
        the same particle count, grid, degrees, dt and candidates is applied
        without timing, and a new one is stored there.
        """
        blocks = [int(b) for b in blocks]
        sort_intervals = [int(k) for k in sort_intervals]
        if n_steps is None:
            n_steps = max(20, 2 * max(sort_intervals))
        key = autotune_key(self.dims[1], self.particle_group.n_particles,
                           self.kernel_smoother_0.n_dofs, self.kernel_smoother_0.spline_degree,
                           self.kernel_smoother_1.spline_degree, dt, n_steps, blocks,
                           sort_intervals, self.kernels_per_step, self.arithmetic, device_info())
        if cache is not None:
            hit = autotune_cache_lookup(cache, key)
            if hit is not None:
                self.set_launch(hit["blocks"])
                self.set_sort_interval(hit["sort_interval"])
                return dict(hit, cached=True)
        if self.sync:
            self.to_device()
        cb = (ctypes.c_int * len(blocks))(*blocks)
        ci = (ctypes.c_int64 * len(sort_intervals))(*sort_intervals)
        ms = np.zeros(len(blocks) * len(sort_intervals))
        bb = ctypes.c_int()
        bi = ctypes.c_int64()
        check(_lib.load().gempic_autotune(self._h, float(dt), int(n_steps), cb, len(blocks), ci,
                                          len(sort_intervals), ctypes.byref(bb),
                                          ctypes.byref(bi), ptr(ms)))
        res = {"blocks": int(bb.value), "sort_interval": int(bi.value),
               "ms_per_step": ms.reshape(len(blocks), len(sort_intervals)).tolist(),
               "candidates": {"blocks": blocks, "sort_intervals": sort_intervals},
               "n_steps": int(n_steps)}
        if cache is not None:
            autotune_cache_store(cache, key, res)
        return dict(res, cached=False)


def operatorHB(h: HamiltonianSplitting, dt: float):
    """operatorHB (hamiltonian_splitting_1d2v.jl:234-236; _1d1v.jl:113-125)."""
    h._run(_lib.load().gempic_operator_hb, float(dt))


def operatorHE(h: HamiltonianSplitting, dt: float):
    """operatorHE (hamiltonian_splitting_1d2v.jl:191-219)."""
    h._run(_lib.load().gempic_operator_he, float(dt))


def operatorHp1(h: HamiltonianSplitting, dt: float):
    """operatorHp1 (hamiltonian_splitting_1d2v.jl:41-112; _1d1v.jl:63-98)."""
    h._run(_lib.load().gempic_operator_hp1, float(dt))


def operatorHp2(h: HamiltonianSplitting, dt: float):
    """operatorHp2 (hamiltonian_splitting_1d2v.jl:129-176)."""
    h._run(_lib.load().gempic_operator_hp2, float(dt))


def strang_splitting_(h: HamiltonianSplitting, dt: float, number_steps: int,
                      substep: bool = False):
    """strang_splitting! (hamiltonian_splitting.jl:98-108; _1d1v.jl:11-21).
    The default fused path gives the same particle states as the sub-step
    sequence (substep=True)."""
    if isinstance(h, HamiltonianSplittingBoris):
        # hamiltonian_splitting_boris.jl:132-177 (one particle pass per step)
        h._run(_lib.load().gempic_boris_strang_splitting, float(dt), int(number_steps), 0)
        return
    flags = _lib.STRANG_SUBSTEP if substep else _lib.STRANG_FUSED
    h._run(_lib.load().gempic_strang_splitting, float(dt), int(number_steps), flags)


class Ensemble:
    """Independent HamiltonianSplitting{1,2}s advanced together: one
    particle-pass launch and one field-part launch per Strang step for all
    of them (no counterpart in the reference: the batched form of
    strang_splitting!). Each splitting keeps its own particles, grid,
    domain and fields, and its results equal a stand-alone
    strang_splitting_ call's bit for bit."""

    def __init__(self, splittings):
        self.splittings = list(splittings)
        if not self.splittings:
            raise ValueError("an ensemble needs at least one splitting")
        arr = (ctypes.c_void_p * len(self.splittings))(*[h._h.value for h in self.splittings])
        e = ctypes.c_void_p()
        check(_lib.load().gempic_ensemble_create(len(self.splittings), arr, ctypes.byref(e)))
        self._e = e

    def __del__(self):
        e = getattr(self, "_e", None)
        if e is not None and e.value:
            _lib.load().gempic_ensemble_destroy(e)
            self._e = None


def ensemble_strang_splitting_(ens: Ensemble, dt, number_steps: int, graph: bool = False):
    """strang_splitting!(h, dt_i, number_steps) for every splitting h of ens
    (dt: one value for all, or one per splitting). graph=True replays the
    regular steps as a HIP graph of the step's launches (the same results)."""
    dts = np.broadcast_to(np.asarray(dt, dtype=np.float64), (len(ens.splittings),)).copy()
    for h in ens.splittings:
        if h.sync:
            h.to_device()
    check(_lib.load().gempic_ensemble_strang_splitting(ens._e, ptr(dts), int(number_steps),
                                                       _lib.GRAPH if graph else 0))
    for h in ens.splittings:
        if h.sync:
            h.to_host()


def ensemble_autotune(ens: Ensemble, dt, n_steps: int = 20,
                      particles_per_workgroup=(2048, 4096, 8192, 16384)) -> dict:
    """Times n_steps batched steps for each candidate number of particles
    per workgroup (each instance's pass then takes ceil(N / ppw)
    workgroups), twice each from the same state, which is restored; the
    fastest stays set on the member splittings. Returns the choice and the
    ms per step of every candidate."""
    m = len(ens.splittings)
    dts = np.broadcast_to(np.asarray(dt, dtype=np.float64), (m,)).copy()
    for h in ens.splittings:
        if h.sync:
            h.to_device()
    cand = (ctypes.c_int * len(particles_per_workgroup))(*[int(p) for p in particles_per_workgroup])
    best = ctypes.c_int(0)
    ms = np.zeros(len(particles_per_workgroup))
    check(_lib.load().gempic_ensemble_autotune(ens._e, ptr(dts), int(n_steps), cand,
                                               len(particles_per_workgroup), ctypes.byref(best),
                                               ptr(ms)))
    return {"particles_per_workgroup": int(best.value), "n_steps": int(n_steps),
            "candidates": [int(p) for p in particles_per_workgroup],
            "ms_per_step": [float(v) for v in ms]}


def ensemble_strang_splitting_adjoint_(ens: Ensemble, dt, number_steps: int,
                                       checkpoint_every: int = 50, loss_weight=None,
                                       fetch: bool = True):
    """strang_splitting_adjoint_ for every splitting of ens, batched (no
    cotangents of the final state; loss_weight defaults to each dt). Returns
    one AdjointResult per splitting (or None with fetch=False)."""
    m = len(ens.splittings)
    dts = np.broadcast_to(np.asarray(dt, dtype=np.float64), (m,)).copy()
    wl = dts.copy() if loss_weight is None else \
        np.broadcast_to(np.asarray(loss_weight, dtype=np.float64), (m,)).copy()
    for h in ens.splittings:
        if h.sync:
            h.to_device()
    check(_lib.load().gempic_ensemble_strang_splitting_adjoint(
        ens._e, ptr(dts), int(number_steps), int(checkpoint_every), ptr(wl), 0))
    for h in ens.splittings:
        if h.sync:
            h.to_host()
    return [get_adjoint(h) for h in ens.splittings] if fetch else None


class AdjointResult:
    """Loss and cotangents of the state a reverse-mode call started from:
    g_particles has the particle array's layout (rows x, v1, v2; the weight
    row is zero), g_e1, g_e2, g_b the fields'."""

    def __init__(self, loss, g_particles, g_e1, g_e2, g_b, g_length=0.0, g_weight_sum=0.0):
        self.loss, self.g_particles = loss, g_particles
        self.g_e1, self.g_e2, self.g_b = g_e1, g_e2, g_b
        # explicit d loss / d Lx (state held fixed) and sum_i d loss / d w_i
        self.g_length, self.g_weight_sum = g_length, g_weight_sum


def strang_splitting_adjoint_(h: HamiltonianSplitting, dt: float, number_steps: int,
                              checkpoint_every: int = 50, loss_weight: float | None = None,
                              ct_particles=None, ct_e1=None, ct_e2=None, ct_b=None,
                              fetch: bool = True):
    """Reverse mode of strang_splitting!(h, dt, number_steps) for 1d2v (the
    reference has no derivatives; this is the pullback of an rrule). Runs the
    steps forward, leaving h in the state strang_splitting_ would (without
    the periodic particle sort), and returns the gradient of

        loss_weight * sum_s 1/2 l2norm_squared(B_s, p-1) + <ct, final state>

    with respect to the particles and fields before the call (loss_weight
    defaults to dt: the time-integrated magnetic energy). Checkpoints every
    `checkpoint_every` steps bound the memory; the intervals between them are
    re-run on the GPU. With fetch=False the result stays on the GPU
    (get_adjoint reads it)."""
    if h.dims != (1, 2) or isinstance(h, HamiltonianSplittingBoris):
        raise ValueError("the reverse mode covers HamiltonianSplitting{1,2}")
    wl = float(dt) if loss_weight is None else float(loss_weight)
    cp = None
    if ct_particles is not None:
        cp = np.asfortranarray(ct_particles, dtype=np.float64)
        if cp.shape != h.particle_group.array.shape:
            raise AssertionError("ct_particles must have the particle array's shape")
    f = [None if c is None else np.ascontiguousarray(c, dtype=np.float64) for c in (ct_e1, ct_e2, ct_b)]
    if h.sync:
        h.to_device()
    ld = h.particle_group.array.shape[0]
    check(_lib.load().gempic_strang_splitting_adjoint(h._h, float(dt), int(number_steps),
                                                      int(checkpoint_every), wl, ptr(cp), ld,
                                                      ptr(f[0]), ptr(f[1]), ptr(f[2]), 0))
    if h.sync:
        h.to_host()
    return get_adjoint(h) if fetch else None


def get_adjoint(h: HamiltonianSplitting) -> AdjointResult:
    """The result of the last strang_splitting_adjoint_ on h."""
    a = h.particle_group.array
    g = np.zeros_like(a, order="F")
    nx = h.kernel_smoother_0.n_dofs
    ge1, ge2, gb, loss = np.zeros(nx), np.zeros(nx), np.zeros(nx), np.zeros(1)
    check(_lib.load().gempic_get_adjoint(h._h, ptr(loss), ptr(g), a.shape[0], a.shape[1],
                                         ptr(ge1), ptr(ge2), ptr(gb)))
    gl, gw = np.zeros(1), np.zeros(1)
    check(_lib.load().gempic_get_adjoint_domain(h._h, ptr(gl), ptr(gw)))
    return AdjointResult(float(loss[0]), g, ge1, ge2, gb, float(gl[0]), float(gw[0]))


def solve_poisson_adjoint(h: HamiltonianSplitting, ct_e) -> np.ndarray:
    """Reverse mode of solve_poisson! (diagnostics.jl:15-31) in the particle
    positions: d <ct_e, E> / dx_i, in the particle group's order."""
    if h.sync:
        h.to_device(fields=False)
    g = np.zeros(h.particle_group.n_particles)
    check(_lib.load().gempic_solve_poisson_adjoint(
        h._h, ptr(np.ascontiguousarray(ct_e, dtype=np.float64)), ptr(g)))
    return g


# ---------------------------------------------------------------------------
# particle_sampling.jl

def sample_(pg: ParticleGroup, alpha: float, k: float, sigma: float, mesh: OneDGrid):
    """sample!(pg, alpha, k, sigma, mesh) (particle_sampling.jl:266-311) for
    ParticleGroup{1,1} and {1,2}, computed on the GPU into pg.array."""
    D, V = pg.dims
    if D != 1 or pg.n_weights != 1:
        raise TypeError("sample!(pg, alpha, k, sigma, mesh) needs a ParticleGroup{1,V} with one weight")
    a = pg.array
    check(_lib.load().gempic_sample_landau(V, pg.n_particles, float(alpha), float(k), float(sigma),
                                           float(mesh.dimx), ptr(a), a.shape[0]))


def sample_particles_(h: "HamiltonianSplitting", alpha: float, k: float, sigma: float):
    """sample! straight into a splitting's GPU particles (no host round
    trip); with sync=True the particle group is updated too."""
    check(_lib.load().gempic_sample_landau_particles(h._h, float(alpha), float(k), float(sigma)))
    if h.sync:
        h.to_host(fields=False)


# ---------------------------------------------------------------------------
# hamiltonian_splitting_boris.jl

class HamiltonianSplittingBoris(HamiltonianSplitting):
    """HamiltonianSplittingBoris(maxwell_solver, kernel_smoother_0,
    kernel_smoother_1, particle_group, e_dofs, b_dofs)
    (hamiltonian_splitting_boris.jl:23-88), for a ParticleGroup{1,2}.
    e_dofs_mid and b_dofs_mid start at zero (:50-60) and live on the GPU; with
    sync=True they are mirrored here after every call."""

    def __init__(self, maxwell_solver: Maxwell1DFEM, kernel_smoother_0: ParticleMeshCoupling1D,
                 kernel_smoother_1: ParticleMeshCoupling1D, particle_group: ParticleGroup,
                 e_dofs, b_dofs, sync: bool = True):
        super().__init__(1, 2, maxwell_solver, kernel_smoother_0, kernel_smoother_1,
                         particle_group, e_dofs, b_dofs, sync=sync)
        nx = kernel_smoother_1.n_dofs
        self.e_dofs_mid = [np.zeros(nx), np.zeros(nx)]
        self.b_dofs_mid = np.zeros(nx)
        self.cell_integrals_1 = np.array([0.5, 2.0, 0.5]) / 3.0
        self.cell_integrals_0 = np.array([1.0, 11.0, 11.0, 1.0]) / 24.0

    def to_device(self, particles: bool = True, fields: bool = True):
        super().to_device(particles, fields)
        if fields and hasattr(self, "b_dofs_mid"):
            check(_lib.load().gempic_boris_set_mid(self._h, ptr(self.e_dofs_mid[0]),
                                                   ptr(self.e_dofs_mid[1]), ptr(self.b_dofs_mid)))

    def to_host(self, particles: bool = True, fields: bool = True):
        super().to_host(particles, fields)
        if fields:
            check(_lib.load().gempic_boris_get_mid(self._h, ptr(self.e_dofs_mid[0]),
                                                   ptr(self.e_dofs_mid[1]), ptr(self.b_dofs_mid)))


def staggering_(splitting: HamiltonianSplittingBoris, dt: float):
    """staggering! (hamiltonian_splitting_boris.jl:99-122)."""
    splitting._run(_lib.load().gempic_boris_staggering, float(dt))


def push_v_epart_(splitting: HamiltonianSplittingBoris, dt: float):
    """push_v_epart! (hamiltonian_splitting_boris.jl:189-204)."""
    splitting._run(_lib.load().gempic_boris_push_v_epart, float(dt))


def push_v_bpart_(splitting: HamiltonianSplittingBoris, dt: float):
    """push_v_bpart! (hamiltonian_splitting_boris.jl:211-233)."""
    splitting._run(_lib.load().gempic_boris_push_v_bpart, float(dt))


def push_x_accumulate_j_(splitting: HamiltonianSplittingBoris, dt: float):
    """push_x_accumulate_j! (hamiltonian_splitting_boris.jl:250-288)."""
    splitting._run(_lib.load().gempic_boris_push_x_accumulate_j, float(dt))


# ---------------------------------------------------------------------------
# diagnostics.jl

def _splitting_of(particle_group, kernel_smoother_0=None, maxwell_solver=None):
    """The GPU splitting that holds this particle group (and, when given,
    these smoother and solver objects)."""
    ref = _ON_GPU.get(particle_group)
    h = ref() if ref is not None else None
    if h is None or (kernel_smoother_0 is not None and h.kernel_smoother_0 is not kernel_smoother_0) \
            or (maxwell_solver is not None and h.maxwell_solver is not maxwell_solver):
        raise ValueError("the particle group is not held by a HamiltonianSplitting on the GPU "
                         "with this smoother and solver (the path has no CPU fallback)")
    return h


def solve_poisson_(efield_dofs, *args):
    """solve_poisson!(efield_dofs, particle_group, kernel_smoother_0,
    maxwell_solver, rho) (diagnostics.jl:15-31), the reference's signature:
    the group's particles are deposited on the GPU by the splitting that holds
    them. solve_poisson_(efield_dofs, splitting, rho) names the splitting
    directly."""
    if len(args) == 4:
        pg, ks0, mx, rho = args
        h = _splitting_of(pg, ks0, mx)
    elif len(args) == 2:
        h, rho = args
    else:
        raise TypeError("solve_poisson_(efield_dofs, particle_group, kernel_smoother_0, "
                        "maxwell_solver, rho)")
    if h.sync:
        h.to_device(fields=False)
    check(_lib.load().gempic_solve_poisson(h._h, ptr(efield_dofs), ptr(rho)))
    return efield_dofs


class TimeHistoryDiagnostics:
    """TimeHistoryDiagnostics(particle_group, maxwell_solver,
    kernel_smoother_0, kernel_smoother_1) (diagnostics.jl:128-170): rows of
    the 11 columns, filled by write_step_ on the GPU splitting that holds the
    particle group. TimeHistoryDiagnostics(splitting) names it directly."""

    COLUMNS = ("Time", "KineticEnergy", "Momentum1", "Momentum2", "PotentialEnergyE1",
               "PotentialEnergyE2", "PotentialEnergyB3", "Transfer", "VVB", "Poynting",
               "ErrorPoisson")

    def __init__(self, *args):
        if len(args) == 1:
            self.splitting = args[0]
        elif len(args) == 4:
            pg, mx, ks0, ks1 = args
            self.splitting = _splitting_of(pg, ks0, mx)
            if self.splitting.kernel_smoother_1 is not ks1:
                raise ValueError("kernel_smoother_1 is not the splitting's")
        else:
            raise TypeError("TimeHistoryDiagnostics(particle_group, maxwell_solver, "
                            "kernel_smoother_0, kernel_smoother_1)")
        self.particle_group = self.splitting.particle_group
        self.data = []


def write_step_(thdiag: TimeHistoryDiagnostics, time: float, degree: int, *fields,
                efield_dofs_n=None, efield_poisson=None):
    """write_step!(thdiag, time, degree, efield_dofs, bfield_dofs,
    efield_dofs_n, efield_poisson) (diagnostics.jl:186-250), the reference's
    signature: the given fields are uploaded and the row computed on the GPU.
    Without the four field arguments the splitting's current fields are used
    (efield_dofs_n and efield_poisson by keyword; defaults: the current
    fields, and Poisson recomputed)."""
    h = thdiag.splitting
    if len(fields) == 4:
        efield_dofs, bfield_dofs, efield_dofs_n, efield_poisson = fields
    elif fields:
        raise TypeError("write_step_(thdiag, time, degree, efield_dofs, bfield_dofs, "
                        "efield_dofs_n, efield_poisson)")
    if h.sync:
        h.to_device()
    if len(fields) == 4:
        check(_lib.load().gempic_set_fields(
            h._h, ptr(np.ascontiguousarray(efield_dofs[0], dtype=np.float64)),
            ptr(np.ascontiguousarray(efield_dofs[1], dtype=np.float64)),
            ptr(np.ascontiguousarray(bfield_dofs, dtype=np.float64))))
    out = np.zeros(11)
    e1n = e2n = None
    if efield_dofs_n is not None:
        e1n = np.ascontiguousarray(efield_dofs_n[0], dtype=np.float64)
        e2n = np.ascontiguousarray(efield_dofs_n[1], dtype=np.float64)
    ep = None if efield_poisson is None else np.ascontiguousarray(efield_poisson, dtype=np.float64)
    check(_lib.load().gempic_write_step(h._h, float(time), int(degree), ptr(e1n), ptr(e2n),
                                        ptr(ep), ptr(out)))
    thdiag.data.append(out)
    return out


# 2D coupling and field solver (defined in core2d, dispatched from the verbs above)
from .core2d import (  # noqa: E402
    ParticleMeshCoupling2D, TwoDMaxwell, add_charge_2d_, evaluate_2d, compute_e_from_rho_2d_,
    add_charge_pp_2d_, evaluate_pp_2d,
    compute_e_from_j_2d_, compute_e_from_b_2d_, compute_b_from_e_2d_, inner_product_2d)
from .core2d import __all__ as _all2d  # noqa: E402
from .core2d import *  # noqa: E402,F401,F403

__all__ = __all__ + list(_all2d)
