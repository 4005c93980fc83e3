"""Host-side mirror of GEMPIC.jl's 2D particle-mesh coupling and 2D Maxwell
solver over the HIP library (particle_mesh_coupling_2d.jl, maxwell_2d_fem.jl,
poisson_2d_fem.jl, linear_solver_spline_mass_2d.jl, mesh.jl).

Grid arrays are float64 vectors of nx*ny entries, x fastest, exactly the
reference's vec(reshape(..., nx, ny)). The verbs that share a name with the
1D ones (add_charge_, evaluate, compute_e_from_b_, inner_product, ...) are
dispatched here from core by the type of their solver/coupling argument, as
Julia dispatches them. Set-up arithmetic (mass lines, quadrature right-hand
sides of user functions) stays on the host, as in the reference; every
field and particle operation runs on the GPU.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import check, ptr
from .core import ParticleGroup, _SMOOTHING, _bspline_basis

__all__ = [
    "TwoDGrid", "ParticleMeshCoupling2D", "evaluate_multiple", "TwoDMaxwell",
    "TwoDLinearSolverSplineMass", "solve", "multiply_mass_2dkron_", "compute_rho_from_e_",
    "compute_fem_rhs_", "l2projection", "spline_fem_mass_line", "spline_fem_compute_mass_eig",
    "add_charge_device_", "add_charge_device_noalloc_", "evaluate_device", "evaluate_multiple_device", "pmc2d_sync",
    "b_to_pp_2d_", "HamiltonianSplitting2D3V",
]


# ---------------------------------------------------------------------------
# mesh.jl

class TwoDGrid:
    """TwoDGrid(xmin, xmax, nx, ymin, ymax, ny) or TwoDGrid(dimx, nx, dimy, ny)
    (mesh.jl:17-45)."""

    def __init__(self, *args):
        if len(args) == 4:
            dimx, nx, dimy, ny = args
            xmin, xmax, ymin, ymax = 0.0, dimx, 0.0, dimy
        elif len(args) == 6:
            xmin, xmax, nx, ymin, ymax, ny = args
        else:
            raise TypeError("TwoDGrid(xmin, xmax, nx, ymin, ymax, ny) or TwoDGrid(dimx, nx, dimy, ny)")
        self.nx, self.ny = int(nx), int(ny)
        self.xmin, self.xmax = float(xmin), float(xmax)
        self.ymin, self.ymax = float(ymin), float(ymax)
        self.dimx = self.xmax - self.xmin
        self.dimy = self.ymax - self.ymin
        self.x = np.linspace(self.xmin, self.xmax, self.nx + 1)
        self.y = np.linspace(self.ymin, self.ymax, self.ny + 1)
        self.dx = self.dimx / self.nx
        self.dy = self.dimy / self.ny


def _grid(a, nn, name):
    """A float64 contiguous vector of nn entries (checked before the C call,
    which reads and writes exactly nn doubles)."""
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64):
        raise TypeError(f"{name}: float64 numpy array expected")
    if a.size != nn or not (a.flags.c_contiguous or a.flags.f_contiguous):
        raise ValueError(f"{name}: contiguous array of nx*ny = {nn} entries expected, "
                         f"got shape {a.shape}")
    return a


def _vec(x):
    return np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64)))


# ---------------------------------------------------------------------------
# particle_mesh_coupling_2d.jl

class ParticleMeshCoupling2D:
    """ParticleMeshCoupling2D(pg, grid, degree, smoothing_type)
    (particle_mesh_coupling_2d.jl:12-45)."""

    def __init__(self, pg: ParticleGroup, grid: TwoDGrid, degree: int, smoothing_type: str):
        # {2,2} as in the reference; {2,3} for the 2D3V splitting, which
        # couples the same 2D splines to particles with three velocities
        if tuple(pg.dims) not in ((2, 2), (2, 3)):
            raise TypeError("ParticleMeshCoupling2D needs a ParticleGroup{2,2} or {2,3}")
        if smoothing_type not in _SMOOTHING:
            raise ValueError(f"Smoothing Type {smoothing_type} not implemented for "
                             "kernel_smoother_spline_2d.")
        self.grid = grid
        self.npart = pg.n_particles
        self.n_span = int(degree) + 1
        self.degree = int(degree)
        self.smoothing = _SMOOTHING[smoothing_type]
        self.scaling = 1.0 / (grid.dx * grid.dy) if smoothing_type == "collocation" else 1.0

    def _args(self):
        g = self.grid
        return (g.nx, g.ny, g.xmin, g.xmax, g.ymin, g.ymax, self.degree, self.smoothing)


def add_charge_2d_(rho, pm: ParticleMeshCoupling2D, xp, yp, wp):
    """add_charge! (particle_mesh_coupling_2d.jl:94-105), batched over arrays
    of positions and weights (scalars are one particle)."""
    x, y = _vec(xp), _vec(yp)
    w = _vec(wp) * np.ones_like(x)
    if y.size != x.size or w.size != x.size:
        raise ValueError("xp, yp and wp must have the same length")
    _grid(rho, pm.grid.nx * pm.grid.ny, "rho")
    check(_lib.load().gempic_pmc2d_add_charge(*pm._args(), x.size, ptr(x), ptr(y), ptr(w),
                                              ptr(rho)))


def evaluate_2d(pm: ParticleMeshCoupling2D, xp, yp, field_dofs):
    """evaluate (particle_mesh_coupling_2d.jl:189-203)."""
    x, y = _vec(xp), _vec(yp)
    if y.size != x.size:
        raise ValueError("xp and yp must have the same length")
    f = _grid(np.ascontiguousarray(field_dofs, dtype=np.float64), pm.grid.nx * pm.grid.ny,
              "field_dofs")
    out = np.zeros_like(x)
    check(_lib.load().gempic_pmc2d_evaluate(*pm._args(), x.size, ptr(x), ptr(y), ptr(f),
                                            ptr(out)))
    return float(out[0]) if np.ndim(xp) == 0 else out


def add_charge_pp_2d_(rho, pm: ParticleMeshCoupling2D, xp, yp, wp):
    """add_charge_pp! (particle_mesh_coupling_2d.jl:127-148), batched."""
    x, y = _vec(xp), _vec(yp)
    w = _vec(wp) * np.ones_like(x)
    if y.size != x.size or w.size != x.size:
        raise ValueError("xp, yp and wp must have the same length")
    _grid(rho, pm.grid.nx * pm.grid.ny, "rho")
    check(_lib.load().gempic_pmc2d_add_charge_pp(*pm._args(), x.size, ptr(x), ptr(y), ptr(w),
                                                 ptr(rho)))


def b_to_pp_2d_(pp, pm: ParticleMeshCoupling2D, b):
    """b_to_pp_2d!(pp, spl1, spl2, b) (splinepp.jl:392-399) with the
    coupling's two SplinePP of its degree: pp is the ((degree+1)^2, nx*ny)
    column-major matrix (a Fortran-ordered array, or one of that size)."""
    g = pm.grid
    ld = (pm.degree + 1) ** 2
    bb = _grid(np.ascontiguousarray(b, dtype=np.float64), g.nx * g.ny, "b")
    if not (isinstance(pp, np.ndarray) and pp.dtype == np.float64 and pp.size == ld * g.nx * g.ny
            and (pp.flags.f_contiguous if pp.ndim == 2 else pp.flags.c_contiguous)):
        raise ValueError(f"pp: Fortran-ordered float64 ({ld}, {g.nx * g.ny}) array expected")
    check(_lib.load().gempic_b_to_pp_2d(pm.degree, g.nx, g.ny, ptr(bb), ptr(pp)))
    return pp


def evaluate_pp_2d(pm: ParticleMeshCoupling2D, xp, yp, pp):
    """evaluate_pp (particle_mesh_coupling_2d.jl:162-178); a position outside
    the table raises IndexError (Julia's BoundsError) through ValueError."""
    x, y = _vec(xp), _vec(yp)
    if y.size != x.size:
        raise ValueError("xp and yp must have the same length")
    g = pm.grid
    ld = (pm.degree + 1) ** 2
    if not (isinstance(pp, np.ndarray) and pp.dtype == np.float64 and pp.size == ld * g.nx * g.ny):
        raise ValueError(f"pp: float64 ({ld}, {g.nx * g.ny}) matrix expected")
    pp = np.asfortranarray(pp)
    out = np.zeros_like(x)
    check(_lib.load().gempic_pmc2d_evaluate_pp(*pm._args(), x.size, ptr(x), ptr(y), ptr(pp),
                                               ptr(out)))
    return float(out[0]) if np.ndim(xp) == 0 else out


def evaluate_multiple(pm: ParticleMeshCoupling2D, position, field_dofs):
    """evaluate_multiple (particle_mesh_coupling_2d.jl:214-231): position =
    (x, y), field_dofs = (field1, field2); returns (value1, value2)."""
    xp, yp = position
    x, y = _vec(xp), _vec(yp)
    if y.size != x.size:
        raise ValueError("x and y must have the same length")
    nn = pm.grid.nx * pm.grid.ny
    f1 = _grid(np.ascontiguousarray(field_dofs[0], dtype=np.float64), nn, "field_dofs[1]")
    f2 = _grid(np.ascontiguousarray(field_dofs[1], dtype=np.float64), nn, "field_dofs[2]")
    o1, o2 = np.zeros_like(x), np.zeros_like(x)
    check(_lib.load().gempic_pmc2d_evaluate_multiple(*pm._args(), x.size, ptr(x), ptr(y),
                                                     ptr(f1), ptr(f2), ptr(o1), ptr(o2)))
    if np.ndim(xp) == 0:
        return float(o1[0]), float(o2[0])
    return o1, o2


# device-resident forms: arrays are gempic_amd.device.DeviceArray (GPU memory
# of the library's own HIP runtime), streams gempic_amd.device.Stream

def _dev(t, n, name):
    from .device import DeviceArray
    if not isinstance(t, DeviceArray):
        raise TypeError(f"{name}: DeviceArray expected")
    if t.size != n:
        raise ValueError(f"{name}: {n} entries expected, got {t.size}")
    return t.ptr


def _flags(deterministic, async_):
    return (_lib.DETERMINISTIC if deterministic else 0) | (_lib.ASYNC if async_ else 0)


def _stream(stream):
    return None if stream is None else stream.handle


def add_charge_device_(rho, pm: ParticleMeshCoupling2D, x, y, wp, deterministic=False,
                       async_=False, stream=None):
    """add_charge! (particle_mesh_coupling_2d.jl:94-105) on particles in GPU
    memory: rho (nx*ny), x, y, wp (n) are DeviceArrays. The work runs on
    `stream` (the null stream by default)."""
    n = x.size
    args = (_dev(x, n, "x"), _dev(y, n, "y"), _dev(wp, n, "wp"),
            _dev(rho, pm.grid.nx * pm.grid.ny, "rho"))
    check(_lib.load().gempic_pmc2d_add_charge_dev(*pm._args(), n, *args,
                                                  _flags(deterministic, async_), _stream(stream)))


def add_charge_device_noalloc_(rho, pm: ParticleMeshCoupling2D, x, y, wp, deterministic=False,
                               async_=False, stream=None):
    """add_charge_device_ through the allocation-free entry point: the
    deposit's scratch is the library's fixed per-GPU array, so the call
    allocates no device memory (calls overlapping on one GPU race)."""
    n = x.size
    args = (_dev(x, n, "x"), _dev(y, n, "y"), _dev(wp, n, "wp"),
            _dev(rho, pm.grid.nx * pm.grid.ny, "rho"))
    check(_lib.load().gempic_pmc2d_add_charge_dev_noalloc(
        *pm._args(), n, *args, _flags(deterministic, async_), _stream(stream)))


def evaluate_device(pm: ParticleMeshCoupling2D, x, y, field_dofs, out, async_=False,
                    stream=None):
    """evaluate (particle_mesh_coupling_2d.jl:189-203) into `out` on the GPU."""
    n = x.size
    args = (_dev(x, n, "x"), _dev(y, n, "y"),
            _dev(field_dofs, pm.grid.nx * pm.grid.ny, "field_dofs"), _dev(out, n, "out"))
    check(_lib.load().gempic_pmc2d_evaluate_dev(*pm._args(), n, *args, _flags(False, async_),
                                                _stream(stream)))
    return out


def evaluate_multiple_device(pm: ParticleMeshCoupling2D, position, field_dofs, out,
                             async_=False, stream=None):
    """evaluate_multiple (particle_mesh_coupling_2d.jl:214-231) into the two
    tensors of `out` on the GPU."""
    x, y = position
    n = x.size
    nn = pm.grid.nx * pm.grid.ny
    args = (_dev(x, n, "x"), _dev(y, n, "y"), _dev(field_dofs[0], nn, "field_dofs[1]"),
            _dev(field_dofs[1], nn, "field_dofs[2]"), _dev(out[0], n, "out[1]"),
            _dev(out[1], n, "out[2]"))
    check(_lib.load().gempic_pmc2d_evaluate_multiple_dev(*pm._args(), n, *args,
                                                         _flags(False, async_), _stream(stream)))
    return out


def pmc2d_sync(stream=None):
    """Wait for the stream; raise errors of earlier asynchronous 2D calls."""
    check(_lib.load().gempic_pmc2d_sync(_stream(stream)))


# ---------------------------------------------------------------------------
# poisson_2d_fem.jl set-up helpers (host)

def _gauss_legendre(n: int):
    """Gauss-Legendre nodes/weights on [-1, 1] for n <= 5, the closed forms
    of FastGaussQuadrature.gausslegendre's small rules."""
    if n == 1:
        return [0.0], [2.0]
    if n == 2:
        return [-1.0 / math.sqrt(3.0), 1.0 / math.sqrt(3.0)], [1.0, 1.0]
    if n == 3:
        return ([-math.sqrt(3.0 / 5.0), 0.0, math.sqrt(3.0 / 5.0)],
                [5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0])
    if n == 4:
        a = 2.0 / 7.0 * math.sqrt(6.0 / 5.0)
        x = [-math.sqrt(3.0 / 7.0 + a), -math.sqrt(3.0 / 7.0 - a),
             math.sqrt(3.0 / 7.0 - a), math.sqrt(3.0 / 7.0 + a)]
        w0, w1 = (18.0 - math.sqrt(30.0)) / 36.0, (18.0 + math.sqrt(30.0)) / 36.0
        return x, [w0, w1, w1, w0]
    if n == 5:
        b = 2.0 * math.sqrt(10.0 / 7.0)
        x = [-math.sqrt(5.0 + b) / 3.0, -math.sqrt(5.0 - b) / 3.0, 0.0,
             math.sqrt(5.0 - b) / 3.0, math.sqrt(5.0 + b) / 3.0]
        w0 = (322.0 - 13.0 * math.sqrt(70.0)) / 900.0
        w1 = (322.0 + 13.0 * math.sqrt(70.0)) / 900.0
        return x, [w0, w1, 128.0 / 225.0, w1, w0]
    raise ValueError("Gauss-Legendre rules of 1..5 points only")


def spline_fem_mass_line(degree: int) -> np.ndarray:
    """spline_fem_mass_line (poisson_2d_fem.jl:87-106)."""
    q = degree + 1
    x, w = _gauss_legendre(q)
    x = [(xk + 1.0) * 0.5 for xk in x]
    w = [wk * 0.5 for wk in w]
    val = [_bspline_basis(degree, xk) for xk in x]       # val[k][i]
    line = [0.0] * (degree + 1)
    for j in range(1, degree + 2):
        for i in range(j, degree + 2):
            for k in range(1, q + 1):
                line[j - 1] += val[k - 1][i - 1] * val[k - 1][i - j] * w[k - 1]
    return np.array(line)


def spline_fem_compute_mass_eig(n_cells: int, degree: int, mass_line) -> np.ndarray:
    """spline_fem_compute_mass_eig (poisson_2d_fem.jl:113-126)."""
    factor = 6.283185307179586 / n_cells
    eig = np.zeros(n_cells)
    for k in range(1, n_cells + 1):
        e = float(mass_line[0])
        for j in range(1, degree + 1):
            e += float(mass_line[j]) * 2.0 * math.cos(factor * (k - 1) * j)
        eig[k - 1] = e
    return eig


# ---------------------------------------------------------------------------
# linear_solver_spline_mass_2d.jl, maxwell_2d_fem.jl

class TwoDLinearSolverSplineMass:
    """The Kronecker mass solver (linear_solver_spline_mass_2d.jl:1-14) of one
    (form, component) of a TwoDMaxwell; solved on the GPU."""

    def __init__(self, maxwell: "TwoDMaxwell", form: int, component: int, eigvals1, eigvals2):
        self._maxwell = maxwell
        self.form, self.component = form, component
        self.nx1, self.nx2 = maxwell.mesh.nx, maxwell.mesh.ny
        self.factor = 1.0
        self.eigvals1, self.eigvals2 = eigvals1, eigvals2


def solve(solver: TwoDLinearSolverSplineMass, rhs) -> np.ndarray:
    """solve (linear_solver_spline_mass_2d.jl:16-32): real(ifft2(fft2(rhs) ./
    (eig1[i] * eig2[j] * factor)))."""
    m = solver._maxwell
    r = _grid(np.ascontiguousarray(rhs, dtype=np.float64), m._nn, "rhs")
    out = np.zeros(m._nn)
    check(_lib.load().gempic_maxwell2d_solve(m._h, solver.form, solver.component, ptr(r),
                                             ptr(out)))
    return out


class TwoDMaxwell:
    """TwoDMaxwell(mesh, degree) (maxwell_2d_fem.jl:11-88) with its
    TwoDPoisson (poisson_2d_fem.jl:3-85); fields on the GPU per call."""

    def __init__(self, mesh: TwoDGrid, degree: int):
        degree = int(degree)
        self.s_deg_0 = degree
        self.s_deg_1 = degree - 1
        self.mesh = mesh
        h = ctypes.c_void_p()
        check(_lib.load().gempic_maxwell2d_create(mesh.nx, mesh.xmin, mesh.xmax, mesh.ny,
                                                  mesh.ymin, mesh.ymax, degree,
                                                  ctypes.byref(h)))
        self._h = h
        self._nn = mesh.nx * mesh.ny
        l0 = spline_fem_mass_line(self.s_deg_0)
        l1 = spline_fem_mass_line(self.s_deg_1)
        self.mass_line_0 = [l0 * mesh.dx, l0 * mesh.dy]
        self.mass_line_1 = [l1 * mesh.dx, l1 * mesh.dy]
        e0x, e0y = np.zeros(mesh.nx), np.zeros(mesh.ny)
        e1x, e1y = np.zeros(mesh.nx), np.zeros(mesh.ny)
        check(_lib.load().gempic_maxwell2d_tables(self._h, ptr(e0x), ptr(e0y), ptr(e1x),
                                                  ptr(e1y)))
        S = TwoDLinearSolverSplineMass
        self.inv_mass_1 = [S(self, 1, 1, e1x, e0y), S(self, 1, 2, e0x, e1y),
                           S(self, 1, 3, e0x, e0y)]
        self.inv_mass_2 = [S(self, 2, 1, e0x, e1y), S(self, 2, 2, e1x, e0y),
                           S(self, 2, 3, e1x, e1y)]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().gempic_maxwell2d_destroy(h)
            self._h = None


def _form_of_lines(solver: TwoDMaxwell, mass_line_1, mass_line_2):
    """(form, component) of the C ABI whose mass lines are these of the solver."""
    if mass_line_1 is solver.mass_line_0[0]:
        ax = 0
    elif mass_line_1 is solver.mass_line_1[0]:
        ax = 1
    else:
        raise ValueError("mass_line_1 must be solver.mass_line_0[1] or solver.mass_line_1[1]")
    if mass_line_2 is solver.mass_line_0[1]:
        ay = 0
    elif mass_line_2 is solver.mass_line_1[1]:
        ay = 1
    else:
        raise ValueError("mass_line_2 must be solver.mass_line_0[2] or solver.mass_line_1[2]")
    return {(0, 0): (0, 1), (1, 0): (1, 1), (0, 1): (1, 2), (1, 1): (3, 1)}[(ax, ay)]


def multiply_mass_2dkron_(c_out, solver: TwoDMaxwell, mass_line_1, mass_line_2, c_in):
    """multiply_mass_2dkron! (maxwell_2d_fem.jl:343-363) with two of the
    solver's own mass lines (the reference's only call sites)."""
    form, comp = _form_of_lines(solver, mass_line_1, mass_line_2)
    _grid(c_out, solver._nn, "c_out")
    ci = _grid(np.ascontiguousarray(c_in, dtype=np.float64), solver._nn, "c_in")
    check(_lib.load().gempic_maxwell2d_multiply_mass(solver._h, form, comp, ptr(ci), ptr(c_out)))
    return c_out


def _three(v, nn, name):
    if len(v) != 3:
        raise ValueError(f"{name}: three components expected")
    return [_grid(a, nn, f"{name}[{k + 1}]") for k, a in enumerate(v)]


def compute_e_from_b_2d_(e, solver: TwoDMaxwell, dt: float, b):
    """compute_e_from_b! (maxwell_2d_fem.jl:370-412): e[k] += dt * M1^-1 curl M2 b."""
    e = _three(e, solver._nn, "e")
    b = _three(b, solver._nn, "b")
    check(_lib.load().gempic_maxwell2d_compute_e_from_b(solver._h, float(dt),
                                                        *[ptr(a) for a in b],
                                                        *[ptr(a) for a in e]))
    return e[2]


def compute_b_from_e_2d_(b, solver: TwoDMaxwell, dt: float, e):
    """compute_b_from_e! (maxwell_2d_fem.jl:423-444)."""
    e = _three(e, solver._nn, "e")
    b = _three(b, solver._nn, "b")
    check(_lib.load().gempic_maxwell2d_compute_b_from_e(solver._h, float(dt),
                                                        *[ptr(a) for a in e],
                                                        *[ptr(a) for a in b]))


def compute_e_from_j_2d_(e, solver: TwoDMaxwell, current, component: int):
    """compute_e_from_j! (maxwell_2d_fem.jl:455-459)."""
    _grid(e, solver._nn, "e")
    j = _grid(np.ascontiguousarray(current, dtype=np.float64), solver._nn, "current")
    check(_lib.load().gempic_maxwell2d_compute_e_from_j(solver._h, ptr(j), int(component),
                                                        ptr(e)))
    return e


def compute_rho_from_e_(rho, solver: TwoDMaxwell, efield):
    """compute_rho_from_e! (maxwell_2d_fem.jl:468-499): rho = -(G^T M1 e)."""
    _grid(rho, solver._nn, "rho")
    e1 = _grid(np.ascontiguousarray(efield[0], dtype=np.float64), solver._nn, "efield[1]")
    e2 = _grid(np.ascontiguousarray(efield[1], dtype=np.float64), solver._nn, "efield[2]")
    check(_lib.load().gempic_maxwell2d_compute_rho_from_e(solver._h, ptr(e1), ptr(e2), None,
                                                          ptr(rho)))
    return rho


def compute_e_from_rho_2d_(efield, solver: TwoDMaxwell, rho):
    """compute_e_from_rho! (maxwell_2d_fem.jl:199-201 -> poisson_2d_fem.jl:237-262):
    efield[1], efield[2] from rho."""
    e1 = _grid(efield[0], solver._nn, "efield[1]")
    e2 = _grid(efield[1], solver._nn, "efield[2]")
    r = _grid(np.ascontiguousarray(rho, dtype=np.float64), solver._nn, "rho")
    check(_lib.load().gempic_maxwell2d_compute_e_from_rho(solver._h, ptr(r), ptr(e1), ptr(e2)))
    return e2


def inner_product_2d(solver: TwoDMaxwell, coefs1_dofs, coefs2_dofs, component: int,
                     form: int) -> float:
    """inner_product (maxwell_2d_fem.jl:511-575): sum(c1 .* M c2)."""
    c1 = _grid(np.ascontiguousarray(coefs1_dofs, dtype=np.float64), solver._nn, "coefs1_dofs")
    c2 = _grid(np.ascontiguousarray(coefs2_dofs, dtype=np.float64), solver._nn, "coefs2_dofs")
    out = np.zeros(1)
    check(_lib.load().gempic_maxwell2d_inner_product(solver._h, ptr(c1), ptr(c2),
                                                     int(component), int(form), ptr(out)))
    return float(out[0])


def compute_fem_rhs_(coefs_dofs, solver: TwoDMaxwell, f, component: int, form: int):
    """compute_fem_rhs! (maxwell_2d_fem.jl:212-282): host set-up (the
    reference evaluates the user's function on the host)."""
    s0, s1 = solver.s_deg_0, solver.s_deg_1
    if form == 0:
        d = [s0, s0]
    elif form == 1:
        d = [s0, s0]
        if component < 3:
            d[component - 1] = s1
    elif form == 2:
        d = [s1, s1]
        if component < 3:
            d[component - 1] = s0
    elif form == 3:
        d = [s1, s1]
    else:
        raise ValueError("form must be 0, 1, 2 or 3")
    d1, d2 = d
    x1, w1 = _gauss_legendre(d1 + 1)
    x2, w2 = _gauss_legendre(d2 + 1)
    x1 = [(a + 1.0) / 2.0 for a in x1]
    w1 = [a / 2.0 for a in w1]
    x2 = [(a + 1.0) / 2.0 for a in x2]
    w2 = [a / 2.0 for a in w2]
    b1 = [_bspline_basis(d1, a) for a in x1]     # b1[k][c]
    b2 = [_bspline_basis(d2, a) for a in x2]
    g = solver.mesh
    dx, dy = g.dx, g.dy
    ind = 0
    for i2 in range(1, g.ny + 1):
        for i1 in range(1, g.nx + 1):
            coef = 0.0
            for j1 in range(1, d1 + 2):
                for j2 in range(1, d2 + 2):
                    for k1 in range(1, d1 + 2):
                        for k2 in range(1, d2 + 2):
                            xg1 = dx * (((x1[k1 - 1] + i1) + j1) - 2)
                            xg2 = dy * (((x2[k2 - 1] + i2) + j2) - 2)
                            coef += (w1[k1 - 1] * w2[k2 - 1] * f(xg1, xg2) *
                                     b1[k1 - 1][d1 + 2 - j1 - 1] * b2[k2 - 1][d2 + 2 - j2 - 1])
            coefs_dofs[ind] = coef * dx * dy
            ind += 1
    return coefs_dofs


def l2projection(solver: TwoDMaxwell, f, component: int, form: int):
    """l2projection (maxwell_2d_fem.jl:288-298): the FEM rhs of f, solved
    with the inverse mass of a 1-form or 2-form; None for forms 0 and 3, as
    the reference returns nothing for them."""
    rhs = compute_fem_rhs_(np.zeros(solver._nn), solver, f, component, form)
    if form == 1:
        return solve(solver.inv_mass_1[component - 1], rhs)
    if form == 2:
        return solve(solver.inv_mass_2[component - 1], rhs)
    return None


# ---------------------------------------------------------------------------
# The 2D3V Hamiltonian splitting (hamiltonian_splitting_2d3v.jl is empty in
# the reference; the splitting of the GEMPIC paper on its 2D pieces)

class HamiltonianSplitting2D3V:
    """HamiltonianSplitting{2,3}(maxwell_solver, kernel_smoother_0,
    kernel_smoother_1, particle_group, e_dofs, b_dofs), the 2D3V form of
    hamiltonian_splitting.jl:20-86: ParticleGroup{2,3} (rows x1, x2, v1, v2,
    v3, w), a TwoDMaxwell of degree p, the two smoothers of degrees p and
    p-1 (galerkin), and the aliased field vectors e_dofs = [E1, E2, E3],
    b_dofs = [B1, B2, B3] (nx*ny each, E1 of degrees (p-1, p), E2 (p, p-1),
    E3 (p, p), B1 (p, p-1), B2 (p-1, p), B3 (p-1, p-1)). As the 1D splitting,
    every call uploads the state, runs on the GPU and downloads it back
    (sync=True), or leaves it on the GPU (sync=False, to_host/to_device)."""

    def __init__(self, maxwell_solver: TwoDMaxwell, kernel_smoother_0, kernel_smoother_1,
                 particle_group: ParticleGroup, e_dofs, b_dofs, sync: bool = True):
        if tuple(particle_group.dims) != (2, 3):
            raise AssertionError("the 2D3V splitting needs a ParticleGroup{2,3}")
        p = maxwell_solver.s_deg_0
        if kernel_smoother_0.degree != p or kernel_smoother_1.degree != p - 1:
            raise AssertionError("smoothers of degrees p and p-1 of the Maxwell solver")
        self.dims = (2, 3)
        self.maxwell_solver = maxwell_solver
        self.kernel_smoother_0 = kernel_smoother_0
        self.kernel_smoother_1 = kernel_smoother_1
        self.particle_group = particle_group
        mesh = maxwell_solver.mesh
        nn = mesh.nx * mesh.ny
        self.e_dofs = [_grid(e, nn, "e_dofs") for e in e_dofs]
        self.b_dofs = [_grid(b, nn, "b_dofs") for b in b_dofs]
        self.j_dofs = [np.zeros(nn) for _ in range(3)]
        self.sync = sync
        pg = particle_group
        h = ctypes.c_void_p()
        check(_lib.load().gempic_splitting2d_create(
            mesh.nx, mesh.ny, mesh.xmin, mesh.xmax, mesh.ymin, mesh.ymax, p, pg.n_particles,
            pg.charge, pg.mass, pg.common_weight, ctypes.byref(h)))
        self._h = h
        self.to_device()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().gempic_splitting2d_destroy(h)
            self._h = None

    def to_device(self, particles: bool = True, fields: bool = True):
        L = _lib.load()
        if particles:
            a = self.particle_group.array
            assert a.flags.f_contiguous
            check(L.gempic_splitting2d_set_particles(self._h, ptr(a), a.shape[0], a.shape[1]))
        if fields:
            check(L.gempic_splitting2d_set_fields(self._h, *(ptr(f) for f in self.e_dofs),
                                                  *(ptr(f) for f in self.b_dofs)))

    def to_host(self, particles: bool = True, fields: bool = True):
        L = _lib.load()
        if particles:
            a = self.particle_group.array
            check(L.gempic_splitting2d_get_particles(self._h, ptr(a), a.shape[0], a.shape[1]))
        if fields:
            check(L.gempic_splitting2d_get_fields(self._h, *(ptr(f) for f in self.e_dofs),
                                                  *(ptr(f) for f in self.b_dofs)))
            check(L.gempic_splitting2d_get_j(self._h, *(ptr(j) for j in self.j_dofs)))

    def _run(self, fn, *args):
        if self.sync:
            self.to_device()
        check(fn(self._h, *args))
        if self.sync:
            self.to_host()

    def set_sort_interval(self, k: int):
        check(_lib.load().gempic_splitting2d_set_sort_interval(self._h, int(k)))

    def operatorHB(self, dt: float):
        self._run(_lib.load().gempic_splitting2d_operator_hb, float(dt))

    def operatorHE(self, dt: float):
        self._run(_lib.load().gempic_splitting2d_operator_he, float(dt))

    def operatorHp1(self, dt: float):
        self._run(_lib.load().gempic_splitting2d_operator_hp1, float(dt))

    def operatorHp2(self, dt: float):
        self._run(_lib.load().gempic_splitting2d_operator_hp2, float(dt))

    def operatorHp3(self, dt: float):
        self._run(_lib.load().gempic_splitting2d_operator_hp3, float(dt))

    def strang_splitting(self, dt: float, number_steps: int, async_: bool = False):
        """strang_splitting!(h, dt, number_steps): HB HE Hp3 Hp2 Hp1 Hp2 Hp3 HE HB."""
        self._run(_lib.load().gempic_splitting2d_strang_splitting, float(dt), int(number_steps),
                  16 if async_ else 0)

    def synchronize(self):
        check(_lib.load().gempic_splitting2d_synchronize(self._h))

    def rho(self) -> np.ndarray:
        """add_charge! of the particles on the GPU (degree (p, p), q = charge w
        common_weight)."""
        if self.sync:
            self.to_device(fields=False)
        out = np.zeros(self.maxwell_solver._nn)
        check(_lib.load().gempic_splitting2d_rho(self._h, ptr(out)))
        return out

    def rho_from_e(self) -> np.ndarray:
        """compute_rho_from_e! of the GPU fields."""
        if self.sync:
            self.to_device(particles=False)
        out = np.zeros(self.maxwell_solver._nn)
        check(_lib.load().gempic_splitting2d_rho_from_e(self._h, ptr(out)))
        return out
