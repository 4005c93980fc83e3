"""The 2D oracle (ParticleMeshCoupling2D, TwoDMaxwell, TwoDPoisson,
TwoDLinearSolverSplineMass) against the reference's own 2D tests and against
what the mathematics fixes: Kronecker mass matrices built from the mass
lines, Gauss quadrature of the splines, and the discrete Gauss law."""
import math

import numpy as np
import pytest

import pyoracle as orc
from conftest import golden


# ---------------------------------------------------------------------------
# test/test_particle_mesh_coupling_spline_2d.jl

def _kat2d():
    kat = golden("kat_pmc2d.json")
    n = kat["n_cells"]
    p = orc.PMC2D(kat["xmin"], kat["xmax"], n, kat["ymin"], kat["ymax"], n,
                  kat["spline_degree"], orc.COLLOCATION)
    return kat, p, n


def _rho_ref(kat, n):
    """rho_ref as the reference test assembles it (:67-75)."""
    vx = [np.array(v) for v in kat["v_grid_x"]]
    ref = np.zeros(n * n)
    ref[7:10] = vx[0][0:3]
    ref[0] = vx[0][3]
    ref[0:4] += vx[1] + vx[2]
    ref[4:8] += vx[3]
    ref[70:80] = ref[0:10] / 6.0
    ref[80:90] = ref[0:10] * 2.0 / 3.0
    ref[90:100] = ref[0:10] / 6.0
    ref[0:10] = 0.0
    volume = (kat["xmax"] - kat["xmin"]) * (kat["ymax"] - kat["ymin"])
    ref *= n ** 2 / volume / kat["n_particles"]
    return ref


def test_pmc2d_rho_kat():
    kat, p, n = _kat2d()
    rho = np.zeros(n * n)
    rho_pp = np.zeros(n * n)
    for x, y in zip(kat["x"], kat["y"]):
        p.add_charge(rho, x, y, kat["weight"])
        p.add_charge_pp(rho_pp, x, y, kat["weight"])
    ref = _rho_ref(kat, n)
    np.testing.assert_allclose(rho, ref, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(rho_pp, ref, rtol=1e-14, atol=1e-13)


def test_pmc2d_evaluate_kat():
    kat, p, n = _kat2d()
    rho = _rho_ref(kat, n)
    pp = p.b_to_pp(rho)
    vx = [np.array(v) for v in kat["v_grid_x"]]
    vy = np.array(kat["v_grid_y"])
    for ip, (x, y) in enumerate(zip(kat["x"], kat["y"])):
        ref = 0.0
        for i in range(4):
            i1 = (kat["index_grid_x"][ip] + i - 1) % n
            for j in range(4):
                i2 = (kat["index_grid_y"][ip] + j - 1) % n
                ref += vx[ip][i] * vy[j] * rho[i1 + i2 * n]
        assert abs(p.evaluate(x, y, rho) - ref) < 1e-13 * max(1.0, abs(ref))
        assert abs(p.evaluate_pp(x, y, pp) - ref) < 1e-12 * max(1.0, abs(ref))
        f1, f2 = p.evaluate_multiple(x, y, rho, 2.0 * rho)
        assert f1 == p.evaluate(x, y, rho) or abs(f1 - ref) < 1e-13 * max(1.0, abs(ref))
        assert abs(f2 - 2.0 * ref) < 2e-13 * max(1.0, abs(ref))


def test_pmc2d_partition_of_unity_and_total_charge():
    """Galerkin deposit of weight w spreads exactly w over the grid, and a
    constant field evaluates to that constant, at any position."""
    rng = np.random.default_rng(2)
    p = orc.PMC2D(-1.0, 2.0, 12, 0.5, 2.5, 9, 3, orc.GALERKIN)
    rho = np.zeros(12 * 9)
    tot = 0.0
    for _ in range(200):
        x = -1.0 + 3.0 * rng.random()
        y = 0.5 + 2.0 * rng.random()
        w = rng.random()
        p.add_charge(rho, x, y, w)
        tot += w
        assert abs(p.evaluate(x, y, np.full(12 * 9, 0.7)) - 0.7) < 1e-15
    assert abs(rho.sum() - tot) < 1e-12 * tot


# ---------------------------------------------------------------------------
# Mass lines, eigenvalues, Kronecker solves

def test_mass_line_is_the_spline_autocorrelation():
    """spline_fem_mass_line (poisson_2d_fem.jl:87-106): int N_0 N_j over the
    line, the Eulerian-number lines of Maxwell1DFEM (maxwell_1d_fem.jl:78-88)."""
    np.testing.assert_allclose(orc.spline_fem_mass_line(3),
                               np.array([2416, 1191, 120, 1]) / 5040, rtol=1e-14)
    np.testing.assert_allclose(orc.spline_fem_mass_line(2), np.array([66, 26, 1]) / 120,
                               rtol=1e-14)
    np.testing.assert_allclose(orc.spline_fem_mass_line(1), np.array([4, 1]) / 6, rtol=1e-14)


def _circulant_from_line(n, line):
    C = np.zeros((n, n))
    for i in range(n):
        C[i, i] += line[0]
        for c in range(1, len(line)):
            C[i, (i + c) % n] += line[c]
            C[i, (i - c) % n] += line[c]
    return C


def test_mass_eigenvalues_are_those_of_the_circulant():
    n = 11
    line = orc.spline_fem_mass_line(3) * 0.3
    eig = orc.spline_fem_compute_mass_eig(n, 3, line)
    got = np.sort(eig)
    want = np.sort(np.linalg.eigvalsh(_circulant_from_line(n, line)))
    np.testing.assert_allclose(got, want, rtol=1e-13)


@pytest.mark.parametrize("form,component", [(1, 1), (1, 2), (1, 3), (2, 1), (2, 2), (2, 3)])
def test_kron_solve_inverts_the_kronecker_mass_matrix(form, component):
    """TwoDLinearSolverSplineMass with the eigenvalues TwoDMaxwell assigns to
    each form and component (maxwell_2d_fem.jl:56-66) inverts M_x (x) M_y."""
    nx, ny, deg = 8, 6, 3
    m = orc.Maxwell2D(0.0, 2.0, nx, 0.0, 1.5, ny, deg)
    dx, dy = 2.0 / nx, 1.5 / ny
    l0, l1 = orc.spline_fem_mass_line(deg), orc.spline_fem_mass_line(deg - 1)
    if form == 1:
        ax = l1 if component == 1 else l0
        ay = l1 if component == 2 else l0
    else:
        ax = l0 if component == 1 else l1
        ay = l0 if component == 2 else l1
    Mx = _circulant_from_line(nx, ax * dx)
    My = _circulant_from_line(ny, ay * dy)
    M = np.kron(My, Mx)                    # x fastest
    rng = np.random.default_rng(form * 10 + component)
    xs = rng.normal(size=nx * ny)
    got = m.solve(form, component, M @ xs)
    np.testing.assert_allclose(got, xs, rtol=0, atol=1e-12 * np.abs(xs).max())
    # and multiply_mass_2dkron is that matrix
    np.testing.assert_allclose(m.multiply_mass(ax * dx, ay * dy, xs), M @ xs, atol=1e-13)


def _quad_inner(c1, c2, nx, ny, dx, dy, degx, degy):
    """int (sum c1 N)(sum c2 N) over the torus by Gauss quadrature of the
    tensor B-splines (exact)."""
    qx, wx = np.polynomial.legendre.leggauss(degx + 1)
    qy, wy = np.polynomial.legendre.leggauss(degy + 1)
    tot = 0.0
    C1 = c1.reshape(ny, nx)
    C2 = c2.reshape(ny, nx)
    for cy in range(ny):
        for cx in range(nx):
            for tx, ux in zip(0.5 * (qx + 1), wx):
                bx = orc.bspline_basis(degx, tx)
                ix = [(cx - degx + a) % nx for a in range(degx + 1)]
                for ty, uy in zip(0.5 * (qy + 1), wy):
                    by = orc.bspline_basis(degy, ty)
                    iy = [(cy - degy + a) % ny for a in range(degy + 1)]
                    f = sum(C1[jy, jx] * bx[a] * by[b] for a, jx in enumerate(ix)
                            for b, jy in enumerate(iy))
                    g = sum(C2[jy, jx] * bx[a] * by[b] for a, jx in enumerate(ix)
                            for b, jy in enumerate(iy))
                    tot += 0.25 * ux * uy * f * g * dx * dy
    return tot


@pytest.mark.parametrize("form,component", [(0, 1), (1, 1), (1, 2), (2, 1), (2, 3), (3, 1)])
def test_inner_product_matches_quadrature(form, component):
    nx, ny, deg = 6, 5, 3
    m = orc.Maxwell2D(0.0, 1.2, nx, 0.0, 1.0, ny, deg)
    rng = np.random.default_rng(form + 3 * component)
    c1 = rng.normal(size=nx * ny)
    c2 = rng.normal(size=nx * ny)
    s0, s1 = deg, deg - 1
    if form == 0:
        dgx = dgy = s0
    elif form == 1:
        dgx = s1 if component == 1 else s0
        dgy = s1 if component == 2 else s0
    elif form == 2:
        dgx = s0 if component == 1 else s1
        dgy = s0 if component == 2 else s1
    else:
        dgx = dgy = s1
    want = _quad_inner(c1, c2, nx, ny, 1.2 / nx, 1.0 / ny, dgx, dgy)
    assert abs(m.inner_product(c1, c2, component, form) - want) < 1e-12 * max(1.0, abs(want))


# ---------------------------------------------------------------------------
# Gauss's law: rho_from_e(e_from_rho(rho)) = rho without its mean

def test_poisson_then_divergence_is_identity_on_zero_mean():
    nx, ny = 8, 10
    m = orc.Maxwell2D(0.0, 2.0, nx, 0.0, 3.0, ny, 3)
    rng = np.random.default_rng(4)
    rho = rng.normal(size=nx * ny)
    e = [np.zeros(nx * ny) for _ in range(3)]
    m.compute_e_from_rho(e, rho)
    back = np.zeros(nx * ny)
    m.compute_rho_from_e(back, e)
    np.testing.assert_allclose(back, rho - rho.mean(), atol=1e-12 * np.abs(rho).max())


# ---------------------------------------------------------------------------
# test/test_maxwell_2d_fem.jl (analytic, 1e-4)

def _eval_spline_2d(nx, ny, degs, dofs):
    vals = dofs.reshape(ny, nx).copy()
    for j in range(ny):
        vals[j, :] = orc.eval_uniform_periodic_spline_curve(degs[0], np.ascontiguousarray(vals[j, :]))
    for i in range(nx):
        vals[:, i] = orc.eval_uniform_periodic_spline_curve(degs[1], np.ascontiguousarray(vals[:, i]))
    return vals.reshape(-1)


def test_maxwell_2d_analytic():
    x1min, x1max, nx1 = 0.0, 2 * math.pi, 16
    x2min, x2max, nx2 = 0.0, 2 * math.pi, 32
    deg, dt, nsteps = 3, 0.01, 300
    m = orc.Maxwell2D(x1min, x1max, nx1, x2min, x2max, nx2, deg)
    xg = np.tile(np.linspace(x1min, x1max, nx1 + 1)[:-1], nx2)
    yg = np.repeat(np.linspace(x2min, x2max, nx2 + 1)[:-1], nx1)
    w1 = math.sqrt(3)
    t = 0.0
    sin_k = lambda x, y: math.sin((x + y) - w1 * t)  # noqa: E731
    cos_k = lambda x, y: math.cos((x + y) - w1 * t)  # noqa: E731
    # Poisson (:48-61)
    rho = m.fem_rhs(cos_k, 1, 0)
    e = [np.zeros(nx1 * nx2) for _ in range(3)]
    m.compute_e_from_rho(e, rho)
    ref = np.sin(xg + yg) / 2
    np.testing.assert_allclose(_eval_spline_2d(nx1, nx2, (deg - 1, deg), e[0]), ref, rtol=1e-4,
                               atol=1e-4)
    np.testing.assert_allclose(_eval_spline_2d(nx1, nx2, (deg, deg - 1), e[1]), ref, rtol=1e-4,
                               atol=1e-4)
    # Leapfrog (:63-113)
    s2 = math.sqrt(2)
    tt = [-0.5 * dt]
    e1 = lambda x, y: math.cos(x) * math.sin(y) * math.sin(s2 * tt[0]) / s2  # noqa: E731
    e2 = lambda x, y: -math.sin(x) * math.cos(y) * math.sin(s2 * tt[0]) / s2  # noqa: E731
    b3 = lambda x, y: -math.cos(x) * math.cos(y) * math.cos(s2 * tt[0])  # noqa: E731
    b = [m.l2projection(e1, 1, 2), m.l2projection(e2, 2, 2), m.l2projection(b3, 3, 2)]
    tt[0] = 0.0
    e = [m.l2projection(e1, 1, 1), m.l2projection(e2, 2, 1), -m.l2projection(b3, 3, 1)]
    for _ in range(nsteps):
        m.compute_b_from_e(b, dt, e)
        m.compute_e_from_b(e, dt, b)
    tt[0] = (nsteps - 0.5) * dt
    for comp, degs, f in ((0, (deg, deg - 1), e1), (1, (deg - 1, deg), e2)):
        got = _eval_spline_2d(nx1, nx2, degs, b[comp])
        want = np.array([f(x, y) for x, y in zip(xg, yg)])
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4)
    got = _eval_spline_2d(nx1, nx2, (deg - 1, deg - 1), b[2])
    want = np.array([b3(x, y) for x, y in zip(xg, yg)])
    np.testing.assert_allclose(got, want, rtol=1e-3, atol=1e-3)
    tt[0] = nsteps * dt
    got = _eval_spline_2d(nx1, nx2, (deg - 1, deg), e[0])
    want = np.array([e1(x, y) for x, y in zip(xg, yg)])
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4)
    # L2 norm (:115-119)
    t = 0.0
    ef = m.l2projection(cos_k, 1, 1)
    assert abs(m.inner_product(ef, ef, 1, 1) - 2 * math.pi ** 2) < 1e-5
    # Ampere (:121-133)
    j = m.fem_rhs(sin_k, 1, 1)
    m.compute_e_from_j(ef, j, 1)
    got = _eval_spline_2d(nx1, nx2, (deg - 1, deg), ef)
    want = np.cos(xg + yg) - np.sin(xg + yg)
    assert np.max(np.abs(got - want)) < 1e-2
    # Gauss (:135-146)
    rho_ref = 2.0 * m.fem_rhs(cos_k, 1, 0)
    ef = [m.l2projection(sin_k, c, 1) for c in (1, 2, 3)]
    rho = np.zeros(nx1 * nx2)
    m.compute_rho_from_e(rho, ef)
    np.testing.assert_allclose(rho, rho_ref, rtol=1e-6, atol=1e-8)


def test_pmc2d_batched_loops_are_the_per_particle_calls():
    rng = np.random.default_rng(9)
    p = orc.PMC2D(0.0, 1.0, 7, -1.0, 1.0, 6, 2, orc.GALERKIN)
    x, y, w = rng.random(300), 2.0 * rng.random(300) - 1.0, rng.random(300)
    r1, r2 = np.zeros(42), np.zeros(42)
    p.add_charge_n(r1, x, y, w)
    for a, b, c in zip(x, y, w):
        p.add_charge(r2, a, b, c)
    np.testing.assert_array_equal(r1, r2)
    f1, f2 = rng.normal(size=42), rng.normal(size=42)
    np.testing.assert_array_equal(p.evaluate_n(x, y, f1), [p.evaluate(a, b, f1) for a, b in zip(x, y)])
    o1, o2 = p.evaluate_multiple_n(x, y, f1, f2)
    ref = np.array([p.evaluate_multiple(a, b, f1, f2) for a, b in zip(x, y)])
    np.testing.assert_array_equal(o1, ref[:, 0])
    np.testing.assert_array_equal(o2, ref[:, 1])
