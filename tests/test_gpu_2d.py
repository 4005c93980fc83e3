"""GPU parity of the 2D rows: ParticleMeshCoupling2D's batched verbs and
TwoDMaxwell (with TwoDPoisson and TwoDLinearSolverSplineMass), through the
C ABI, against the oracle and the reference's own 2D tests.

Tolerances: evaluation, the field solves and the field updates use the
oracle's operations in its order and are compared bit for bit. The deposit
sums each particle's contribution in a different order than the reference's
sequential loop (deterministic per-lane, per-wave and per-workgroup sums, or
LDS atomics in the fast device form), so it is compared at 1e-13 relative to
the largest grid value; the same holds for inner_product's sum over the grid.
"""
import math

import numpy as np
import pytest

from conftest import golden

import gempic_amd as G
import pyoracle as orc

pytestmark = pytest.mark.gpu

DEP_REL = 1e-13


def _coupling(nx, ny, bounds, degree, smoothing, n=1):
    grid = G.TwoDGrid(bounds[0], bounds[1], nx, bounds[2], bounds[3], ny)
    pg = G.ParticleGroup(2, 2, max(n, 1))
    return grid, G.ParticleMeshCoupling2D(pg, grid, degree, smoothing)


def _oracle_deposit(p, rho, x, y, w):
    p.add_charge_n(rho, x, y, w)          # the reference's loop over particles, in order


def _assert_deposit(got, want):
    scale = max(np.abs(want).max(), 1e-300)
    err = np.abs(got - want).max() / scale
    assert err <= DEP_REL, f"deposit differs by {err:.3e} (relative to max |rho|)"


# ---------------------------------------------------------------------------
# test/test_particle_mesh_coupling_spline_2d.jl known answers on the GPU

def test_gpu_pmc2d_kat():
    kat = golden("kat_pmc2d.json")
    n = kat["n_cells"]
    bounds = (kat["xmin"], kat["xmax"], kat["ymin"], kat["ymax"])
    _, pm = _coupling(n, n, bounds, kat["spline_degree"], "collocation", kat["n_particles"])
    x, y = np.array(kat["x"]), np.array(kat["y"])
    rho = np.zeros(n * n)
    G.add_charge_(rho, pm, x, y, kat["weight"])
    # the reference test's rho_ref (:67-75)
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
    ref *= n ** 2 / ((bounds[1] - bounds[0]) * (bounds[3] - bounds[2])) / kat["n_particles"]
    np.testing.assert_allclose(rho, ref, rtol=1e-14, atol=1e-14)
    # evaluate at the particles (:95-107), bitwise equal to the oracle
    p = orc.PMC2D(*bounds[:2], n, *bounds[2:], n, kat["spline_degree"], orc.COLLOCATION)
    got = G.evaluate(pm, x, y, ref)
    want = np.array([p.evaluate(a, b, ref) for a, b in zip(x, y)])
    np.testing.assert_array_equal(got, want)
    g1, g2 = G.evaluate_multiple(pm, (x, y), (ref, 2.0 * ref))
    w12 = np.array([p.evaluate_multiple(a, b, ref, 2.0 * ref) for a, b in zip(x, y)])
    np.testing.assert_array_equal(g1, w12[:, 0])
    np.testing.assert_array_equal(g2, w12[:, 1])


def _particles(rng, n, bounds, nx, ny):
    """Uniform positions plus the edge cases: the domain's lower corner,
    positions exactly on grid nodes, and positions just below the upper
    bounds."""
    x = bounds[0] + (bounds[1] - bounds[0]) * rng.random(n)
    y = bounds[2] + (bounds[3] - bounds[2]) * rng.random(n)
    dx = (bounds[1] - bounds[0]) / nx
    dy = (bounds[3] - bounds[2]) / ny
    k = min(n, 64)
    x[:k] = bounds[0] + dx * rng.integers(0, nx, k)
    y[:k] = bounds[2] + dy * rng.integers(0, ny, k)
    if n > k + 2:
        x[k], y[k] = bounds[0], bounds[2]
        x[k + 1], y[k + 1] = np.nextafter(bounds[1], -np.inf), np.nextafter(bounds[3], -np.inf)
    return x, y


@pytest.mark.parametrize("degree", [1, 2, 3])
@pytest.mark.parametrize("smoothing", ["galerkin", "collocation"])
@pytest.mark.parametrize("nx,ny", [(16, 12), (5, 7)])
def test_gpu_pmc2d_matches_oracle(degree, smoothing, nx, ny):
    if min(nx, ny) < degree + 1:
        pytest.skip("grid smaller than the spline support")
    rng = np.random.default_rng(100 * degree + nx)
    bounds = (-0.5, 1.75, 0.25, 2.0)
    n = 5000 + 37                                    # ragged: not a multiple of 64
    x, y = _particles(rng, n, bounds, nx, ny)
    w = rng.random(n) - 0.3
    _, pm = _coupling(nx, ny, bounds, degree, smoothing, n)
    sm = orc.COLLOCATION if smoothing == "collocation" else orc.GALERKIN
    p = orc.PMC2D(bounds[0], bounds[1], nx, bounds[2], bounds[3], ny, degree, sm)
    rho0 = rng.normal(size=nx * ny)                  # add_charge! accumulates
    rho, want = rho0.copy(), rho0.copy()
    G.add_charge_(rho, pm, x, y, w)
    _oracle_deposit(p, want, x, y, w)
    _assert_deposit(rho, want)
    f1, f2 = rng.normal(size=nx * ny), rng.normal(size=nx * ny)
    got = G.evaluate(pm, x, y, f1)
    np.testing.assert_array_equal(got, [p.evaluate(a, b, f1) for a, b in zip(x, y)])
    g1, g2 = G.evaluate_multiple(pm, (x, y), (f1, f2))
    wm = np.array([p.evaluate_multiple(a, b, f1, f2) for a, b in zip(x, y)])
    np.testing.assert_array_equal(g1, wm[:, 0])
    np.testing.assert_array_equal(g2, wm[:, 1])


@pytest.mark.parametrize("order", ["random", "sorted"])
def test_gpu_pmc2d_one_million_particles(order):
    """1,048,576 particles on the 64 x 64 grid of the 2D3V configuration,
    degree 3: deposit within 1e-13 of the oracle, total charge conserved,
    evaluation bitwise equal."""
    rng = np.random.default_rng(7)
    nx = ny = 64
    bounds = (0.0, 4 * math.pi, 0.0, 4 * math.pi)
    n = 1 << 20
    x = bounds[1] * rng.random(n)
    y = bounds[3] * rng.random(n)
    if order == "sorted":
        cell = np.floor(x / (bounds[1] / nx)) + nx * np.floor(y / (bounds[3] / ny))
        idx = np.argsort(cell, kind="stable")
        x, y = x[idx], y[idx]
    w = np.full(n, 1.0 / n)
    _, pm = _coupling(nx, ny, bounds, 3, "galerkin", n)
    p = orc.PMC2D(*bounds[:2], nx, *bounds[2:], ny, 3, orc.GALERKIN)
    rho, want = np.zeros(nx * ny), np.zeros(nx * ny)
    G.add_charge_(rho, pm, x, y, w)
    _oracle_deposit(p, want, x, y, w)
    _assert_deposit(rho, want)
    assert abs(rho.sum() - 1.0) < 1e-12
    f = rng.normal(size=nx * ny)
    got = G.evaluate(pm, x[:20000], y[:20000], f)
    np.testing.assert_array_equal(got, [p.evaluate(a, b, f) for a, b in zip(x[:20000], y[:20000])])


def test_gpu_pmc2d_edge_cases():
    bounds = (0.0, 1.0, 0.0, 1.0)
    _, pm = _coupling(8, 8, bounds, 3, "galerkin")
    rho = np.arange(64.0)
    G.add_charge_(rho, pm, np.zeros(0), np.zeros(0), np.zeros(0))      # empty input
    np.testing.assert_array_equal(rho, np.arange(64.0))
    # one particle (scalar form)
    p = orc.PMC2D(0.0, 1.0, 8, 0.0, 1.0, 8, 3, orc.GALERKIN)
    want = rho.copy()
    p.add_charge(want, 0.3, 0.7, 0.5)
    G.add_charge_(rho, pm, 0.3, 0.7, 0.5)
    _assert_deposit(rho, want)
    assert G.evaluate(pm, 0.3, 0.7, want) == p.evaluate(0.3, 0.7, want)
    # non-finite positions: Julia's ceil(Int, NaN) InexactError
    with pytest.raises(G.InexactError):
        G.add_charge_(rho, pm, np.array([0.5, np.nan]), np.array([0.5, 0.5]), 1.0)
    with pytest.raises(G.InexactError):
        G.evaluate(pm, np.array([0.5]), np.array([np.inf]), want)
    # evaluation is not limited by the deposit's LDS grid
    _, big = _coupling(100, 90, bounds, 2, "galerkin")
    pb = orc.PMC2D(0.0, 1.0, 100, 0.0, 1.0, 90, 2, orc.GALERKIN)
    f = np.random.default_rng(3).normal(size=9000)
    xs = np.linspace(0.0, 0.999, 257)
    ys = np.linspace(0.999, 0.0, 257)
    np.testing.assert_array_equal(G.evaluate(big, xs, ys, f),
                                  [pb.evaluate(a, b, f) for a, b in zip(xs, ys)])


def test_gpu_pmc2d_deposit_is_deterministic():
    rng = np.random.default_rng(11)
    n = 300000
    x, y = rng.random(n), rng.random(n)
    w = rng.random(n)
    _, pm = _coupling(32, 32, (0.0, 1.0, 0.0, 1.0), 3, "galerkin", n)
    r1, r2 = np.zeros(1024), np.zeros(1024)
    G.add_charge_(r1, pm, x, y, w)
    G.add_charge_(r2, pm, x, y, w)
    np.testing.assert_array_equal(r1, r2)


# ---------------------------------------------------------------------------
# TwoDMaxwell

def _pair(nx, ny, degree, bounds=(0.0, 2.0, -0.5, 1.0)):
    m = G.TwoDMaxwell(G.TwoDGrid(bounds[0], bounds[1], nx, bounds[2], bounds[3], ny), degree)
    o = orc.Maxwell2D(bounds[0], bounds[1], nx, bounds[2], bounds[3], ny, degree)
    return m, o


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_gpu_maxwell2d_tables_and_lines(degree):
    m, o = _pair(12, 10, degree)
    np.testing.assert_array_equal(m.inv_mass_1[2].eigvals1, o.table(0, 12))   # eig0x
    np.testing.assert_array_equal(m.inv_mass_1[2].eigvals2, o.table(1, 10))   # eig0y
    np.testing.assert_array_equal(m.inv_mass_1[0].eigvals1, o.table(2, 12))   # eig1x
    np.testing.assert_array_equal(m.inv_mass_1[1].eigvals2, o.table(3, 10))   # eig1y
    np.testing.assert_array_equal(m.mass_line_0[0], o.table(4, degree + 1))
    np.testing.assert_array_equal(m.mass_line_0[1], o.table(5, degree + 1))
    np.testing.assert_array_equal(m.mass_line_1[0], o.table(6, degree))
    np.testing.assert_array_equal(m.mass_line_1[1], o.table(7, degree))


@pytest.mark.parametrize("degree", [1, 2, 3])
@pytest.mark.parametrize("nx,ny", [(12, 10), (7, 32)])
def test_gpu_maxwell2d_verbs_match_oracle(degree, nx, ny):
    if min(nx, ny) < 2 * degree + 1:
        pytest.skip("grid too small for the banded mass product")
    m, o = _pair(nx, ny, degree)
    nn = nx * ny
    rng = np.random.default_rng(degree * 1000 + nx)
    rnd = lambda: rng.normal(size=nn)  # noqa: E731
    # the Kronecker solves of every form and component
    for form, inv in ((1, m.inv_mass_1), (2, m.inv_mass_2)):
        for c in (1, 2, 3):
            r = rnd()
            np.testing.assert_array_equal(G.solve(inv[c - 1], r), o.solve(form, c, r))
    # multiply_mass_2dkron! with each pair of lines
    for l1 in (m.mass_line_0[0], m.mass_line_1[0]):
        for l2 in (m.mass_line_0[1], m.mass_line_1[1]):
            cin, out = rnd(), np.zeros(nn)
            G.multiply_mass_2dkron_(out, m, l1, l2, cin)
            np.testing.assert_array_equal(out, o.multiply_mass(l1, l2, cin))
    # inner_product: sum over the grid in another order
    for form, comp in ((0, 1), (1, 1), (1, 2), (1, 3), (2, 1), (2, 2), (2, 3), (3, 1)):
        c1, c2 = rnd(), rnd()
        got = G.inner_product(m, c1, c2, comp, form)
        want = o.inner_product(c1, c2, comp, form)
        assert abs(got - want) <= 1e-13 * np.sum(np.abs(c1) * np.abs(c2)) * max(
            1.0, float(np.abs(m.mass_line_0[0]).max())), (form, comp, got, want)
    # field updates, in place
    e = [rnd(), rnd(), rnd()]
    b = [rnd(), rnd(), rnd()]
    eo = [a.copy() for a in e]
    G.compute_e_from_b_(e, m, 0.05, b)
    o.compute_e_from_b(eo, 0.05, b)
    for a, z in zip(e, eo):
        np.testing.assert_array_equal(a, z)
    bo = [a.copy() for a in b]
    G.compute_b_from_e_(b, m, 0.05, e)
    o.compute_b_from_e(bo, 0.05, e)
    for a, z in zip(b, bo):
        np.testing.assert_array_equal(a, z)
    for c in (1, 2, 3):
        ej, j = rnd(), rnd()
        ejo = ej.copy()
        G.compute_e_from_j_(ej, m, j, c)
        o.compute_e_from_j(ejo, j, c)
        np.testing.assert_array_equal(ej, ejo)
    rho, rhoo = np.zeros(nn), np.zeros(nn)
    G.compute_rho_from_e_(rho, m, e)
    o.compute_rho_from_e(rhoo, e)
    np.testing.assert_array_equal(rho, rhoo)
    r = rnd()
    ep = [np.zeros(nn), np.zeros(nn)]
    epo = [np.zeros(nn), np.zeros(nn), np.zeros(nn)]
    G.compute_e_from_rho_(ep, m, r)
    o.compute_e_from_rho(epo, r)
    np.testing.assert_array_equal(ep[0], epo[0])
    np.testing.assert_array_equal(ep[1], epo[1])


def test_gpu_gauss_law_at_256x256():
    """At the largest grid: rho_from_e(e_from_rho(rho)) = rho - mean(rho)."""
    m = G.TwoDMaxwell(G.TwoDGrid(0.0, 4.0, 256, 0.0, 3.0, 256), 3)
    rng = np.random.default_rng(5)
    rho = rng.normal(size=256 * 256)
    e = [np.zeros(256 * 256), np.zeros(256 * 256)]
    G.compute_e_from_rho_(e, m, rho)
    back = np.zeros(256 * 256)
    G.compute_rho_from_e_(back, m, e)
    np.testing.assert_allclose(back, rho - rho.mean(), atol=1e-10 * np.abs(rho).max())


# ---------------------------------------------------------------------------
# test/test_maxwell_2d_fem.jl through the GPU path (analytic, 1e-4)

def _eval_spline_2d(nx, ny, degs, dofs):
    vals = dofs.reshape(ny, nx).copy()
    for j in range(ny):
        vals[j, :] = orc.eval_uniform_periodic_spline_curve(degs[0], np.ascontiguousarray(vals[j, :]))
    for i in range(nx):
        vals[:, i] = orc.eval_uniform_periodic_spline_curve(degs[1], np.ascontiguousarray(vals[:, i]))
    return vals.reshape(-1)


def test_gpu_maxwell_2d_analytic():
    x1min, x1max, nx1 = 0.0, 2 * math.pi, 16
    x2min, x2max, nx2 = 0.0, 2 * math.pi, 32
    deg, dt, nsteps = 3, 0.01, 300
    m = G.TwoDMaxwell(G.TwoDGrid(x1min, x1max, nx1, x2min, x2max, nx2), deg)
    nn = nx1 * nx2
    xg = np.tile(np.linspace(x1min, x1max, nx1 + 1)[:-1], nx2)
    yg = np.repeat(np.linspace(x2min, x2max, nx2 + 1)[:-1], nx1)
    w1 = math.sqrt(3)
    t = 0.0
    sin_k = lambda x, y: math.sin((x + y) - w1 * t)  # noqa: E731
    cos_k = lambda x, y: math.cos((x + y) - w1 * t)  # noqa: E731
    # Poisson (:48-61)
    rho = G.compute_fem_rhs_(np.zeros(nn), m, cos_k, 1, 0)
    e = [np.zeros(nn), np.zeros(nn), np.zeros(nn)]
    G.compute_e_from_rho_(e, m, rho)
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
    b = [G.l2projection(m, e1, 1, 2), G.l2projection(m, e2, 2, 2), G.l2projection(m, b3, 3, 2)]
    tt[0] = 0.0
    e = [G.l2projection(m, e1, 1, 1), G.l2projection(m, e2, 2, 1), -G.l2projection(m, b3, 3, 1)]
    for _ in range(nsteps):
        G.compute_b_from_e_(b, m, dt, e)
        G.compute_e_from_b_(e, m, dt, b)
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
    ef = G.l2projection(m, cos_k, 1, 1)
    assert abs(G.inner_product(m, ef, ef, 1, 1) - 2 * math.pi ** 2) < 1e-5
    # Ampere (:121-133)
    j = G.compute_fem_rhs_(np.zeros(nn), m, sin_k, 1, 1)
    G.compute_e_from_j_(ef, m, j, 1)
    got = _eval_spline_2d(nx1, nx2, (deg - 1, deg), ef)
    assert np.max(np.abs(got - (np.cos(xg + yg) - np.sin(xg + yg)))) < 1e-2
    # Gauss (:135-146)
    rho_ref = 2.0 * G.compute_fem_rhs_(np.zeros(nn), m, cos_k, 1, 0)
    ef = [G.l2projection(m, sin_k, c, 1) for c in (1, 2, 3)]
    rho = np.zeros(nn)
    G.compute_rho_from_e_(rho, m, ef)
    np.testing.assert_allclose(rho, rho_ref, rtol=1e-6, atol=1e-8)


# ---------------------------------------------------------------------------
# device-resident forms (particles in GPU memory)

from gempic_amd.device import DeviceArray  # noqa: E402


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_gpu_pmc2d_device_forms(degree):
    """The device-resident verbs: the LDS-atomic deposit within 1e-13 of the
    oracle, the deterministic deposit bitwise equal to the host form, and
    the staged and unstaged evaluations bitwise equal to the oracle."""
    rng = np.random.default_rng(21 + degree)
    nx, ny = 48, 40
    bounds = (0.0, 3.0, -1.0, 1.0)
    n = 200003
    x, y = _particles(rng, n, bounds, nx, ny)
    w = rng.random(n)
    _, pm = _coupling(nx, ny, bounds, degree, "galerkin", n)
    p = orc.PMC2D(bounds[0], bounds[1], nx, bounds[2], bounds[3], ny, degree, orc.GALERKIN)
    want = np.zeros(nx * ny)
    _oracle_deposit(p, want, x, y, w)
    dx_, dy_, dw = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y), DeviceArray.from_numpy(w)
    rho = DeviceArray.zeros(nx * ny)
    G.add_charge_device_(rho, pm, dx_, dy_, dw)
    _assert_deposit(rho.numpy(), want)
    rho_det = DeviceArray.zeros(nx * ny)
    G.add_charge_device_(rho_det, pm, dx_, dy_, dw, deterministic=True)
    host = np.zeros(nx * ny)
    G.add_charge_(host, pm, x, y, w)
    np.testing.assert_array_equal(rho_det.numpy(), host)
    # 1920 nodes: one field (15 KiB) and two fields are staged in LDS
    f1, f2 = rng.normal(size=nx * ny), rng.normal(size=nx * ny)
    out = DeviceArray(n)
    G.evaluate_device(pm, dx_, dy_, DeviceArray.from_numpy(f1), out)
    np.testing.assert_array_equal(out.numpy(), p.evaluate_n(x, y, f1))
    o1, o2 = DeviceArray(n), DeviceArray(n)
    G.evaluate_multiple_device(pm, (dx_, dy_), (DeviceArray.from_numpy(f1),
                                                DeviceArray.from_numpy(f2)), (o1, o2))
    r1, r2 = p.evaluate_multiple_n(x, y, f1, f2)
    np.testing.assert_array_equal(o1.numpy(), r1)
    np.testing.assert_array_equal(o2.numpy(), r2)


def test_gpu_pmc2d_device_unstaged_eval_and_async_errors():
    rng = np.random.default_rng(33)
    nx, ny = 96, 80                    # 7680 nodes: two fields are read from global memory
    _, pm = _coupling(nx, ny, (0.0, 1.0, 0.0, 1.0), 3, "galerkin")
    p = orc.PMC2D(0.0, 1.0, nx, 0.0, 1.0, ny, 3, orc.GALERKIN)
    n = 10000
    x, y = rng.random(n), rng.random(n)
    f1, f2 = rng.normal(size=nx * ny), rng.normal(size=nx * ny)
    D = DeviceArray.from_numpy
    o1, o2 = DeviceArray(n), DeviceArray(n)
    G.evaluate_device(pm, D(x), D(y), D(f1), o1)
    np.testing.assert_array_equal(o1.numpy(), p.evaluate_n(x, y, f1))
    G.evaluate_multiple_device(pm, (D(x), D(y)), (D(f1), D(f2)), (o1, o2))
    r1, r2 = p.evaluate_multiple_n(x, y, f1, f2)
    np.testing.assert_array_equal(o1.numpy(), r1)
    np.testing.assert_array_equal(o2.numpy(), r2)
    # an InexactError of an asynchronous call surfaces at the sync, once
    _, small = _coupling(16, 16, (0.0, 1.0, 0.0, 1.0), 3, "galerkin")
    xb = x.copy()
    xb[17] = np.nan
    rho = DeviceArray.zeros(256)
    G.add_charge_device_(rho, small, D(xb), D(y), D(np.ones(n)), async_=True)
    with pytest.raises(G.InexactError):
        G.pmc2d_sync()
    G.pmc2d_sync()
    with pytest.raises(ValueError):                          # the deposit's LDS grid limit
        G.add_charge_device_(DeviceArray.zeros(nx * ny), pm, D(x), D(y), D(np.ones(n)))
    with pytest.raises(TypeError):
        G.add_charge_device_(rho, small, x, y, np.ones(n))


# ---------------------------------------------------------------------------
# the pp (piecewise-polynomial, Horner) variants

def test_gpu_pmc2d_pp_kat():
    """test_particle_mesh_coupling_spline_2d.jl: add_charge_pp! gives rho_ref
    and evaluate_pp on b_to_pp_2d! gives the evaluate values (:67-118)."""
    kat = golden("kat_pmc2d.json")
    n = kat["n_cells"]
    bounds = (kat["xmin"], kat["xmax"], kat["ymin"], kat["ymax"])
    _, pm = _coupling(n, n, bounds, kat["spline_degree"], "collocation", kat["n_particles"])
    p = orc.PMC2D(*bounds[:2], n, *bounds[2:], n, kat["spline_degree"], orc.COLLOCATION)
    x, y = np.array(kat["x"]), np.array(kat["y"])
    rho = np.zeros(n * n)
    G.add_charge_pp_(rho, pm, x, y, kat["weight"])
    want = np.zeros(n * n)
    for a, b in zip(x, y):
        p.add_charge(want, a, b, kat["weight"])
    np.testing.assert_allclose(rho, want, rtol=1e-14, atol=1e-13)
    pp = np.zeros((16, n * n), order="F")
    G.b_to_pp_2d_(pp, pm, want)
    np.testing.assert_array_equal(pp, p.b_to_pp(want))
    got = G.evaluate_pp(pm, x, y, pp)
    np.testing.assert_array_equal(got, [p.evaluate_pp(a, b, pp) for a, b in zip(x, y)])
    ev = np.array([p.evaluate(a, b, want) for a, b in zip(x, y)])
    np.testing.assert_allclose(got, ev, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("degree", [1, 2, 3])
@pytest.mark.parametrize("nx,ny", [(16, 12), (5, 7)])
def test_gpu_pmc2d_pp_match_oracle(degree, nx, ny):
    if min(nx, ny) < degree + 1:
        pytest.skip("grid smaller than the spline support")
    rng = np.random.default_rng(7 * degree + nx)
    bounds = (-0.5, 1.75, 0.25, 2.0)
    n = 3001
    x, y = _particles(rng, n, bounds, nx, ny)
    w = rng.random(n) - 0.2
    _, pm = _coupling(nx, ny, bounds, degree, "galerkin", n)
    p = orc.PMC2D(bounds[0], bounds[1], nx, bounds[2], bounds[3], ny, degree, orc.GALERKIN)
    rho, want = np.zeros(nx * ny), np.zeros(nx * ny)
    G.add_charge_pp_(rho, pm, x, y, w)
    for a, b, c in zip(x, y, w):
        p.add_charge_pp(want, a, b, c)
    _assert_deposit(rho, want)
    f = rng.normal(size=nx * ny)
    ld = (degree + 1) ** 2
    pp = np.zeros((ld, nx * ny), order="F")
    G.b_to_pp_2d_(pp, pm, f)
    np.testing.assert_array_equal(pp, p.b_to_pp(f))
    # positions whose floor()+1 index leaves the table are a BoundsError in
    # the reference (evaluate_pp does not wrap): the one just below the upper
    # bound can round onto index nx+1
    inside = []
    for i, (a, b) in enumerate(zip(x, y)):
        try:
            p.evaluate_pp(a, b, pp)
            inside.append(i)
        except IndexError:
            with pytest.raises(ValueError, match="BoundsError"):
                G.evaluate_pp(pm, np.array([a]), np.array([b]), pp)
    xi, yi = x[inside], y[inside]
    got = G.evaluate_pp(pm, xi, yi, pp)
    np.testing.assert_array_equal(got, [p.evaluate_pp(a, b, pp) for a, b in zip(xi, yi)])
    # evaluate_pp of the pp form is evaluate of the B-spline form
    np.testing.assert_allclose(got, p.evaluate_n(xi, yi, f), rtol=0, atol=1e-12 * np.abs(f).max())
    with pytest.raises(ValueError, match="BoundsError"):
        G.evaluate_pp(pm, np.array([0.1, bounds[1] + 0.5]), np.array([0.5, 0.5]), pp)
