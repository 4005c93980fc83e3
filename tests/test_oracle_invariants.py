"""Pins of the oracle functions that the reference's known-answer tests leave
open: the 1d1v operators and the no-B current deposit, the Maxwell1DFEM
norms, the Strang composition and the write_step! diagnostics.

Each check is fixed by the mathematics rather than by the oracle's own
formula: charge continuity of the line-integral deposit, Gauss's law kept by
Ampere's update, mass-matrix norms against Gauss quadrature of the spline
itself, second-order energy conservation of the symmetric splitting, and the
diagnostics as exact rates of change of the energies they belong to.
"""
import math

import numpy as np
import pytest

import pyoracle as orc


# ---------------------------------------------------------------------------
# add_current_update_v! without B (particle_mesh_coupling_1d.jl:471-580)

def test_current_nob_is_charge_conserving_and_leaves_v():
    """The 1d1v deposit takes floor() for the new cell (:487), so unlike the
    1d2v variant (trunc, :307,313) it satisfies the discrete continuity
    equation rho_new[k] - rho_old[k] = (j[k] - j[k+1]) / dx for every move,
    including moves out through x < 0; it returns v unchanged."""
    rng = np.random.default_rng(7)
    nx, L = 16, 3.0
    ks0 = orc.PMC1D(0.0, L, nx, 3)
    ks1 = orc.PMC1D(0.0, L, nx, 2)
    dx = L / nx
    n_neg = 0
    for _ in range(400):
        x = rng.random() * L
        dxp = (rng.random() - 0.5) * 1.5
        q = rng.random()
        j = np.zeros(nx)
        v = ks1.add_current_update_v_nob(j, x, x + dxp, q, 1.0, 0.37)
        assert v == 0.37
        r0 = np.zeros(nx)
        r1 = np.zeros(nx)
        ks0.add_charge(r0, x, q)
        ks0.add_charge(r1, orc.jl_mod(x + dxp, L), q)
        resid = np.max(np.abs((r1 - r0) - (j - np.roll(j, -1)) / dx))
        assert resid < 1e-13
        n_neg += x + dxp < 0.0
    assert n_neg > 0


# ---------------------------------------------------------------------------
# Maxwell1DFEM l2norm_squared / inner_product (maxwell_1d_fem.jl:299-313,461-475)

def _spline_integral(f_coefs, g_coefs, degree, L):
    """Integral over the period of (sum_k f_k N_k)(sum_k g_k N_k) by
    Gauss-Legendre quadrature of the B-spline pieces cell by cell (exact:
    degree+1 points integrate the degree-2p products)."""
    n = len(f_coefs)
    dx = L / n
    xq, wq = np.polynomial.legendre.leggauss(degree + 1)
    tq = 0.5 * (xq + 1.0)
    total = 0.0
    for cell in range(n):
        for t, w in zip(tq, wq):
            b = orc.bspline_basis(degree, t)
            idx = [(cell - degree + m) % n for m in range(degree + 1)]
            f = sum(f_coefs[i] * bm for i, bm in zip(idx, b))
            g = sum(g_coefs[i] * bm for i, bm in zip(idx, b))
            total += 0.5 * w * f * g * dx
    return total


@pytest.mark.parametrize("maxwell_degree", [1, 2, 3])
def test_mass_norms_match_quadrature(maxwell_degree):
    """Both degrees the solver holds mass matrices for: s_deg_0 and s_deg_0 - 1."""
    rng = np.random.default_rng(maxwell_degree)
    n, L = 12, 2.7
    m = orc.Maxwell1D(0.0, L, n, maxwell_degree)
    for degree in (maxwell_degree, maxwell_degree - 1):
        c1 = rng.normal(size=n)
        c2 = rng.normal(size=n)
        ref = _spline_integral(c1, c1, degree, L)
        assert abs(m.l2norm_squared(c1, degree) - ref) < 1e-12 * max(1.0, ref)
        ref = _spline_integral(c1, c2, degree, L)
        assert abs(m.inner_product(c1, c2, degree) - ref) < 1e-12 * max(1.0, abs(ref))


# ---------------------------------------------------------------------------
# 1d1v operators (hamiltonian_splitting_1d1v.jl:63-125)

def _landau_1d1v(n=2000, nx=16, seed=5):
    rng = np.random.default_rng(seed)
    L = 2 * math.pi / 0.5
    x = rng.random(n) * L
    v = rng.normal(0.0, 1.0, n)
    parts = np.asfortranarray(np.vstack([x, v, np.full(n, L)]))
    e1, e2, b = np.zeros(nx), np.zeros(nx), np.zeros(nx)
    h = orc.Splitting(nx, 0.0, L, 3, 3, 2, parts, e1, e2, b)
    rho = np.zeros(nx)
    h.solve_poisson(e1, rho)
    return h, parts, e1


def test_1d1v_hp1_keeps_gauss_law():
    """Ampere's update with the line-integral current keeps E1 equal to the
    Poisson field of the moved charge (up to the mean, which follows the net
    current)."""
    h, parts, e1 = _landau_1d1v()
    nx = len(e1)
    for _ in range(5):
        h.operatorHp1(0.05)
        ep, rho = np.zeros(nx), np.zeros(nx)
        h.solve_poisson(ep, rho)
        assert np.max(np.abs((e1 - e1.mean()) - (ep - ep.mean()))) < 1e-13 * np.max(np.abs(ep))
        assert np.max(np.abs(ep)) > 0.1


def test_1d1v_hb_kicks_by_dt_times_field():
    """operatorHB of 1d1v is v += dt * E1(x) (no q/m); for a constant field the
    splines' partition of unity makes the kick exactly dt * c."""
    h, parts, e1 = _landau_1d1v()
    e1[:] = 0.7
    v0 = parts[1].copy()
    x0 = parts[0].copy()
    h.operatorHB(0.05)
    np.testing.assert_allclose(parts[1], v0 + 0.05 * 0.7, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(parts[0], x0)


# ---------------------------------------------------------------------------
# strang_splitting! and write_step! (hamiltonian_splitting.jl:98-108,
# diagnostics.jl:186-250)

def _hot_weibel(n=4096, nx=16, seed=3, mass=2.0, charge=-1.0):
    rng = np.random.default_rng(seed)
    k = 1.25
    L = 2 * math.pi / k
    x = rng.random(n) * L
    parts = np.asfortranarray(np.vstack([x, rng.normal(0, 0.3, n), rng.normal(0, 0.5, n),
                                         np.full(n, L)]))
    m = orc.Maxwell1D(0.0, L, nx, 3)
    b = np.zeros(nx)
    m.l2projection(b, 2, a_cos=0.3, k=k)
    e1, e2 = np.zeros(nx), np.zeros(nx)
    h = orc.Splitting(nx, 0.0, L, 3, 3, 2, parts, e1, e2, b, charge=charge, mass=mass)
    rho = np.zeros(nx)
    h.solve_poisson(e1, rho)
    return h, parts, e1, e2, b


def _diag(h, e1, e2, b):
    return h.write_step(0.0, 3, e1, e2, b, e1, e2, np.zeros(len(e1)))


def _energy(o):
    # (KineticEnergy + PotentialEnergyE1 + E2 + B3) / 2 (diagnostics.jl:206,226-236)
    return 0.5 * (o[1] + o[4] + o[5] + o[6])


def test_strang_energy_error_is_second_order():
    """The splitting conserves the total energy up to O(dt^2) with no drift,
    while kinetic and magnetic energy exchange a sizeable fraction of it; a
    non-symmetric composition (first order) or a mis-weighted energy term
    breaks one of the two. Mass 2 and charge -1 make the q/m, charge and
    mass weights all matter."""
    devs = []
    for dt in (0.1, 0.05, 0.025):
        h, parts, e1, e2, b = _hot_weibel()
        o0 = _diag(h, e1, e2, b)
        e0 = _energy(o0)
        worst = 0.0
        for _ in range(int(round(2.0 / dt))):
            h.strang_splitting(dt, 1)
            worst = max(worst, abs(_energy(_diag(h, e1, e2, b)) - e0) / e0)
        o = _diag(h, e1, e2, b)
        assert abs(o[1] - o0[1]) / e0 > 5e-2          # energy really flows
        devs.append(worst)
    assert devs[1] < 1e-4
    for a, c in zip(devs, devs[1:]):
        assert 3.5 < a / c < 4.5


def _rate(step, quantity, dt=1e-3):
    """d quantity / d t of a sub-step that is a polynomial of degree 2 in dt,
    by Richardson extrapolation of two one-sided differences."""
    t1 = (quantity(step(dt)) - quantity(step(0.0))) / dt
    t2 = (quantity(step(0.5 * dt)) - quantity(step(0.0))) / (0.5 * dt)
    return 2.0 * t2 - t1


def _fresh_after(op, dt):
    h, parts, e1, e2, b = _hot_weibel(n=512, nx=12)
    e2[:] = np.cos(np.arange(len(e2)))          # make every field nonzero
    if dt != 0.0:
        getattr(h, op)(dt)
    return h, e1, e2, b


def test_transfer_is_the_rate_of_kinetic_energy_under_he():
    """operatorHE: v += dt q/m E(x), so d(KineticEnergy)/dt = 2 Transfer."""
    def ke(s):
        h, e1, e2, b = s
        return _diag(h, e1, e2, b)[1]
    rate = _rate(lambda dt: _fresh_after("operatorHE", dt), ke)
    h, e1, e2, b = _fresh_after("operatorHE", 0.0)
    transfer = _diag(h, e1, e2, b)[7]
    assert abs(rate - 2.0 * transfer) < 1e-9 * max(1.0, abs(transfer))
    assert abs(transfer) > 1e-3


def test_vvb_is_the_rate_of_kinetic_energy_under_hp2():
    """operatorHp2: v1 += dt q/m v2 B(x) (v2 fixed), so d(KineticEnergy)/dt = 2 VVB."""
    def ke(s):
        h, e1, e2, b = s
        return _diag(h, e1, e2, b)[1]
    rate = _rate(lambda dt: _fresh_after("operatorHp2", dt), ke)
    h, e1, e2, b = _fresh_after("operatorHp2", 0.0)
    vvb = _diag(h, e1, e2, b)[8]
    assert abs(rate - 2.0 * vvb) < 1e-9 * max(1.0, abs(vvb))
    assert abs(vvb) > 1e-3


def test_poynting_is_the_rate_of_e2_energy_under_hb():
    """operatorHB: E2 += dt M0^-1 R^T B, so d(PotentialEnergyE2)/dt = 2 Poynting."""
    def pe2(s):
        h, e1, e2, b = s
        return _diag(h, e1, e2, b)[5]
    rate = _rate(lambda dt: _fresh_after("operatorHB", dt), pe2)
    h, e1, e2, b = _fresh_after("operatorHB", 0.0)
    poynting = _diag(h, e1, e2, b)[9]
    assert abs(rate - 2.0 * poynting) < 1e-9 * max(1.0, abs(poynting))
    assert abs(poynting) > 1e-3


def test_momentum_rate_under_he_is_the_force():
    """operatorHE: d(Momentum_i)/dt = sum_p q_p E_i(x_p) (Newton), with the
    field evaluated by the smoothers HE uses (E1: degree p-1, E2: degree p)."""
    h, e1, e2, b = _fresh_after("operatorHE", 0.0)
    parts = h.particles
    q = parts[3] * (-1.0) * (1.0 / parts.shape[1])
    L = 2 * math.pi / 1.25
    ks0 = orc.PMC1D(0.0, L, len(e1), 3)
    ks1 = orc.PMC1D(0.0, L, len(e1), 2)
    f1 = sum(qi * ks1.evaluate(xi, e1) for xi, qi in zip(parts[0], q))
    f2 = sum(qi * ks0.evaluate(xi, e2) for xi, qi in zip(parts[0], q))
    for col, force in ((2, f1), (3, f2)):
        def mom(s, col=col):
            hh, a, c, d = s
            return _diag(hh, a, c, d)[col]
        rate = _rate(lambda dt: _fresh_after("operatorHE", dt), mom)
        assert abs(rate - force) < 1e-9 * max(1.0, abs(force))
        assert abs(force) > 1e-4


def test_error_poisson_is_zero_after_steps_and_sees_stale_fields():
    """ErrorPoisson = max |E1 - E_poisson|: round-off when E_poisson is the
    Poisson field of the current particles, since the line-integral current
    keeps Gauss's law (up to E1's mean, which follows the net current), and
    large against a stale E_poisson. The particles stay clear of x = 0, where
    the reference's trunc() breaks charge continuity (see
    test_line_integral_current_is_charge_conserving)."""
    h, parts, e1, e2, b = _hot_weibel()
    L = 2 * math.pi / 1.25
    parts[0] = 0.2 * L + 0.6 * L * (parts[0] / L)
    parts[1] *= 0.1
    nx = len(e1)
    rho = np.zeros(nx)
    h.solve_poisson(e1, rho)
    stale = e1.copy()
    h.strang_splitting(0.05, 10)
    assert parts[0].min() > 0.1 * L and parts[0].max() < 0.9 * L
    ep, rho = np.zeros(nx), np.zeros(nx)
    h.solve_poisson(ep, rho)
    shift = e1.mean() - ep.mean()
    out = h.write_step(0.0, 3, e1, e2, b, e1, e2, ep + shift)
    assert out[10] < 1e-13 * np.max(np.abs(ep))
    out = h.write_step(0.0, 3, e1, e2, b, e1, e2, stale)
    assert out[10] > 1e-3 * np.max(np.abs(ep))
