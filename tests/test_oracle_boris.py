"""The oracle's HamiltonianSplittingBoris against the reference's own test
(test/test_hamiltonian_splitting_boris.jl) and against what the mathematics
fixes: the B rotation is orthogonal, the E kick of a constant field is
v + dt qm E, the current is the charge deposit at the midpoint, and a
uniform B turns every velocity by the Boris angle 2 atan(qm B dt / 2) per
step."""
import math

import numpy as np
import pyoracle as orc
from conftest import golden


def _boris_kat_setup():
    kat = golden("kat_boris.json")
    n, nc, deg = kat["n_particles"], kat["num_cells"], kat["degree_smoother"]
    xmax = kat["eta_max_over_pi"] * math.pi
    parts = np.asfortranarray(np.array(kat["particle_info_initial"]).reshape(n, 4).T)
    e1, e2, b = np.zeros(nc), np.ones(nc), np.ones(nc)
    sp = orc.Splitting(nc, kat["eta_min"], xmax, deg, deg, deg - 1, parts, e1, e2, b,
                       common_weight=1.0)
    # rho of the two particles with kernel_smoother_0, E1 from Poisson (:61-71)
    p0 = orc.PMC1D(kat["eta_min"], xmax, nc, deg, orc.GALERKIN)
    rho = np.zeros(nc)
    for i in range(n):
        p0.add_charge(rho, parts[0, i], 1.0 * parts[3, i] * 1.0)
    m = orc.Maxwell1D(kat["eta_min"], xmax, nc, deg)
    m.compute_e_from_rho(e1, rho)
    return kat, sp, parts, e1, e2, b


def test_boris_kat():
    kat, sp, parts, e1, e2, b = _boris_kat_setup()
    dt = kat["delta_t"]
    sp.boris_staggering(0.5 * dt)
    sp.boris_strang(dt, 1)
    ref = np.array(kat["particle_info_ref"]).reshape(kat["n_particles"], 4).T
    np.testing.assert_allclose(parts, ref, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(b, kat["bfield_ref"], rtol=0, atol=1e-14)
    eref = np.array(kat["efield_ref"]).reshape(2, kat["num_cells"])
    np.testing.assert_allclose(e1, eref[0], rtol=0, atol=1e-14)
    np.testing.assert_allclose(e2, eref[1], rtol=0, atol=1e-14)


def _hot(n, nx=16, seed=3, L=4 * math.pi):
    rng = np.random.default_rng(seed)
    parts = np.asfortranarray(np.vstack([L * rng.random(n), rng.normal(size=n),
                                         rng.normal(size=n), rng.random(n) + 0.5]))
    e1, e2, b = (rng.normal(size=nx) * 0.3 for _ in range(3))
    return parts, e1, e2, b, L


def test_boris_b_rotation_is_orthogonal_and_e_kick_is_linear():
    parts, e1, e2, b, L = _hot(400)
    sp = orc.Splitting(16, 0.0, L, 3, 3, 2, parts, e1, e2, b, common_weight=0.01)
    rng = np.random.default_rng(1)
    sp.boris_set_mid(rng.normal(size=16), rng.normal(size=16), 2.0 * rng.normal(size=16))
    v0 = parts[1:3].copy()
    sp.boris_push_v_bpart(0.3)
    np.testing.assert_allclose((parts[1:3] ** 2).sum(0), (v0 ** 2).sum(0), rtol=2e-15)
    assert np.abs(parts[1:3] - v0).max() > 1e-3            # it does rotate
    np.testing.assert_array_equal(parts[0], parts[0])       # x untouched
    # constant E_mid: v += dt * qm * E exactly as the reference rounds it
    sp.boris_set_mid(np.full(16, 0.7), np.full(16, -0.2), np.zeros(16))
    v1 = parts[1:3].copy()
    sp.boris_push_v_epart(0.05)
    e_at = np.array([sum(0.7 * s for s in orc.bspline_basis(2, 0.5))])  # partition of unity
    assert abs(e_at[0] - 0.7) < 1e-15
    np.testing.assert_allclose(parts[1], v1[0] + 0.05 * 0.7, rtol=0, atol=1e-15)
    np.testing.assert_allclose(parts[2], v1[1] + 0.05 * -0.2, rtol=0, atol=1e-15)


def test_boris_current_is_the_midpoint_charge_deposit():
    parts, e1, e2, b, L = _hot(300, seed=5)
    x0, v = parts[0].copy(), parts[1:3].copy()
    sp = orc.Splitting(16, 0.0, L, 3, 3, 2, parts, e1, e2, b, common_weight=0.02)
    sp.boris_push_x_accumulate_j(0.1)
    p1 = orc.PMC1D(0.0, L, 16, 2, orc.GALERKIN)
    p0 = orc.PMC1D(0.0, L, 16, 3, orc.GALERKIN)
    j1, j2 = np.zeros(16), np.zeros(16)
    for i in range(300):
        q = 1.0 * parts[3, i] * 0.02
        xm = (x0[i] + (x0[i] + 0.1 * v[0, i])) * 0.5
        p1.add_charge(j1, xm, q * v[0, i])
        p0.add_charge(j2, xm, q * v[1, i])
    np.testing.assert_array_equal(sp.j(1), j1)
    np.testing.assert_array_equal(sp.j(2), j2)
    np.testing.assert_allclose(parts[0], np.mod(x0 + 0.1 * v[0], L), rtol=0, atol=1e-13)
    # total current = sum of q v (partition of unity)
    q = parts[3] * 0.02
    assert abs(j1.sum() - (q * v[0]).sum()) < 1e-12 * np.abs(q * v[0]).sum()
    assert abs(j2.sum() - (q * v[1]).sum()) < 1e-12 * np.abs(q * v[1]).sum()


def test_boris_gyration_in_a_uniform_b():
    """With a uniform B and negligible self-fields (weights 1e-30) every
    velocity turns by exactly theta = 2 atan(qm B dt / 2) per step, clockwise
    as dv/dt = qm v x B for B along z, and x moves with the staggered v1."""
    n, nx, L, dt, B0, steps = 64, 16, 4 * math.pi, 0.1, 2.0, 40
    rng = np.random.default_rng(2)
    x0, v0 = L * rng.random(n), rng.normal(size=(2, n))
    parts = np.asfortranarray(np.vstack([x0, v0, np.ones(n)]))
    e1, e2, b = np.zeros(nx), np.zeros(nx), np.full(nx, B0)
    sp = orc.Splitting(nx, 0.0, L, 3, 3, 2, parts, e1, e2, b, common_weight=1e-30)
    sp.boris_staggering(0.5 * dt)
    sp.boris_strang(dt, steps)
    th = 2.0 * math.atan(1.0 * B0 * dt / 2.0)
    c, s_ = math.cos(steps * th), math.sin(steps * th)
    np.testing.assert_allclose(parts[1], c * v0[0] + s_ * v0[1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(parts[2], -s_ * v0[0] + c * v0[1], rtol=0, atol=1e-12)
    # x: half a step with v0 (staggering), then dt times each step's v1
    x = x0 + 0.25 * dt * v0[0]
    for k in range(1, steps + 1):
        ck, sk = math.cos(k * th), math.sin(k * th)
        x = x + dt * (ck * v0[0] + sk * v0[1])
    d = np.abs(np.mod(x, L) - parts[0])
    assert np.minimum(d, L - d).max() < 1e-11
    np.testing.assert_allclose(b, B0, rtol=0, atol=1e-25)
