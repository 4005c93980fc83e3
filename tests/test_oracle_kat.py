"""Pins the CPU oracle against the reference's own known-answer tests.

Every vector below comes from the reference's test suite (fixtures under
tests/golden/ cite file:line); tolerances are the reference's own.
"""
import math

import numpy as np
import pytest

from conftest import golden

import pyoracle as orc


# ---------------------------------------------------------------------------
# B-spline values: test/test_particle_mesh_coupling_spline_1d.jl:25-43

def test_bspline_values_kat():
    kat = golden("kat_pmc1d.json")
    dx = (kat["xmax"] - kat["xmin"]) / kat["n_cells"]
    for x, ref in zip(kat["x"], kat["values_grid"]):
        xi = (x - kat["xmin"]) / dx
        off = xi - math.trunc(xi)
        got = orc.bspline_basis(kat["spline_degree"], off)
        assert np.max(np.abs(got - np.array(ref))) < 1e-15


# ---------------------------------------------------------------------------
# rho / j / evaluate KAT: test/test_particle_mesh_coupling_spline_1d.jl:45-144

def _pmc1d_kat_setup():
    kat = golden("kat_pmc1d.json")
    p = orc.PMC1D(kat["xmin"], kat["xmax"], kat["n_cells"], kat["spline_degree"], orc.COLLOCATION)
    cw = 1.0 / kat["n_particles"]          # ParticleGroup default common_weight
    q = 1.0 * kat["weight"] * cw            # get_charge
    return kat, p, q


def test_pmc1d_rho_kat():
    kat, p, q = _pmc1d_kat_setup()
    n = kat["n_cells"]
    rho = np.zeros(n)
    rho_pp = np.zeros(n)
    for x in kat["x"]:
        p.add_charge(rho, x, q)
        p.add_charge_pp(rho_pp, x, q)
    vg = np.array(kat["values_grid"])      # vg[particle, k] = values_grid[k, 1, particle]
    ref = np.zeros(n)
    ref[7:10] = vg[0, 0:3]
    ref[0] = vg[0, 3]
    ref[0:4] = ref[0:4] + vg[1, :] + vg[2, :]
    ref[4:8] = ref[4:8] + vg[3, :]
    ref = ref / kat["n_particles"] * n / kat["xmax"]
    assert np.max(np.abs(rho - ref)) < kat["tolerance_abs"]
    assert np.max(np.abs(rho_pp - ref)) < kat["tolerance_abs"]


def test_pmc1d_current_kat():
    kat, p, q = _pmc1d_kat_setup()
    n = kat["n_cells"]
    j = np.zeros(n)
    j_pp = np.zeros(n)
    b = np.zeros(n)
    for x, v in zip(kat["x"], kat["v"]):
        x_new = x + v[0] / 10.0
        v2 = p.add_current_update_v(j, x, x_new, q, 1.0, b, v[1])
        assert v2 == v[1]                  # B = 0: no kick
        p.add_current_update_v_pp(j_pp, x, x_new, q, 1.0, b, v[1])
    ref = np.array(kat["j_parts"][0])
    ref = ref + np.array(kat["j_parts"][1])
    ref = ref + np.array(kat["j_parts"][2])
    assert np.max(np.abs(j - ref)) < kat["tolerance_abs"]
    assert np.max(np.abs(j_pp - ref)) < kat["tolerance_abs"]


def test_pmc1d_evaluate_kat():
    kat, p, q = _pmc1d_kat_setup()
    n = kat["n_cells"]
    rho = np.zeros(n)
    for x in kat["x"]:
        p.add_charge(rho, x, q)
    rho_pp = orc.b_to_pp(kat["spline_degree"], rho)
    ref = np.array(kat["particle_values_ref_times_xmax"]) / kat["xmax"]
    got = np.array([p.evaluate(x, rho) for x in kat["x"]])
    got_pp = np.array([p.evaluate_pp(x, rho_pp) for x in kat["x"]])
    assert np.max(np.abs(got - ref)) < kat["tolerance_abs"]
    assert np.max(np.abs(got_pp - ref)) < kat["tolerance_abs"]


# ---------------------------------------------------------------------------
# 1D2V splitting KAT: test/test_hamiltonian_splitting.jl

def _splitting_kat():
    kat = golden("kat_splitting_1d2v.json")
    nx = kat["num_cells"]
    xmax = kat["eta_max_over_pi"] * math.pi     # 4π
    parts = np.asfortranarray(np.array(kat["initial"]).T)   # (4, 2)
    e1 = np.ones(nx)
    e2 = np.ones(nx)
    b = np.ones(nx)
    deg = kat["degree_smoother"]
    h = orc.Splitting(nx, kat["eta_min"], xmax, deg, deg, deg - 1, parts, e1, e2, b,
                      common_weight=kat["common_weight"])
    # rho by add_charge with ks0 and get_charge, then Poisson (:70-77)
    p0 = orc.PMC1D(kat["eta_min"], xmax, nx, deg)
    rho = np.zeros(nx)
    for i in range(parts.shape[1]):
        p0.add_charge(rho, parts[0, i], 1.0 * parts[3, i] * kat["common_weight"])
    m = orc.Maxwell1D(kat["eta_min"], xmax, nx, deg)
    m.compute_e_from_rho(e1, rho)
    return kat, h, parts, e1, e2, b


def test_splitting_1d2v_kat():
    kat, h, parts, e1, e2, b = _splitting_kat()
    dt = kat["delta_t"]
    worst = 0.0
    for op, key in (("operatorHp1", "after_hp1"), ("operatorHp2", "after_hp2"),
                    ("operatorHE", "after_he"), ("operatorHB", "after_hb")):
        getattr(h, op)(dt)
        ref = np.array(kat[key]).T
        rel = np.abs(parts - ref) / np.maximum(np.abs(ref), 1e-300)
        worst = max(worst, rel.max())
        np.testing.assert_allclose(parts, ref, rtol=kat["particle_rtol"], atol=0)
    # the Fortran-printed references carry ~17 digits: the oracle is far tighter
    assert worst < 1e-13
    assert np.max(np.abs(b - np.array(kat["bfield_ref"]))) < kat["field_atol"]
    assert np.max(np.abs(e1 - np.array(kat["efield1_ref"]))) < kat["field_atol"]
    assert np.max(np.abs(e2 - np.array(kat["efield2_ref"]))) < kat["field_atol"]


def test_splitting_kat_detects_plausible_mistakes():
    """Each plausible slip in a sub-step moves the KAT state far beyond its
    tolerance: e.g. skipping the v x B kick (a dropped term) or the sign of
    the current update. Recomputed from the oracle's pieces on the KAT
    particle 0 (test/test_hamiltonian_splitting.jl:88-121)."""
    kat = golden("kat_splitting_1d2v.json")
    nx = kat["num_cells"]
    xmax = kat["eta_max_over_pi"] * math.pi
    ks1 = orc.PMC1D(0.0, xmax, nx, kat["degree_smoother"] - 1)
    x0, v1, v2, w = kat["initial"][0]
    dt = kat["delta_t"]
    b = np.ones(nx)
    j = np.zeros(nx)
    good = ks1.add_current_update_v(j, x0, x0 + dt * v1, w, 1.0, b, v2)
    ref = kat["after_hp1"][0][2]
    assert abs(good - ref) < 1e-14
    # dropped kick
    assert abs(v2 - ref) > 1e-3
    # wrong direction (x_old/x_new transposed)
    j2 = np.zeros(nx)
    wrong = ks1.add_current_update_v(j2, x0 + dt * v1, x0, w, 1.0, b, v2)
    assert abs(wrong - ref) > 1e-3


# ---------------------------------------------------------------------------
# Closed forms and invariants for what the KATs leave open


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4, 5])
def test_bspline_partition_of_unity_and_symmetry(deg):
    rng = np.random.default_rng(deg)
    for off in rng.random(50):
        b = orc.bspline_basis(deg, off)
        assert abs(b.sum() - 1.0) < 1e-15
        # symmetry (low_level_bsplines.jl:31-35): values at 1-off are reversed
        br = orc.bspline_basis(deg, 1.0 - off)
        assert np.max(np.abs(b - br[::-1])) < 1e-14
    # offset 0, degree 3: [1/6, 2/3, 1/6, 0]
    if deg == 3:
        np.testing.assert_allclose(orc.bspline_basis(3, 0.0), [1 / 6, 2 / 3, 1 / 6, 0], atol=1e-16)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 10])
def test_gausslegendre_exactness(n):
    x, w = orc.gausslegendre(n)
    ref_x, ref_w = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(x, ref_x, atol=2e-16 * 4)
    np.testing.assert_allclose(w, ref_w, atol=2e-16 * 8)
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) < 1e-14


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
def test_splinepp_tables_match_de_boor(deg):
    pc, pf = orc.splinepp_tables(deg)
    rng = np.random.default_rng(10 + deg)
    for x in rng.random(20):
        b = orc.bspline_basis(deg, x)
        powers = np.array([x ** (deg - r) for r in range(deg + 1)])
        np.testing.assert_allclose(powers @ pc, b, atol=1e-15)
        # primitive / x: x * P_fp(x) = int_0^x piece(t) dt (check by Gauss-Legendre)
        gx, gw = np.polynomial.legendre.leggauss(8)
        t = 0.5 * x * (gx + 1.0)
        integ = np.zeros(deg + 1)
        for ti, wi in zip(t, gw):
            integ += 0.5 * x * wi * orc.bspline_basis(deg, ti)
        np.testing.assert_allclose(x * (powers @ pf), integ, atol=1e-15)
    # the reference's degree-3 constants (splinepp.jl:52-88) as exact rationals
    if deg == 3:
        assert pc[3, 1] == 4 * (1.0 / 6.0)
        assert pf[0, 0] == -(1.0 / 24.0)


def test_b_to_pp_consistency():
    rng = np.random.default_rng(42)
    for deg in (1, 2, 3, 4, 5):
        nc = 12
        p = orc.PMC1D(0.0, 1.0, nc, deg)
        b = rng.random(nc)
        pp = orc.b_to_pp(deg, b)
        for x in rng.random(30) * 0.999:
            assert abs(p.evaluate(x, b) - p.evaluate_pp(x, pp)) < 1e-14


def test_julia_mod_semantics():
    L = 2.0
    assert orc.jl_mod(3.5, L) == 1.5
    assert orc.jl_mod(-0.5, L) == 1.5
    assert orc.jl_mod(-4.0, L) == 0.0 and math.copysign(1, orc.jl_mod(-4.0, L)) == 1.0
    # r + L rounds up to L itself for a tiny negative r
    assert orc.jl_mod(-1e-20, L) == L
    assert orc.lib().or_jl_mod_i64(-1, 10) == 9
    assert orc.lib().or_jl_mod_i64(-10, 10) == 0
    assert orc.lib().or_jl_mod_i64(23, 10) == 3


def test_maxwell_eigenvalues_match_dft_of_mass_lines():
    for deg in (1, 2, 3):
        for n in (10, 16, 32):
            m = orc.Maxwell1D(0.0, 1.0, n, deg)
            t = m.tables()
            line0 = np.zeros(n)
            line1 = np.zeros(n)
            m0 = {1: [4 / 6, 1 / 6], 2: [66 / 120, 26 / 120, 1 / 120],
                  3: [2416 / 5040, 1191 / 5040, 120 / 5040, 1 / 5040]}[deg]
            m1 = {1: [1.0], 2: [4 / 6, 1 / 6], 3: [66 / 120, 26 / 120, 1 / 120]}[deg]
            for k, c in enumerate(m0):
                line0[k] += c
                if k:
                    line0[-k] += c
            for k, c in enumerate(m1):
                line1[k] += c
                if k:
                    line1[-k] += c
            f0 = np.fft.rfft(line0).real
            f1 = np.fft.rfft(line1).real
            np.testing.assert_allclose(t["eig_mass0"][: n // 2 + 1], f0, atol=1e-15)
            np.testing.assert_allclose(t["eig_mass1"][: n // 2 + 1], f1, atol=1e-15)
            # weak Ampere M0^{-1} D^T M1: the strong curl (compute_b_from_e!)
            # is (D e)_i = e_i - e_{i-1}, so D^T has symbol 1 - e^{+i theta}
            kk = np.arange(1, n // 2)
            th = 2 * np.pi * kk / n
            lam = (f1[kk] / f0[kk]) * (1 - np.exp(1j * th))
            np.testing.assert_allclose(t["eig_weak_ampere"][kk], lam.real, atol=1e-14)
            np.testing.assert_allclose(t["eig_weak_ampere"][n - kk], lam.imag, atol=1e-14)


def test_r2hc_matches_numpy_and_circulant_solve_is_dense_solve():
    rng = np.random.default_rng(3)
    for n in (10, 16, 32, 64):
        m = orc.Maxwell1D(0.0, 2.0, n, 3)
        x = rng.standard_normal(n)
        hc = m.r2hc(x)
        f = np.fft.rfft(x)
        np.testing.assert_allclose(hc[: n // 2 + 1], f.real, atol=1e-13)
        np.testing.assert_allclose(hc[n - np.arange(1, n // 2)], f.imag[1: n // 2], atol=1e-13)
        np.testing.assert_allclose(m.hc2r(hc) / n, x, atol=1e-14)
        # inverse mass solve == dense circulant solve
        t = m.tables()
        eig = np.zeros(n)
        eig[: n // 2 + 1] = 1.0 / t["eig_mass0"][: n // 2 + 1]
        got = m.solve_circulant(eig, x)
        line = np.zeros(n)
        for k, c in enumerate([2416 / 5040, 1191 / 5040, 120 / 5040, 1 / 5040]):
            line[k] += c
            if k:
                line[-k] += c
        C = np.array([[line[(i - j) % n] for j in range(n)] for i in range(n)])
        np.testing.assert_allclose(C @ got, x, atol=1e-13)


# ---------------------------------------------------------------------------
# Maxwell 1D analytic test: test/test_maxwell_1d_fem.jl:20-125

def test_maxwell_1d_analytic():
    mode = 2
    n = 256
    L = 2 * math.pi
    deg = 3
    m = orc.Maxwell1D(0.0, L, n, deg)
    k = mode * 2 * math.pi / L
    dx = L / n
    xs = np.arange(n) * dx
    # Poisson (:48-65)
    rho = np.zeros(n)
    m.compute_rhs_from_function(rho, deg, a_cos=1.0, k=k)
    ex = np.zeros(n)
    m.compute_e_from_rho(ex, rho)
    sval = orc.eval_uniform_periodic_spline_curve(deg - 1, ex)
    assert np.max(np.abs(sval - np.sin(k * xs) / k)) < 1e-6
    # Ampere (:67-86)
    dt = 0.5 * dx
    rho = np.zeros(n)
    m.compute_rhs_from_function(rho, deg - 1, a_cos=1.0, k=k)
    ex = np.zeros(n)
    m.compute_e_from_j(ex, dt * rho, 1)
    sval = orc.eval_uniform_periodic_spline_curve(deg - 1, ex)
    assert np.max(np.abs(sval + np.cos(k * xs) * dt)) < 1e-6
    # Leapfrog (:92-125); l2projection! as written leaves the rhs / dx
    bz = np.zeros(n)
    ey = np.zeros(n)
    m.l2projection(bz, deg - 1, a_cos=1.0, k=k)
    t = 0.0
    for _ in range(10):
        m.compute_b_from_e(bz, 0.5 * dt, ey)
        m.compute_e_from_b(ey, dt, bz)
        m.compute_b_from_e(bz, 0.5 * dt, ey)
        t += dt
        ey_ex = np.sin(k * xs) * np.sin(k * t)
        bz_ex = np.cos(k * xs) * np.cos(k * t)
        err_e = np.linalg.norm(orc.eval_uniform_periodic_spline_curve(deg, ey) - ey_ex)
        err_b = np.linalg.norm(orc.eval_uniform_periodic_spline_curve(deg - 1, bz) - bz_ex)
        assert err_e < 1e-2 and err_b < 1e-2


def test_l2projection_leaves_rhs_over_dx():
    """maxwell_1d_fem.jl:347-374 solves into the workspace but returns
    coefs_dofs ./ delta_x, i.e. the right-hand side without the inverse mass."""
    n = 32
    m = orc.Maxwell1D(0.0, 2 * math.pi / 1.25, n, 3)
    rhs = np.zeros(n)
    m.compute_rhs_from_function(rhs, 2, a_cos=1e-4, k=1.25)
    b = np.zeros(n)
    m.l2projection(b, 2, a_cos=1e-4, k=1.25)
    np.testing.assert_array_equal(b, rhs / m.delta_x)


# ---------------------------------------------------------------------------
# Charge conservation of the line-integral deposition (invariant)

def test_line_integral_current_is_charge_conserving():
    """With exact line-integral deposition, the deposited current j (degree
    p-1) and the charge rho (degree p) before/after the move satisfy the
    discrete continuity equation rho_new[k] - rho_old[k] = (j[k] - j[k+1]) / dx.
    The reference's trunc() (particle_mesh_coupling_1d.jl:307,313) breaks it
    for a particle that leaves through x < 0: its new cell index is 0 and the
    cell-0 polynomial is integrated over a negative offset."""
    rng = np.random.default_rng(7)
    nx, L = 16, 3.0
    ks0 = orc.PMC1D(0.0, L, nx, 3)
    ks1 = orc.PMC1D(0.0, L, nx, 2)
    b = np.zeros(nx)
    dx = L / nx
    n_quirk = 0
    for _ in range(400):
        x = rng.random() * L
        dxp = (rng.random() - 0.5) * 1.5
        q = rng.random()
        j = np.zeros(nx)
        ks1.add_current_update_v(j, x, x + dxp, q, 1.0, b, 0.0)
        r0 = np.zeros(nx)
        r1 = np.zeros(nx)
        ks0.add_charge(r0, x, q)
        ks0.add_charge(r1, orc.jl_mod(x + dxp, L), q)
        resid = np.max(np.abs((r1 - r0) - (j - np.roll(j, -1)) / dx))
        if x + dxp >= 0.0:
            assert resid < 1e-13
        else:
            n_quirk += 1
            assert resid > 1e-6
    assert n_quirk > 0
