"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Per-particle arithmetic is the reference's, rounded identically (no FMA
contraction on either side), so particle states after a sub-step from the
same input agree to the last bit. Deposited grid quantities (rho, j) differ
only in summation order; they are compared normwise against the sum of the
absolute contributions: |j_gpu - j_ref|_inf <= 1e-12 * |sum|contrib||_inf.
Fields derived from them inherit that bound. Tolerances are written in each
test.
"""
import math

import numpy as np
import pytest

from conftest import golden

import gempic_amd as G
from gempic_amd import workloads
import pyoracle as orc

pytestmark = pytest.mark.gpu

REL = 1e-12


def _setup_pair(w, degree0=3, degree1=2, sync=True, charge=1.0, mass=1.0):
    """A GPU splitting and an oracle splitting on copies of the same state."""
    nx, xmin, xmax = w["nx"], w["xmin"], w["xmax"]
    parts = w["particles"]
    V = parts.shape[0] - 2
    n = parts.shape[1]
    mesh = G.OneDGrid(xmin, xmax, nx)
    pg = G.ParticleGroup(1, V, n, charge=charge, mass=mass, common_weight=w["common_weight"])
    pg.array[:] = parts
    ks0 = G.ParticleMeshCoupling1D(mesh, n, degree0, "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, n, degree1, "galerkin")
    mx = G.Maxwell1DFEM(mesh, 3)
    e1 = np.zeros(nx)
    e2 = np.zeros(nx)
    b = w.get("b0", np.zeros(nx)).copy()
    # E1 from Poisson on the oracle side, fed to both
    o_parts = np.asfortranarray(parts.copy())
    o_e1, o_e2, o_b = e1.copy(), e2.copy(), b.copy()
    osp = orc.Splitting(nx, xmin, xmax, 3, degree0, degree1, o_parts, o_e1, o_e2, o_b,
                        charge=charge, mass=mass, common_weight=w["common_weight"])
    rho = np.zeros(nx)
    osp.solve_poisson(o_e1, rho)
    e1[:] = o_e1
    h = G.HamiltonianSplitting(1, V, mx, ks0, ks1, pg, [e1, e2], b, sync=sync)
    return h, pg, osp, o_parts, (o_e1, o_e2, o_b)


def _rel_field(a, b, scale):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(scale)), 1e-300)


def assert_rows_close(got, ref, rel):
    """Each particle-array row (x, v1, ...) within rel * max|row| (normwise:
    a near-zero velocity inherits the absolute rounding of its row)."""
    for r in range(ref.shape[0]):
        scale = max(np.max(np.abs(ref[r])), 1e-300)
        err = np.max(np.abs(got[r] - ref[r]))
        assert err <= rel * scale, f"row {r}: {err} > {rel} * {scale}"


# ---------------------------------------------------------------------------
# Known-answer tests through the GPU path

def test_gpu_pmc1d_kat():
    kat = golden("kat_pmc1d.json")
    mesh = G.OneDGrid(kat["xmin"], kat["xmax"], kat["n_cells"])
    p = G.ParticleMeshCoupling1D(mesh, kat["n_particles"], kat["spline_degree"], "collocation")
    q = 1.0 * kat["weight"] * (1.0 / kat["n_particles"])
    x = np.array(kat["x"])
    n = kat["n_cells"]
    rho = np.zeros(n)
    rho_pp = np.zeros(n)
    G.add_charge_(rho, p, x, q)
    G.add_charge_pp_(rho_pp, p, x, q)
    vg = np.array(kat["values_grid"])
    ref = np.zeros(n)
    ref[7:10] = vg[0, 0:3]
    ref[0] = vg[0, 3]
    ref[0:4] = ref[0:4] + vg[1, :] + vg[2, :]
    ref[4:8] = ref[4:8] + vg[3, :]
    ref = ref / kat["n_particles"] * n / kat["xmax"]
    assert np.max(np.abs(rho - ref)) < 1e-15
    assert np.max(np.abs(rho_pp - ref)) < 1e-15
    v = np.array(kat["v"])
    j = np.zeros(n)
    j_pp = np.zeros(n)
    b = np.zeros(n)
    v2 = G.add_current_update_v_(j, p, x, x + v[:, 0] / 10.0, q, 1.0, b, v[:, 1].copy())
    G.add_current_update_v_pp_(j_pp, p, x, x + v[:, 0] / 10.0, q, 1.0, b, v[:, 1].copy())
    jref = np.array(kat["j_parts"][0]) + np.array(kat["j_parts"][1]) + np.array(kat["j_parts"][2])
    assert np.max(np.abs(j - jref)) < 1e-15
    assert np.max(np.abs(j_pp - jref)) < 1e-15
    np.testing.assert_array_equal(v2, v[:, 1])
    vals = G.evaluate(p, x, rho)
    vals_pp = G.evaluate_pp(p, x, G.b_to_pp(p, n, rho))
    vref = np.array(kat["particle_values_ref_times_xmax"]) / kat["xmax"]
    assert np.max(np.abs(vals - vref)) < 1e-15
    assert np.max(np.abs(vals_pp - vref)) < 1e-15


def test_gpu_splitting_kat():
    kat = golden("kat_splitting_1d2v.json")
    nx = kat["num_cells"]
    xmax = kat["eta_max_over_pi"] * math.pi
    mesh = G.OneDGrid(kat["eta_min"], xmax, nx)
    pg = G.ParticleGroup(1, 2, 2, common_weight=kat["common_weight"])
    pg.array[:] = np.array(kat["initial"]).T
    ks1 = G.ParticleMeshCoupling1D(mesh, 2, 2, "galerkin")
    ks0 = G.ParticleMeshCoupling1D(mesh, 2, 3, "galerkin")
    mx = G.Maxwell1DFEM(mesh, 3)
    e1 = np.ones(nx)
    e2 = np.ones(nx)
    b = np.ones(nx)
    rho = np.zeros(nx)
    for i in range(2):
        G.add_charge_(rho, ks0, pg.array[0, i], G.get_charge(pg, i))
    G.compute_e_from_rho_(e1, mx, rho)
    h = G.HamiltonianSplitting(1, 2, mx, ks0, ks1, pg, [e1, e2], b)
    dt = kat["delta_t"]
    for op, key in ((G.operatorHp1, "after_hp1"), (G.operatorHp2, "after_hp2"),
                    (G.operatorHE, "after_he"), (G.operatorHB, "after_hb")):
        op(h, dt)
        np.testing.assert_allclose(pg.array, np.array(kat[key]).T, rtol=kat["particle_rtol"], atol=0)
    assert np.max(np.abs(b - np.array(kat["bfield_ref"]))) < kat["field_atol"]
    assert np.max(np.abs(e1 - np.array(kat["efield1_ref"]))) < kat["field_atol"]
    assert np.max(np.abs(e2 - np.array(kat["efield2_ref"]))) < kat["field_atol"]


# ---------------------------------------------------------------------------
# Sub-steps against the oracle on seeded inputs

def _hot_weibel(n, seed):
    """Weibel geometry with velocities large enough that particles cross
    cells, leave through x < 0 and x >= L, and skip cells."""
    w = workloads.weibel_1d2v(n, seed=seed, nx=32, sigma1=0.6, sigma2=0.4)
    w["b0"] = w["b0"] * 1e3 + 0.3
    return w


@pytest.mark.parametrize("pair", [(3, 2), (3, 3), (2, 1), (1, 0)])
def test_gpu_substeps_match_oracle(pair):
    w = _hot_weibel(20000, seed=11)
    h, pg, osp, o_parts, (o_e1, o_e2, o_b) = _setup_pair(w, *pair)
    dt = 0.4
    for name in ("operatorHp1", "operatorHp2", "operatorHE", "operatorHB", "operatorHp1"):
        getattr(G, name)(h, dt)
        getattr(osp, name)(dt)
        # particles: same per-particle arithmetic
        assert_rows_close(pg.array, o_parts, REL)
        scale = max(np.max(np.abs(o_e1)), np.max(np.abs(o_e2)), 1.0)
        assert np.max(np.abs(h.e_dofs[0] - o_e1)) <= REL * scale
        assert np.max(np.abs(h.e_dofs[1] - o_e2)) <= REL * scale
        assert np.max(np.abs(h.b_dofs - o_b)) <= REL * max(np.max(np.abs(o_b)), 1.0)
        for c in (1, 2):
            jr = osp.j(c)
            assert np.max(np.abs(h.j_dofs[c - 1] - jr)) <= REL * max(np.max(np.abs(jr)), 1e-300) + 1e-300


def test_gpu_hp1_first_op_is_bitwise_per_particle():
    """From identical inputs, operatorHp1 updates x and v2 with the same
    rounded operations as the oracle: bitwise equal particles."""
    w = _hot_weibel(50000, seed=3)
    h, pg, osp, o_parts, _ = _setup_pair(w)
    G.operatorHp1(h, 0.3)
    osp.operatorHp1(0.3)
    np.testing.assert_array_equal(pg.array, o_parts)


def test_gpu_strang_fused_equals_substep_and_oracle():
    w = _hot_weibel(8192, seed=5)
    hf, pgf, osp, o_parts, (o_e1, o_e2, o_b) = _setup_pair(w)
    hs, pgs, *_ = _setup_pair(w)
    G.strang_splitting_(hf, 0.1, 4)
    G.strang_splitting_(hs, 0.1, 4, substep=True)
    osp.strang_splitting(0.1, 4)
    # fused and sub-step GPU paths sum the deposits in the same order
    np.testing.assert_array_equal(pgf.array, pgs.array)
    np.testing.assert_array_equal(hf.e_dofs[0], hs.e_dofs[0])
    np.testing.assert_array_equal(hf.e_dofs[1], hs.e_dofs[1])
    np.testing.assert_array_equal(hf.b_dofs, hs.b_dofs)
    assert_rows_close(pgf.array, o_parts, 1e-10)
    assert np.max(np.abs(hf.e_dofs[0] - o_e1)) <= 1e-10 * max(np.max(np.abs(o_e1)), 1.0)
    assert np.max(np.abs(hf.e_dofs[1] - o_e2)) <= 1e-10 * max(np.max(np.abs(o_e2)), 1.0)
    assert np.max(np.abs(hf.b_dofs - o_b)) <= 1e-10 * max(np.max(np.abs(o_b)), 1.0)


def test_gpu_1d1v_matches_oracle():
    w = workloads.landau_1d1v(20000, seed=2)
    h, pg, osp, o_parts, (o_e1, _, _) = _setup_pair(w, 3, 2)
    dt = 0.05
    for name in ("operatorHB", "operatorHp1", "operatorHB"):
        getattr(G, name)(h, dt)
        getattr(osp, name)(dt)
        assert_rows_close(pg.array, o_parts, REL)
        assert np.max(np.abs(h.e_dofs[0] - o_e1)) <= REL * max(np.max(np.abs(o_e1)), 1.0)
    G.strang_splitting_(h, dt, 3)
    osp.strang_splitting(dt, 3)
    assert_rows_close(pg.array, o_parts, 1e-10)
    hs, pgs, *_ = _setup_pair(w, 3, 2)
    G.operatorHB(hs, dt)
    G.operatorHp1(hs, dt)
    G.operatorHB(hs, dt)
    G.strang_splitting_(hs, dt, 3, substep=True)
    np.testing.assert_array_equal(pg.array, pgs.array)


def test_gpu_deterministic_runs():
    w = _hot_weibel(30000, seed=8)
    a, pga, *_ = _setup_pair(w)
    b, pgb, *_ = _setup_pair(w)
    G.strang_splitting_(a, 0.05, 3)
    G.strang_splitting_(b, 0.05, 3)
    np.testing.assert_array_equal(pga.array, pgb.array)
    np.testing.assert_array_equal(a.e_dofs[0], b.e_dofs[0])
    np.testing.assert_array_equal(a.b_dofs, b.b_dofs)


def test_gpu_solve_poisson_and_diagnostics():
    w = workloads.weibel_1d2v(20000, seed=4)
    h, pg, osp, o_parts, (o_e1, o_e2, o_b) = _setup_pair(w)
    nx = w["nx"]
    ep = np.zeros(nx)
    rho = np.zeros(nx)
    G.solve_poisson_(ep, h, rho)
    ep_o = np.zeros(nx)
    rho_o = np.zeros(nx)
    osp.solve_poisson(ep_o, rho_o)
    assert np.max(np.abs(rho - rho_o)) <= REL * np.max(np.abs(rho_o))
    assert np.max(np.abs(ep - ep_o)) <= 1e-11 * np.max(np.abs(ep_o))
    # total charge: sum(rho) = sum(q) (partition of unity, galerkin scaling 1)
    assert abs(rho.sum() - (w["particles"][3] * w["common_weight"]).sum()) < 1e-12 * abs(rho.sum())
    th = G.TimeHistoryDiagnostics(h)
    out = G.write_step_(th, 0.0, 3, efield_poisson=ep)
    ref = osp.write_step(0.0, 3, o_e1, o_e2, o_b, o_e1, o_e2, ep_o)
    for k in range(1, 10):
        assert abs(out[k] - ref[k]) <= 1e-10 * max(abs(ref[k]), 1e-30), (k, out[k], ref[k])
    assert abs(out[10] - ref[10]) <= 1e-12 * max(np.max(np.abs(o_e1)), 1e-30)


def test_gpu_maxwell_verbs():
    nx = 32
    L = 5.0
    mesh = G.OneDGrid(0.0, L, nx)
    m = G.Maxwell1DFEM(mesh, 3)
    om = orc.Maxwell1D(0.0, L, nx, 3)
    rng = np.random.default_rng(0)
    for k, v in om.tables().items():
        np.testing.assert_array_equal(getattr(m, k), v)
    rhs = rng.standard_normal(nx)
    for comp in (1, 2):
        e = rng.standard_normal(nx)
        eo = e.copy()
        G.compute_e_from_j_(e, m, rhs, comp)
        om.compute_e_from_j(eo, rhs, comp)
        np.testing.assert_allclose(e, eo, rtol=0, atol=1e-13 * np.max(np.abs(eo)))
    e = rng.standard_normal(nx)
    eo = e.copy()
    G.compute_e_from_b_(e, m, 0.3, rhs)
    om.compute_e_from_b(eo, 0.3, rhs)
    np.testing.assert_allclose(e, eo, rtol=0, atol=1e-13 * np.max(np.abs(eo)))
    b = rng.standard_normal(nx)
    bo = b.copy()
    G.compute_b_from_e_(b, m, 0.3, rhs)
    om.compute_b_from_e(bo, 0.3, rhs)
    np.testing.assert_array_equal(b, bo)
    e = np.zeros(nx)
    eo = np.zeros(nx)
    G.compute_e_from_rho_(e, m, rhs - rhs.mean())
    om.compute_e_from_rho(eo, rhs - rhs.mean())
    np.testing.assert_allclose(e, eo, rtol=0, atol=1e-13 * np.max(np.abs(eo)))
    c2 = rng.standard_normal(nx)
    for deg in (2, 3):
        assert abs(G.inner_product(m, rhs, c2, deg) - om.inner_product(rhs, c2, deg)) < 1e-13
        assert abs(G.l2norm_squared(m, rhs, deg) - om.l2norm_squared(rhs, deg)) < 1e-13


def test_gpu_nan_position_raises_inexact():
    w = workloads.weibel_1d2v(1000, seed=1)
    w["particles"][0, 17] = float("nan")
    nx = w["nx"]
    mesh = G.OneDGrid(w["xmin"], w["xmax"], nx)
    pg = G.ParticleGroup(1, 2, 1000)
    pg.array[:] = w["particles"]
    ks0 = G.ParticleMeshCoupling1D(mesh, 1000, 3, "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, 1000, 2, "galerkin")
    mx = G.Maxwell1DFEM(mesh, 3)
    h = G.HamiltonianSplitting(1, 2, mx, ks0, ks1, pg, [np.zeros(nx), np.zeros(nx)], np.zeros(nx))
    with pytest.raises(G.InexactError):
        G.operatorHE(h, 0.1)
    # the diagnostics evaluate the fields at the positions too
    with pytest.raises(G.InexactError):
        G.solve_poisson_(np.zeros(nx), h, np.zeros(nx))
    th = G.TimeHistoryDiagnostics(h)
    with pytest.raises(G.InexactError):
        G.write_step_(th, 0.0, 3, efield_poisson=np.zeros(nx))
    # the error flag is cleared: a finite state runs again
    pg.array[0, 17] = 0.5
    out = G.write_step_(th, 0.0, 3, efield_poisson=np.zeros(nx))
    assert np.all(np.isfinite(out))


# ---------------------------------------------------------------------------
# Full-size properties (4,194,304 particles, the 1D2V Weibel workload)

def test_gpu_weibel_4m_properties():
    n = 4194304
    w = workloads.weibel_1d2v(n, seed=1)
    nx = w["nx"]
    mesh = G.OneDGrid(w["xmin"], w["xmax"], nx)
    pg = G.ParticleGroup(1, 2, n, common_weight=w["common_weight"])
    pg.array[:] = w["particles"]
    ks0 = G.ParticleMeshCoupling1D(mesh, n, 3, "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, n, 2, "galerkin")
    mx = G.Maxwell1DFEM(mesh, 3)
    e1, e2, b = np.zeros(nx), np.zeros(nx), w["b0"].copy()
    h = G.HamiltonianSplitting(1, 2, mx, ks0, ks1, pg, [e1, e2], b, sync=False)
    rho = np.zeros(nx)
    G.solve_poisson_(e1, h, rho)
    h.to_device(particles=False)
    qsum = (w["particles"][3] * w["common_weight"]).sum()
    assert abs(rho.sum() - qsum) < 1e-12 * qsum
    th = G.TimeHistoryDiagnostics(h)
    d0 = G.write_step_(th, 0.0, 3)
    G.strang_splitting_(h, w["dt"], 10)
    d1 = G.write_step_(th, 0.5, 3)
    energy = lambda d: 0.5 * (d[1] + d[4] + d[5] + d[6])
    # energy drift of the Strang scheme over 10 steps stays small
    assert abs(energy(d1) - energy(d0)) < 1e-6 * energy(d0)
    # charge conservation: total rho unchanged
    rho1 = np.zeros(nx)
    ep = np.zeros(nx)
    G.solve_poisson_(ep, h, rho1)
    assert abs(rho1.sum() - qsum) < 1e-12 * qsum
    # Gauss law: E1 from Ampere tracks the Poisson field up to its mean (mode
    # 0 follows the net current, which Poisson cannot see) and the trunc()
    # quirk of particles leaving through x < 0
    h.to_host(particles=False)
    d = h.e_dofs[0] - ep
    assert np.max(np.abs(d - d.mean())) < 1e-5 * np.max(np.abs(ep))


@pytest.mark.parametrize("nx,degree", [(16, 2), (32, 3), (10, 2)])
def test_gpu_rhs_from_function_and_l2projection(nx, degree):
    """compute_rhs_from_function! and l2projection! (maxwell_1d_fem.jl:188-220,
    347-374) through the library: the function is called back on the host at
    the closed-form Gauss-Legendre points; bitwise equal to the oracle for the
    same function (written as the oracle evaluates it), and the reference's
    rhs / dx result of l2projection!."""
    L = 2 * math.pi / 1.25
    mesh = G.OneDGrid(0.0, L, nx)
    m = G.Maxwell1DFEM(mesh, 3)
    om = orc.Maxwell1D(0.0, L, nx, 3)
    a_cos, a_sin, k, c = 0.3, -0.2, 1.25, 0.1

    def f(x):
        v = 0.0
        v += a_cos * math.cos(k * x)
        v += a_sin * math.sin(k * x)
        return v + c
    rhs = np.zeros(nx)
    G.compute_rhs_from_function_(rhs, m, f, degree)
    ref = np.zeros(nx)
    om.compute_rhs_from_function(ref, degree, a_cos=a_cos, a_sin=a_sin, k=k, c=c)
    np.testing.assert_array_equal(rhs, ref)
    proj = np.zeros(nx)
    G.l2projection_(proj, m, f, degree)
    ref_p = np.zeros(nx)
    om.l2projection(ref_p, degree, a_cos=a_cos, a_sin=a_sin, k=k, c=c)
    np.testing.assert_array_equal(proj, ref_p)
    np.testing.assert_array_equal(proj, rhs / m.delta_x)
    # the benchmark's initial B (host set-up, the same quadrature)
    b0 = workloads.l2projection_cos(nx, L, degree, 1e-4, 1.25)
    G.l2projection_(proj, m, lambda x: 1e-4 * math.cos(1.25 * x), degree)
    np.testing.assert_allclose(b0, proj, rtol=0, atol=1e-19)
    with pytest.raises(ValueError):
        G.l2projection_(proj, m, f, 1)


def test_gpu_save_and_resume(tmp_path):
    """save(file, step, pg) (particle_group.jl:152-165) after 5 steps, then a
    new splitting resumed from the file (and the fields) runs 4 more steps:
    the state of 9 uninterrupted steps, up to the deposits' summation order
    (the resumed splitting sorts its particles by their positions after 5
    steps, the uninterrupted one kept the order of its upload)."""
    w = _hot_weibel(8192, seed=21)
    a, pga, *_ = _setup_pair(w)
    G.strang_splitting_(a, 0.1, 9)
    b, pgb, *_ = _setup_pair(w)
    G.strang_splitting_(b, 0.1, 5)
    path = G.save(str(tmp_path / "weibel"), 5, pgb)
    assert path.endswith("weibel-000005.npz")
    fields = [b.e_dofs[0].copy(), b.e_dofs[1].copy(), b.b_dofs.copy()]
    nx = w["nx"]
    mesh = G.OneDGrid(w["xmin"], w["xmax"], nx)
    pgc = G.ParticleGroup(1, 2, 8192, common_weight=w["common_weight"])
    G.load(path, pgc)
    np.testing.assert_array_equal(pgc.array, pgb.array)
    c = G.HamiltonianSplitting(1, 2, G.Maxwell1DFEM(mesh, 3),
                               G.ParticleMeshCoupling1D(mesh, 8192, 3, "galerkin"),
                               G.ParticleMeshCoupling1D(mesh, 8192, 2, "galerkin"), pgc,
                               fields[:2], fields[2])
    G.strang_splitting_(c, 0.1, 4)
    assert_rows_close(pgc.array, pga.array, 1e-11)
    for got, ref in ((c.e_dofs[0], a.e_dofs[0]), (c.e_dofs[1], a.e_dofs[1]), (c.b_dofs, a.b_dofs)):
        assert np.max(np.abs(got - ref)) <= 1e-11 * max(np.max(np.abs(ref)), 1.0)
