"""GPU parity of HamiltonianSplittingBoris (hamiltonian_splitting_boris.jl)
through the C ABI: the reference's Boris known answers, every pusher
against the oracle, whole runs, and an exact gyration at full size.

Tolerances: the E kick and the B rotation use the oracle's operations in its
order (bitwise per particle); the x push is bitwise too. The current
deposits differ from the reference's loop only in summation order and are
compared at 1e-12 of the largest value; fields and particles after steps
that use them, at 1e-12 row-normwise.
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


def _rows_close(got, ref, rel):
    for r in range(ref.shape[0]):
        scale = max(np.max(np.abs(ref[r])), 1e-300)
        err = np.max(np.abs(got[r] - ref[r]))
        assert err <= rel * scale, f"row {r}: {err} > {rel} * {scale}"


def _close(got, ref, rel=REL, floor=1.0):
    scale = max(np.max(np.abs(ref)), floor)
    assert np.max(np.abs(got - ref)) <= rel * scale, (np.max(np.abs(got - ref)), scale)


def _pair(parts, nx, xmin, xmax, e1, e2, b, degree0=3, degree1=2, cw=None, sync=True):
    """A GPU Boris splitting and an oracle splitting on copies of one state."""
    n = parts.shape[1]
    cw = 1.0 / n if cw is None else cw
    mesh = G.OneDGrid(xmin, xmax, nx)
    pg = G.ParticleGroup(1, 2, n, common_weight=cw)
    pg.array[:] = parts
    ks0 = G.ParticleMeshCoupling1D(mesh, n, degree0, "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, n, degree1, "galerkin")
    mx = G.Maxwell1DFEM(mesh, 3)
    ge = [e1.copy(), e2.copy()]
    gb = b.copy()
    h = G.HamiltonianSplittingBoris(mx, ks0, ks1, pg, ge, gb, sync=sync)
    o_parts = np.asfortranarray(parts.copy())
    o_e1, o_e2, o_b = e1.copy(), e2.copy(), b.copy()
    osp = orc.Splitting(nx, xmin, xmax, 3, degree0, degree1, o_parts, o_e1, o_e2, o_b,
                        common_weight=cw)
    return h, pg, osp, o_parts, (o_e1, o_e2, o_b)


def test_gpu_boris_kat():
    kat = golden("kat_boris.json")
    n, nc, deg = kat["n_particles"], kat["num_cells"], kat["degree_smoother"]
    xmax = kat["eta_max_over_pi"] * math.pi
    parts = np.array(kat["particle_info_initial"]).reshape(n, 4).T
    # E1 from Poisson of the particles' charge (:61-71), E2 = B = 1
    mesh = G.OneDGrid(kat["eta_min"], xmax, nc)
    ks0 = G.ParticleMeshCoupling1D(mesh, n, deg, "galerkin")
    rho = np.zeros(nc)
    G.add_charge_(rho, ks0, parts[0].copy(), parts[3] * 1.0)
    e1 = np.zeros(nc)
    G.compute_e_from_rho_(e1, G.Maxwell1DFEM(mesh, deg), rho)
    h, pg, osp, o_parts, (o_e1, o_e2, o_b) = _pair(parts, nc, kat["eta_min"], xmax, e1,
                                                    np.ones(nc), np.ones(nc), deg, deg - 1,
                                                    cw=1.0)
    dt = kat["delta_t"]
    G.staggering_(h, 0.5 * dt)
    G.strang_splitting_(h, dt, 1)
    ref = np.array(kat["particle_info_ref"]).reshape(n, 4).T
    np.testing.assert_allclose(pg.array, ref, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(h.b_dofs, kat["bfield_ref"], rtol=0, atol=1e-14)
    eref = np.array(kat["efield_ref"]).reshape(2, nc)
    np.testing.assert_allclose(h.e_dofs[0], eref[0], rtol=0, atol=1e-14)
    np.testing.assert_allclose(h.e_dofs[1], eref[1], rtol=0, atol=1e-14)


def _hot(n, seed, nx=32):
    w = workloads.weibel_1d2v(n, seed=seed, nx=nx, sigma1=0.6, sigma2=0.4)
    rng = np.random.default_rng(seed)
    e1, e2 = 0.2 * rng.normal(size=nx), 0.2 * rng.normal(size=nx)
    b = w["b0"] * 1e3 + 0.3
    return np.asfortranarray(w["particles"]), w["xmin"], w["xmax"], e1, e2, b, nx


@pytest.mark.parametrize("pair", [(3, 2), (3, 3), (2, 1), (1, 0)])
def test_gpu_boris_pushers_match_oracle(pair):
    parts, xmin, xmax, e1, e2, b, nx = _hot(20000, seed=5)
    h, pg, osp, o_parts, (o_e1, o_e2, o_b) = _pair(parts, nx, xmin, xmax, e1, e2, b, *pair)
    rng = np.random.default_rng(9)
    mids = [0.3 * rng.normal(size=nx), 0.3 * rng.normal(size=nx), 1.0 + rng.normal(size=nx)]
    h.e_dofs_mid[0][:], h.e_dofs_mid[1][:], h.b_dofs_mid[:] = mids
    osp.boris_set_mid(*mids)
    G.push_v_epart_(h, 0.3)
    osp.boris_push_v_epart(0.3)
    np.testing.assert_array_equal(pg.array, o_parts)
    G.push_v_bpart_(h, 0.4)
    osp.boris_push_v_bpart(0.4)
    np.testing.assert_array_equal(pg.array, o_parts)
    G.push_x_accumulate_j_(h, 0.35)
    osp.boris_push_x_accumulate_j(0.35)
    np.testing.assert_array_equal(pg.array, o_parts)
    for c in (1, 2):
        _close(h.j_dofs[c - 1], osp.j(c), floor=1e-300)


def _check_state(h, pg, osp, o_parts, o_fields):
    o_e1, o_e2, o_b = o_fields
    _rows_close(pg.array, o_parts, REL)
    _close(h.e_dofs[0], o_e1)
    _close(h.e_dofs[1], o_e2)
    _close(h.b_dofs, o_b)
    om = osp.boris_mid()
    _close(h.e_dofs_mid[0], om[0])
    _close(h.e_dofs_mid[1], om[1])
    _close(h.b_dofs_mid, om[2])
    for c in (1, 2):
        _close(h.j_dofs[c - 1], osp.j(c), floor=1e-300)


@pytest.mark.parametrize("pair", [(3, 2), (2, 1)])
def test_gpu_boris_step_matches_oracle_hot(pair):
    """staggering! and one step on a state whose particles cross cells and
    leave the domain on both sides (this state is unstable under Boris, so
    longer runs amplify the deposit's rounding; see the Weibel run below)."""
    parts, xmin, xmax, e1, e2, b, nx = _hot(30000, seed=7)
    h, pg, osp, o_parts, fields = _pair(parts, nx, xmin, xmax, e1, e2, b, *pair)
    G.staggering_(h, 0.2)
    osp.boris_staggering(0.2)
    _check_state(h, pg, osp, o_parts, fields)
    G.strang_splitting_(h, 0.2, 1)
    osp.boris_strang(0.2, 1)
    _check_state(h, pg, osp, o_parts, fields)


@pytest.mark.parametrize("pair", [(3, 2), (2, 1)])
def test_gpu_boris_weibel_run_matches_oracle(pair):
    """The 1D2V Weibel initial condition (E1 from Poisson), staggering! and
    1 + 20 steps in two calls (the second keeps the state on the GPU between
    steps and sorts particles every 20 steps)."""
    w = workloads.weibel_1d2v(40000, seed=7)
    nx = w["nx"]
    parts = np.asfortranarray(w["particles"])
    e1 = np.zeros(nx)
    o = orc.Splitting(nx, w["xmin"], w["xmax"], 3, pair[0], pair[1], parts.copy(order="F"), e1,
                      np.zeros(nx), w["b0"].copy(), common_weight=w["common_weight"])
    o.solve_poisson(e1, np.zeros(nx))
    h, pg, osp, o_parts, fields = _pair(parts, nx, w["xmin"], w["xmax"], e1, np.zeros(nx),
                                        w["b0"], *pair, cw=w["common_weight"])
    G.staggering_(h, w["dt"])
    osp.boris_staggering(w["dt"])
    for steps in (1, 20):
        G.strang_splitting_(h, w["dt"], steps)
        osp.boris_strang(w["dt"], steps)
        _check_state(h, pg, osp, o_parts, fields)


def test_gpu_boris_deterministic_and_needs_v2():
    parts, xmin, xmax, e1, e2, b, nx = _hot(50000, seed=8)
    outs = []
    for _ in range(2):
        h, pg, *_ = _pair(parts, nx, xmin, xmax, e1, e2, b)
        G.staggering_(h, 0.1)
        G.strang_splitting_(h, 0.1, 3)
        outs.append((pg.array.copy(), h.e_dofs_mid[1].copy(), h.j_dofs[0].copy()))
    for a, c in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, c)
    mesh = G.OneDGrid(0.0, 1.0, 16)
    pg = G.ParticleGroup(1, 1, 64)
    pg.array[0] = np.linspace(0.0, 0.9, 64)
    ks0 = G.ParticleMeshCoupling1D(mesh, 64, 3, "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, 64, 2, "galerkin")
    with pytest.raises(AssertionError):
        G.HamiltonianSplittingBoris(G.Maxwell1DFEM(mesh, 3), ks0, ks1, pg,
                                    [np.zeros(16), np.zeros(16)], np.zeros(16))


def test_gpu_boris_gyration_4m():
    """4,194,304 particles in a uniform B with negligible self-fields: every
    velocity turns by exactly 2 atan(qm B dt / 2) per step (size-independent
    check at the benchmark's particle count)."""
    n, nx, L, dt, B0, steps = 1 << 22, 32, 2 * math.pi / 1.25, 0.05, 1.5, 20
    rng = np.random.default_rng(4)
    v0 = rng.normal(size=(2, n))
    parts = np.asfortranarray(np.vstack([L * rng.random(n), v0, np.ones(n)]))
    h, pg, *_ = _pair(parts, nx, 0.0, L, np.zeros(nx), np.zeros(nx), np.full(nx, B0),
                      cw=1e-30, sync=False)
    G.staggering_(h, 0.5 * dt)
    G.strang_splitting_(h, dt, steps)
    h.to_host()
    th = 2.0 * math.atan(B0 * dt / 2.0)
    c, s = math.cos(steps * th), math.sin(steps * th)
    np.testing.assert_allclose(pg.array[1], c * v0[0] + s * v0[1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(pg.array[2], -s * v0[0] + c * v0[1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(h.b_dofs, B0, rtol=0, atol=1e-14)
