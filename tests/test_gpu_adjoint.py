"""GPU reverse mode of the 1D2V Strang splitting through the C ABI, against
the oracle's reverse mode (itself pinned by central differences) and against
central differences of the GPU's own forward run.

Tolerances: the GPU forward equals the oracle's to the deposit summation
order (1e-12), and the cotangents inherit that: GPU and oracle gradients
agree to 1e-9 of each group's largest entry. Against central differences
(h = 1e-6 of the parameter's scale) the directional derivatives agree to
1e-6 relative.
"""
import math

import numpy as np
import pytest

import gempic_amd as G
from gempic_amd import workloads
import pyoracle as orc

pytestmark = pytest.mark.gpu

REL = 1e-9


def _hot(n, nx=16, seed=3, vscale=0.6):
    rng = np.random.default_rng(seed)
    L = 4 * math.pi
    parts = np.asfortranarray(np.vstack([L * rng.random(n), vscale * rng.normal(size=n),
                                         0.5 * rng.normal(size=n), rng.random(n) + 0.5]))
    e1, e2, b = (0.3 * rng.normal(size=nx) for _ in range(3))
    return parts, e1, e2, b, L, nx


def _gpu(parts, e1, e2, b, L, nx, deg, cw, sync=True):
    n = parts.shape[1]
    mesh = G.OneDGrid(0.0, L, nx)
    pg = G.ParticleGroup(1, 2, n, common_weight=cw)
    pg.array[:] = parts
    ks0 = G.ParticleMeshCoupling1D(mesh, n, deg[0], "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, n, deg[1], "galerkin")
    h = G.HamiltonianSplitting(1, 2, G.Maxwell1DFEM(mesh, 3), ks0, ks1, pg,
                               [e1.copy(), e2.copy()], b.copy(), sync=sync)
    return h, pg


def _close(got, ref, rel=REL):
    scale = max(np.max(np.abs(ref)), 1e-300)
    err = np.max(np.abs(got - ref))
    assert err <= rel * scale, (err, scale)


@pytest.mark.parametrize("deg", [(3, 2), (2, 1)])
@pytest.mark.parametrize("K", [1, 4, 6])
def test_gpu_adjoint_matches_oracle(deg, K):
    """Six steps of a hot state (particles cross cells and the domain's
    ends), with cotangents on the final state and the magnetic-energy loss;
    K = 4 re-runs an interval of four steps and a partial one of two."""
    parts, e1, e2, b, L, nx = _hot(20000)
    n = parts.shape[1]
    dt, steps, wl, cw = 0.1, 6, 0.1, 0.01
    rng = np.random.default_rng(11)
    cp = np.zeros_like(parts, order="F")
    cp[:3] = rng.normal(size=(3, n))
    ct = (rng.normal(size=nx), rng.normal(size=nx), rng.normal(size=nx))
    h, pg = _gpu(parts, e1, e2, b, L, nx, deg, cw)
    r = G.strang_splitting_adjoint_(h, dt, steps, checkpoint_every=K, loss_weight=wl,
                                    ct_particles=cp, ct_e1=ct[0], ct_e2=ct[1], ct_b=ct[2])
    o_parts = parts.copy(order="F")
    o_f = [e1.copy(), e2.copy(), b.copy()]
    osp = orc.Splitting(nx, 0.0, L, 3, deg[0], deg[1], o_parts, *o_f, common_weight=cw)
    J, gp, ge1, ge2, gb = osp.strang_adjoint(dt, steps, wl, cp, *ct)
    # the forward state after the call
    np.testing.assert_allclose(pg.array, o_parts, rtol=0, atol=1e-11)
    assert abs(r.loss - J) <= 1e-12 * abs(J)
    for row in range(3):
        _close(r.g_particles[row], gp[row])
    np.testing.assert_array_equal(r.g_particles[3], 0.0)
    _close(r.g_e1, ge1)
    _close(r.g_e2, ge2)
    _close(r.g_b, gb)
    # the explicit derivatives in the domain length and the weights
    gl, gw = osp.strang_adjoint_domain()
    assert abs(r.g_length - gl) <= REL * abs(gl), (r.g_length, gl)
    assert abs(r.g_weight_sum - gw) <= REL * abs(gw), (r.g_weight_sum, gw)


def test_gpu_adjoint_enters_with_a_sort_due():
    """The handle enters the reverse mode with a sort already due, and the
    sort interval spans two checkpoint intervals: a sort then runs at the
    first interval boundary, so two sort epochs are needed although the
    interval would allow only one sort after the first. The results match
    the oracle's reverse mode from the same state."""
    parts, e1, e2, b, L, nx = _hot(20000, seed=6)
    dt, cw = 0.1, 0.01
    h, pg = _gpu(parts, e1, e2, b, L, nx, (3, 2), cw, sync=False)
    h.set_sort_interval(0)
    G.strang_splitting_(h, dt, 5)                 # five steps since the upload's sort
    h.to_host()                                   # (downloads do not sort)
    start = pg.array.copy(order="F")
    f0 = [h.e_dofs[0].copy(), h.e_dofs[1].copy(), h.b_dofs.copy()]
    h.set_sort_interval(5)
    r = G.strang_splitting_adjoint_(h, dt, 6, checkpoint_every=3, loss_weight=0.1)
    osp = orc.Splitting(nx, 0.0, L, 3, 3, 2, start, *f0, common_weight=cw)
    J, gp, ge1, ge2, gb = osp.strang_adjoint(dt, 6, 0.1)
    assert abs(r.loss - J) <= 1e-12 * abs(J)
    for row in range(3):
        _close(r.g_particles[row], gp[row])
    _close(r.g_e1, ge1)
    _close(r.g_b, gb)


def test_gpu_adjoint_loss_only_and_deterministic():
    parts, e1, e2, b, L, nx = _hot(30000, seed=4)
    outs = []
    for _ in range(2):
        h, pg = _gpu(parts, e1, e2, b, L, nx, (3, 2), 0.01)
        r = G.strang_splitting_adjoint_(h, 0.1, 5, checkpoint_every=2)
        outs.append(r)
    for a in ("g_particles", "g_e1", "g_e2", "g_b"):
        np.testing.assert_array_equal(getattr(outs[0], a), getattr(outs[1], a))
    assert outs[0].loss == outs[1].loss
    o_parts = parts.copy(order="F")
    osp = orc.Splitting(nx, 0.0, L, 3, 3, 2, o_parts, e1.copy(), e2.copy(), b.copy(),
                        common_weight=0.01)
    J, gp, ge1, ge2, gb = osp.strang_adjoint(0.1, 5, 0.1)
    for row in range(3):
        _close(outs[0].g_particles[row], gp[row])
    _close(outs[0].g_b, gb)


def test_gpu_solve_poisson_adjoint():
    parts, e1, e2, b, L, nx = _hot(50000, seed=9)
    h, pg = _gpu(parts, e1, e2, b, L, nx, (3, 2), 0.001)
    ct = np.random.default_rng(5).normal(size=nx)
    g = G.solve_poisson_adjoint(h, ct)
    osp = orc.Splitting(nx, 0.0, L, 3, 3, 2, parts.copy(order="F"), e1.copy(), e2.copy(),
                        b.copy(), common_weight=0.001)
    np.testing.assert_allclose(g, osp.solve_poisson_adjoint(ct), rtol=1e-12, atol=1e-15)


def _weibel_loss_and_grad(w, alpha_shift=0.0, beta_shift=0.0, steps=40, K=10, fetch=True):
    """Loss of the Weibel case with E1 from Poisson; the particle positions
    move with alpha to first order through dx/dalpha (exact for the test's
    purposes: the same linearisation on both sides of the difference)."""
    n, nx = w["particles"].shape[1], w["nx"]
    parts = w["particles"].copy(order="F")
    parts[0] = parts[0] + alpha_shift * w["dx_dalpha"]
    b0 = w["b0"] + beta_shift * w["b_pattern"]
    h, pg = _gpu(parts, np.zeros(nx), np.zeros(nx), b0, w["xmax"], nx, (3, 2), w["common_weight"])
    e1, rho = np.zeros(nx), np.zeros(nx)
    G.solve_poisson_(e1, h, rho)
    h.e_dofs[0][:] = e1
    r = G.strang_splitting_adjoint_(h, w["dt"], steps, checkpoint_every=K)
    gpois = G.solve_poisson_adjoint(_gpu(parts, e1, np.zeros(nx), b0, w["xmax"], nx, (3, 2),
                                         w["common_weight"])[0], r.g_e1)
    return r, gpois


def test_gpu_adjoint_weibel_parameter_gradients_vs_central_differences():
    """dJ/dbeta and dJ/dalpha of the time-integrated magnetic energy of a
    200,000-particle Weibel run (40 steps, checkpoints every 10) against
    central differences of the GPU's own forward run."""
    w = workloads.weibel_adjoint_1d2v(200000, seed=2, alpha=1e-2, beta=1e-3)
    r, gpois = _weibel_loss_and_grad(w)
    dj_dbeta = float(r.g_b @ w["b_pattern"])
    dj_dalpha = float((r.g_particles[0] + gpois) @ w["dx_dalpha"])
    hb = 1e-6 * w["beta"]
    fd_b = (_weibel_loss_and_grad(w, beta_shift=hb)[0].loss -
            _weibel_loss_and_grad(w, beta_shift=-hb)[0].loss) / (2 * hb)
    assert abs(fd_b - dj_dbeta) <= 1e-6 * abs(fd_b), (fd_b, dj_dbeta)
    ha = 1e-6 * w["alpha"]
    fd_a = (_weibel_loss_and_grad(w, alpha_shift=ha)[0].loss -
            _weibel_loss_and_grad(w, alpha_shift=-ha)[0].loss) / (2 * ha)
    assert abs(fd_a - dj_dalpha) <= 1e-6 * abs(fd_a), (fd_a, dj_dalpha)


def test_gpu_adjoint_velocity_gradient_vs_central_differences():
    """A random direction in all initial velocities (200,000 particles, 40
    steps): <g_v, d> against the central difference of the GPU loss."""
    w = workloads.weibel_adjoint_1d2v(200000, seed=3, alpha=1e-2, beta=1e-3)
    nx = w["nx"]
    d = np.random.default_rng(1).normal(size=(2, w["particles"].shape[1])) * 0.02

    def run(eps, fetch=True):
        parts = w["particles"].copy(order="F")
        parts[1:3] += eps * d
        h, pg = _gpu(parts, np.zeros(nx), np.zeros(nx), w["b0"], w["xmax"], nx, (3, 2),
                     w["common_weight"])
        return G.strang_splitting_adjoint_(h, w["dt"], 40, checkpoint_every=10)
    r = run(0.0)
    ad = float(np.sum(r.g_particles[1:3] * d))
    he = 1e-6
    fd = (run(he).loss - run(-he).loss) / (2 * he)
    assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)


@pytest.mark.parametrize("K", [2, 3])
def test_gpu_adjoint_across_particle_sorts(K):
    """With a sort interval of one step the particles are re-sorted at every
    checkpoint after the first; the cotangents follow the particles through
    each re-sort back to the caller's order."""
    parts, e1, e2, b, L, nx = _hot(20000, seed=6, vscale=1.5)
    dt, steps, wl, cw = 0.1, 7, 0.1, 0.01
    h, pg = _gpu(parts, e1, e2, b, L, nx, (3, 2), cw)
    h.set_sort_interval(1)
    r = G.strang_splitting_adjoint_(h, dt, steps, checkpoint_every=K, loss_weight=wl)
    o_parts = parts.copy(order="F")
    osp = orc.Splitting(nx, 0.0, L, 3, 3, 2, o_parts, e1.copy(), e2.copy(), b.copy(),
                        common_weight=cw)
    J, gp, ge1, ge2, gb = osp.strang_adjoint(dt, steps, wl)
    np.testing.assert_allclose(pg.array, o_parts, rtol=0, atol=1e-11)
    assert abs(r.loss - J) <= 1e-12 * abs(J)
    for row in range(3):
        _close(r.g_particles[row], gp[row])
    _close(r.g_e1, ge1)
    _close(r.g_e2, ge2)
    _close(r.g_b, gb)


def test_gpu_adjoint_one_step_intervals_store_every_state():
    """checkpoint_every = 1: the fused pass stores each step's state as it
    goes (no re-run), across sorts (sort interval 3)."""
    parts, e1, e2, b, L, nx = _hot(20000, seed=8, vscale=1.5)
    dt, steps, wl, cw = 0.1, 8, 0.1, 0.01
    h, pg = _gpu(parts, e1, e2, b, L, nx, (3, 2), cw)
    h.set_sort_interval(3)
    r = G.strang_splitting_adjoint_(h, dt, steps, checkpoint_every=1, loss_weight=wl)
    o_parts = parts.copy(order="F")
    osp = orc.Splitting(nx, 0.0, L, 3, 3, 2, o_parts, e1.copy(), e2.copy(), b.copy(),
                        common_weight=cw)
    J, gp, ge1, ge2, gb = osp.strang_adjoint(dt, steps, wl)
    np.testing.assert_allclose(pg.array, o_parts, rtol=0, atol=1e-11)
    assert abs(r.loss - J) <= 1e-12 * abs(J)
    for row in range(3):
        _close(r.g_particles[row], gp[row])
    _close(r.g_e1, ge1)
    _close(r.g_e2, ge2)
    _close(r.g_b, gb)


def _k_run(k, n, steps, K, alpha=1e-2, beta=1e-3):
    w = workloads.weibel_adjoint_1d2v(n, seed=2, k=k, alpha=alpha, beta=beta)
    nx = w["nx"]
    h, pg = _gpu(w["particles"], np.zeros(nx), np.zeros(nx), w["b0"], w["xmax"], nx, (3, 2),
                 w["common_weight"])
    e1, rho = np.zeros(nx), np.zeros(nx)
    G.solve_poisson_(e1, h, rho)
    h.e_dofs[0][:] = e1
    r = G.strang_splitting_adjoint_(h, w["dt"], steps, checkpoint_every=K)
    return r, w, e1


def test_gpu_adjoint_wavenumber_gradient_vs_central_differences():
    """dJ/dk of the Weibel run, where k moves Lx = 2 pi / k, dx, the quiet-start
    positions (x = Lx xi), the weights (w = Lx) and E1 (Poisson, proportional
    to Lx): the explicit derivative in Lx from the reverse sweep plus the
    chain rule through the initial state, against central differences in k
    (200,000 particles, 40 steps)."""
    k = 1.25
    r, w, e1 = _k_run(k, 200000, 40, 10)
    L = w["xmax"]
    dj_dl = (r.g_length + r.g_weight_sum + float(r.g_particles[0] @ w["particles"][0]) / L +
             float(r.g_e1 @ e1) / L)
    dj_dk = -(L / k) * dj_dl
    hk = 1e-6 * k
    fd = (_k_run(k + hk, 200000, 40, 10)[0].loss - _k_run(k - hk, 200000, 40, 10)[0].loss) / (2 * hk)
    assert abs(fd - dj_dk) <= 1e-6 * abs(fd), (fd, dj_dk)


@pytest.mark.parametrize("nx", [4, 128])
def test_gpu_adjoint_grid_limits(nx):
    """The reverse mode at the smallest (4) and largest (128) grids it
    supports, against the oracle; 2 and 130 cells are rejected."""
    parts, e1, e2, b, L, _ = _hot(6000, nx=nx, seed=12)
    dt, steps, wl, cw = 0.05, 4, 0.05, 0.01
    h, pg = _gpu(parts, e1, e2, b, L, nx, (3, 2), cw)
    r = G.strang_splitting_adjoint_(h, dt, steps, checkpoint_every=2, loss_weight=wl)
    o_parts = parts.copy(order="F")
    osp = orc.Splitting(nx, 0.0, L, 3, 3, 2, o_parts, e1.copy(), e2.copy(), b.copy(),
                        common_weight=cw)
    J, gp, ge1, ge2, gb = osp.strang_adjoint(dt, steps, wl)
    assert abs(r.loss - J) <= 1e-12 * abs(J)
    for row in range(3):
        _close(r.g_particles[row], gp[row])
    _close(r.g_e2, ge2)
    _close(r.g_b, gb)
    for bad in (2, 130):
        pb, e1b, e2b, bb, Lb, _ = _hot(500, nx=bad, seed=1)
        hb, _ = _gpu(pb, e1b, e2b, bb, Lb, bad, (3, 2), cw)
        with pytest.raises(ValueError):
            G.strang_splitting_adjoint_(hb, dt, 2)


@pytest.mark.parametrize("deg", [(3, 2), (2, 1)])
def test_gpu_adjoint_matches_forward_mode(deg):
    """The GPU reverse mode against dual-number forward mode
    (oracle/tangent1d2v.py, itself equal to the oracle's reverse mode to
    1e-12 and to central differences of its forward run): the gradient in a
    random direction of every input group (positions, velocities, E1, E2, B)
    with cotangents on the final state, 6 steps of a hot state with a
    re-run interval (K = 4), to 1e-9 relative (the forwards differ in the
    deposit summation order only)."""
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle"))
    from tangent1d2v import Tangent1D2V
    parts, e1, e2, b, L, nx = _hot(20000, seed=12)
    n = parts.shape[1]
    dt, steps, wl, cw = 0.1, 6, 0.1, 0.01
    rng = np.random.default_rng(13)
    cp = np.zeros_like(parts, order="F")
    cp[:3] = rng.normal(size=(3, n))
    ct = (rng.normal(size=nx), rng.normal(size=nx), rng.normal(size=nx))
    d = [rng.normal(size=n) for _ in range(3)]
    dF = [rng.normal(size=nx) for _ in range(3)]
    h, pg = _gpu(parts, e1, e2, b, L, nx, deg, cw)
    r = G.strang_splitting_adjoint_(h, dt, steps, checkpoint_every=4, loss_weight=wl,
                                    ct_particles=cp, ct_e1=ct[0], ct_e2=ct[1], ct_b=ct[2])
    g_dot = (sum(float(np.dot(r.g_particles[row], d[row])) for row in range(3)) +
             float(np.dot(r.g_e1, dF[0])) + float(np.dot(r.g_e2, dF[1])) +
             float(np.dot(r.g_b, dF[2])))
    T = Tangent1D2V(nx, 0.0, L, deg[0], deg[1], common_weight=cw)
    _, dJ, _ = T.strang_jvp(parts, e1, e2, b, d, *dF, dt, steps, wl, cp, *ct)
    assert abs(g_dot - dJ) <= 1e-9 * abs(dJ), (g_dot, dJ)


def test_gpu_adjoint_config4_scale_vs_central_differences():
    """The adjoint configuration itself: 1,048,576 Weibel particles, 1000
    steps, checkpoints every 50, alpha = beta = 1e-4, k = 1.25; dJ/dbeta,
    dJ/dalpha, dJ/dk and a random direction in all initial velocities
    against central differences of the GPU's own forward run with
    h = 1e-6 |theta|.

    dJ/dbeta is checked to 1e-6. The other three only to 1e-3, the
    resolution of central differences at this horizon. The forward map is
    C^1 but not C^2 (the degree-2 smoother's second derivative jumps where a
    particle crosses a cell boundary), and over 1000 steps of the unstable
    run J' changes by up to 1e-4 across |s| <= 4e-6 |theta| without a
    smooth trend. Measured on the oracle (8192 particles, velocity
    direction), where the dual-number forward mode equals the reverse mode
    to 2e-12: the central quotient is off by 3e-6, 9e-8, 1e-9, 7e-6 and
    3e-7 relative at h = 1e-9, 1e-8, 1e-6, 4e-6 and 1e-5. The beta direction
    is smooth at this scale because the loss is dominated by the B seed.
    The derivative itself is checked exactly at the full 1000-step horizon
    against the oracle's reverse mode in
    test_gpu_adjoint_full_horizon_matches_oracle_reverse_mode."""
    n, steps, K = 1048576, 1000, 50
    print(f"adjoint at the configuration's size: n={n} steps={steps} K={K}", flush=True)
    w = workloads.weibel_adjoint_1d2v(n, seed=2)
    r, gpois = _weibel_loss_and_grad(w, steps=steps, K=K)
    dj_dbeta = float(r.g_b @ w["b_pattern"])
    dj_dalpha = float((r.g_particles[0] + gpois) @ w["dx_dalpha"])
    hb = 1e-6 * w["beta"]
    fd_b = (_weibel_loss_and_grad(w, beta_shift=hb, steps=steps, K=K)[0].loss -
            _weibel_loss_and_grad(w, beta_shift=-hb, steps=steps, K=K)[0].loss) / (2 * hb)
    ha = 1e-6 * w["alpha"]
    fd_a = (_weibel_loss_and_grad(w, alpha_shift=ha, steps=steps, K=K)[0].loss -
            _weibel_loss_and_grad(w, alpha_shift=-ha, steps=steps, K=K)[0].loss) / (2 * ha)
    # a random direction in the velocities
    nx = w["nx"]
    d = np.random.default_rng(4).normal(size=(2, n)) * 0.02

    def vrun(eps):
        parts = w["particles"].copy(order="F")
        parts[1:3] += eps * d
        h, pg = _gpu(parts, np.zeros(nx), np.zeros(nx), w["b0"], w["xmax"], nx, (3, 2),
                     w["common_weight"])
        e1 = np.zeros(nx)
        G.solve_poisson_(e1, h, np.zeros(nx))
        h.e_dofs[0][:] = e1
        return G.strang_splitting_adjoint_(h, w["dt"], steps, checkpoint_every=K)
    rv = vrun(0.0)
    ad = float(np.sum(rv.g_particles[1:3] * d))
    fd_v = (vrun(1e-6).loss - vrun(-1e-6).loss) / 2e-6
    # the wavenumber
    k = w["k"]
    rk, wk, e1k = _k_run(k, n, steps, K, 1e-4, 1e-4)
    L = wk["xmax"]
    dj_dl = (rk.g_length + rk.g_weight_sum + float(rk.g_particles[0] @ wk["particles"][0]) / L +
             float(rk.g_e1 @ e1k) / L)
    dj_dk = -(L / k) * dj_dl
    hk = 1e-6 * k
    fd_k = (_k_run(k + hk, n, steps, K, 1e-4, 1e-4)[0].loss -
            _k_run(k - hk, n, steps, K, 1e-4, 1e-4)[0].loss) / (2 * hk)
    print(f"beta {dj_dbeta!r} {fd_b!r}; alpha {dj_dalpha!r} {fd_a!r}; velocities {ad!r} "
          f"{fd_v!r}; k {dj_dk!r} {fd_k!r}", flush=True)
    assert abs(fd_b - dj_dbeta) <= 1e-6 * abs(fd_b), (fd_b, dj_dbeta)
    assert abs(fd_a - dj_dalpha) <= 1e-3 * abs(fd_a), (fd_a, dj_dalpha)
    assert abs(fd_v - ad) <= 1e-3 * abs(fd_v), (fd_v, ad)
    assert abs(fd_k - dj_dk) <= 1e-3 * abs(fd_k), (fd_k, dj_dk)


def test_gpu_adjoint_full_horizon_matches_oracle_reverse_mode():
    """The configuration's horizon (1000 steps, checkpoints every 50, the
    Weibel adjoint state with alpha = beta = 1e-4, k = 1.25) on 8192
    particles: the GPU's loss and every cotangent of the initial state
    against the oracle's reverse mode, which equals the dual-number forward
    mode at this horizon (tests/test_oracle_tangent.py). The two forward
    runs differ only in the deposits' summation order, so the loss agrees to
    1e-12 and each cotangent group to 1e-8 of its largest entry (the
    summation-order differences grow through the unstable run and the
    derivative is evaluated on the two slightly different trajectories)."""
    n, steps, K = 8192, 1000, 50
    print(f"adjoint over the full horizon vs the oracle: n={n} steps={steps} K={K}", flush=True)
    w = workloads.weibel_adjoint_1d2v(n, seed=2)
    nx, L, dt = w["nx"], w["xmax"], w["dt"]
    e1 = np.zeros(nx)
    orc.Splitting(nx, 0.0, L, 3, 3, 2, w["particles"].copy(order="F"), e1, np.zeros(nx),
                  w["b0"].copy(), common_weight=w["common_weight"]).solve_poisson(e1, np.zeros(nx))
    h, pg = _gpu(w["particles"], e1, np.zeros(nx), w["b0"], L, nx, (3, 2), w["common_weight"])
    r = G.strang_splitting_adjoint_(h, dt, steps, checkpoint_every=K)
    osp = orc.Splitting(nx, 0.0, L, 3, 3, 2, w["particles"].copy(order="F"), e1.copy(),
                        np.zeros(nx), w["b0"].copy(), common_weight=w["common_weight"])
    loss, gp, ge1, ge2, gb = osp.strang_adjoint(dt, steps, dt)
    assert abs(r.loss - loss) <= 1e-12 * abs(loss), (r.loss, loss)
    for row in range(3):
        _close(r.g_particles[row], gp[row], 1e-8)
    _close(r.g_e1, ge1, 1e-8)
    _close(r.g_e2, ge2, 1e-8)
    _close(r.g_b, gb, 1e-8)
