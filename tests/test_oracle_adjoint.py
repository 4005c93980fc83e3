"""The oracle's reverse mode of the 1D2V Strang splitting (no counterpart in
the reference) against central finite differences of the oracle's forward
functions, which the reference's known answers pin.

Loss: wloss * sum over steps of 1/2 l2norm_squared(B_s, p-1)
(maxwell_1d_fem.jl:299-313, diagnostics.jl:236) plus <ct, final state>.
Every input group (positions, both velocities, E1, E2, B) is perturbed along a
random direction; the directional derivative from the adjoint must match
(J(+h) - J(-h)) / 2h to 1e-6 relative, or to 1e-9 of the loss's size for
directions the loss hardly depends on (h = 1e-5: truncation ~1e-10 relative,
rounding ~1e-11 of the loss)."""
import math

import numpy as np
import pytest

import pyoracle as orc

H = 1e-5
TOL = 1e-6


def _state(n=400, nx=16, seed=3, vscale=0.6):
    rng = np.random.default_rng(seed)
    L = 4 * math.pi
    parts = np.asfortranarray(np.vstack([L * rng.random(n), vscale * rng.normal(size=n),
                                         0.5 * rng.normal(size=n), rng.random(n) + 0.5]))
    e1, e2, b = (0.3 * rng.normal(size=nx) for _ in range(3))
    return parts, e1, e2, b, L, nx


def _split(st, deg, parts=None, e1=None, e2=None, b=None):
    p0, a1, a2, bb, L, nx = st
    p = (p0 if parts is None else parts).copy(order="F")
    f = [(x if y is None else y).copy() for x, y in ((a1, e1), (a2, e2), (bb, b))]
    sp = orc.Splitting(nx, 0.0, L, 3, deg[0], deg[1], p, f[0], f[1], f[2], common_weight=0.01)
    return sp, p, f


def _loss(st, deg, dt, steps, wl, ct, d, s):
    parts = np.asfortranarray(st[0] + s * H * d[0])
    sp, p, f = _split(st, deg, parts, st[1] + s * H * d[1], st[2] + s * H * d[2],
                      st[3] + s * H * d[3])
    J = sp.strang_loss(dt, steps, wl)
    if ct is not None:
        J += np.sum(ct[0][:3] * p[:3]) + ct[1] @ f[0] + ct[2] @ f[1] + ct[3] @ f[2]
    return J


@pytest.mark.parametrize("deg", [(3, 2), (2, 1)])
@pytest.mark.parametrize("with_ct", [False, True])
@pytest.mark.parametrize("steps", [1, 6])
def test_adjoint_matches_central_differences(deg, with_ct, steps):
    st = _state()
    n, nx = st[0].shape[1], st[5]
    dt, wl = 0.1, 0.1
    rng = np.random.default_rng(11)
    ct = None
    if with_ct:
        cp = np.zeros_like(st[0], order="F")
        cp[:3] = rng.normal(size=(3, n))
        ct = (cp, rng.normal(size=nx), rng.normal(size=nx), rng.normal(size=nx))
    sp, _, _ = _split(st, deg)
    args = () if ct is None else ct
    J, gp, ge1, ge2, gb = sp.strang_adjoint(dt, steps, wl, *args)
    if ct is None:
        assert J == _split(st, deg)[0].strang_loss(dt, steps, wl)
    np.testing.assert_array_equal(gp[3], 0.0)
    scale = abs(_loss(st, deg, dt, steps, wl, ct, [0.0] * 4, 0))
    for group in range(6):
        d = [np.zeros_like(st[0]), np.zeros(nx), np.zeros(nx), np.zeros(nx)]
        if group < 3:
            d[0][group] = rng.normal(size=n)
        else:
            d[group - 2] = rng.normal(size=nx)
        fd = (_loss(st, deg, dt, steps, wl, ct, d, 1) - _loss(st, deg, dt, steps, wl, ct, d, -1)) / (2 * H)
        ad = np.sum(gp * d[0]) + ge1 @ d[1] + ge2 @ d[2] + gb @ d[3]
        assert abs(fd - ad) <= TOL * max(abs(fd), abs(ad)) + 1e-9 * scale, (group, fd, ad)


def test_adjoint_particles_cross_cells_and_wrap():
    """Fast particles: every Hp1 crosses cells and some leave through both
    ends of the domain (the trunc() cells and the mod)."""
    st = _state(n=300, seed=5, vscale=4.0)
    L = st[4]
    dt, steps, wl = 0.2, 4, 0.2
    x0 = st[0][0]
    x1 = x0 + dt * st[0][1]
    assert np.any(x1 < 0) and np.any(x1 >= L)
    sp, _, _ = _split(st, (3, 2))
    J, gp, ge1, ge2, gb = sp.strang_adjoint(dt, steps, wl)
    rng = np.random.default_rng(2)
    for row in range(3):
        d = [np.zeros_like(st[0]), np.zeros(16), np.zeros(16), np.zeros(16)]
        d[0][row] = rng.normal(size=st[0].shape[1])
        fd = (_loss(st, (3, 2), dt, steps, wl, None, d, 1) -
              _loss(st, (3, 2), dt, steps, wl, None, d, -1)) / (2 * H)
        ad = np.sum(gp * d[0])
        assert abs(fd - ad) <= TOL * max(abs(fd), abs(ad)), (row, fd, ad)


def test_bspline_derivative():
    for p in range(0, 5):
        for t in (0.0, 0.3, 0.77, -0.2):
            d = orc.bspline_deriv(p, t)
            if p == 0:
                np.testing.assert_array_equal(d, 0.0)
                continue
            h = 1e-6
            fd = (orc.bspline_basis(p, t + h) - orc.bspline_basis(p, t - h)) / (2 * h)
            np.testing.assert_allclose(d, fd, rtol=0, atol=1e-8)
            assert abs(d.sum()) < 1e-15        # d/dt of a partition of unity


def test_solve_poisson_adjoint():
    st = _state(n=500, seed=7)
    nx = st[5]
    rng = np.random.default_rng(3)
    ct = rng.normal(size=nx)
    sp, p, _ = _split(st, (3, 2))
    g = sp.solve_poisson_adjoint(ct)
    d = rng.normal(size=p.shape[1])

    def f(s):
        sp2, p2, _ = _split(st, (3, 2))
        p2[0] += s * H * d
        e, rho = np.zeros(nx), np.zeros(nx)
        sp2.solve_poisson(e, rho)
        return ct @ e
    fd = (f(1) - f(-1)) / (2 * H)
    assert abs(fd - g @ d) <= TOL * abs(fd)


def _k_loss(k, n, steps, with_grad):
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "gempic.jl_amd"))
    from gempic_amd import workloads
    w = workloads.weibel_adjoint_1d2v(n, seed=2, nx=16, k=k, alpha=1e-2, beta=1e-2,
                                      sigma1=0.2)
    nx = w["nx"]
    parts = np.asfortranarray(w["particles"].copy())
    e1, e2, b = np.zeros(nx), np.zeros(nx), w["b0"].copy()
    sp = orc.Splitting(nx, 0.0, w["xmax"], 3, 3, 2, parts, e1, e2, b,
                       common_weight=w["common_weight"])
    sp.solve_poisson(e1, np.zeros(nx))
    e1_0 = e1.copy()
    if not with_grad:
        return sp.strang_loss(0.1, steps, 0.1)
    J, gp, ge1, ge2, gb = sp.strang_adjoint(0.1, steps, 0.1)
    g_len, g_w = sp.strang_adjoint_domain()
    L = w["xmax"]
    dj_dl = g_len + g_w + float(gp[0] @ w["particles"][0]) / L + float(ge1 @ e1_0) / L
    return J, -(L / k) * dj_dl


def test_adjoint_wavenumber_gradient():
    """dJ/dk, where k moves Lx = 2 pi / k and with it dx, the quiet-start
    positions (x = Lx xi), the weights (w = Lx) and E1 (Poisson, ~ Lx): the
    oracle's explicit domain derivatives plus the chain rule through the
    initial state, against central differences in k."""
    k, n, steps = 1.25, 600, 6
    J, dj_dk = _k_loss(k, n, steps, True)
    hk = 1e-6 * k
    fd = (_k_loss(k + hk, n, steps, False) - _k_loss(k - hk, n, steps, False)) / (2 * hk)
    assert abs(fd - dj_dk) <= 1e-6 * abs(fd), (fd, dj_dk)
