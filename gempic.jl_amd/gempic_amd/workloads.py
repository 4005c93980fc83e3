"""Seeded synthetic initial conditions of the benchmark workloads.

There is no data to load: particles are drawn from a seeded numpy
generator, so the same arrays feed the GPU path and any checker.

* weibel_1d2v: 1D2V Weibel instability, x in [0, 2pi/k), density
  1 + alpha*cos(kx) (rejection sampling), v1 ~ N(0, s1^2), v2 ~ N(0, s2^2),
  weight = Lx (the reference sampler's `dimx`, particle_sampling.jl:279),
  common_weight = 1/N, B0 = l2projection!(beta*cos(kx), p-1).
* landau_1d1v: 1D1V strong Landau damping, x in [0, 2pi/k), density
  1 + alpha*cos(kx), v ~ N(0, 1).
"""
from __future__ import annotations

import math

import numpy as np

from .core import _bspline_basis


def sample_cos_density(rng, n, L, k, alpha):
    """Rejection sampling of 1 + alpha*cos(k x) on [0, L)."""
    out = np.empty(n)
    filled = 0
    bound = 1.0 + abs(alpha)
    while filled < n:
        m = int((n - filled) * 1.1) + 16
        u = rng.random(m) * L
        acc = rng.random(m) * bound < (1.0 + alpha * np.cos(k * u))
        take = u[acc][: n - filled]
        out[filled: filled + take.size] = take
        filled += take.size
    return out


def _gauss_legendre(n):
    """FastGaussQuadrature.gausslegendre(n) for n <= 4, in closed form (the
    rule compute_rhs_from_function! takes, maxwell_1d_fem.jl:195)."""
    if n == 1:
        return [0.0], [2.0]
    if n == 2:
        return [-1.0 / math.sqrt(3.0), 1.0 / math.sqrt(3.0)], [1.0, 1.0]
    if n == 3:
        r = math.sqrt(3.0 / 5.0)
        return [-r, 0.0, r], [5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0]
    a = 2.0 / 7.0 * math.sqrt(6.0 / 5.0)
    x = [-math.sqrt(3.0 / 7.0 + a), -math.sqrt(3.0 / 7.0 - a), math.sqrt(3.0 / 7.0 - a),
         math.sqrt(3.0 / 7.0 + a)]
    w0, w1 = (18.0 - math.sqrt(30.0)) / 36.0, (18.0 + math.sqrt(30.0)) / 36.0
    return x, [w0, w1, w1, w0]


def l2projection_cos(nx, L, degree, amp, k):
    """l2projection!(amp*cos(k x), degree) of maxwell_1d_fem.jl:347-374, which
    returns its right-hand side divided by dx (the mass solve lands in the
    workspace only)."""
    dx = L / nx
    xg, wg = _gauss_legendre(degree + 1)
    xg = [0.5 * (x + 1.0) for x in xg]
    wg = [0.5 * w for w in wg]
    bspl = [_bspline_basis(degree, xk) for xk in xg]
    out = np.zeros(nx)
    for i in range(1, nx + 1):
        coef = 0.0
        for j in range(1, degree + 2):
            for kq in range(1, degree + 2):
                xq = ((xg[kq - 1] + i) + j) - 2.0
                coef = coef + wg[kq - 1] * (amp * math.cos(k * (dx * xq))) * bspl[kq - 1][degree + 1 - j]
        out[i - 1] = coef * dx
    return out / dx


def weibel_1d2v(n, seed=1, nx=32, k=1.25, alpha=1e-4, beta=1e-4, sigma1=0.02 / math.sqrt(2.0),
                sigma2=None, degree=3):
    if sigma2 is None:
        sigma2 = math.sqrt(12.0) * sigma1
    L = 2.0 * math.pi / k
    rng = np.random.default_rng(seed)
    parts = np.empty((4, n), order="F")
    parts[0] = sample_cos_density(rng, n, L, k, alpha)
    parts[1] = rng.standard_normal(n) * sigma1
    parts[2] = rng.standard_normal(n) * sigma2
    parts[3] = L
    b0 = l2projection_cos(nx, L, degree - 1, beta, k)
    return dict(xmin=0.0, xmax=L, nx=nx, degree=degree, particles=parts, b0=b0,
                common_weight=1.0 / n, dt=0.05)


def landau_1d1v(n, seed=1, nx=32, k=0.5, alpha=0.5, degree=3):
    L = 2.0 * math.pi / k
    rng = np.random.default_rng(seed)
    parts = np.empty((3, n), order="F")
    parts[0] = sample_cos_density(rng, n, L, k, alpha)
    parts[1] = rng.standard_normal(n)
    parts[2] = L
    return dict(xmin=0.0, xmax=L, nx=nx, degree=degree, particles=parts,
                common_weight=1.0 / n, dt=0.05)


def newton_cos_cdf(u, L, k, alpha, iters=60):
    """x in [0, L) with x + (alpha/k) sin(k x) = u L, the inverse cdf of the
    density (1 + alpha cos(k x)) / L, by newton's method from x = u L
    (vectorised; the same equation as particle_sampling.jl:236-245)."""
    x = u * L
    for _ in range(iters):
        f = x + alpha * np.sin(k * x) / k - u * L
        x = x - f / (1.0 + alpha * np.cos(k * x))
    return x


def weibel_adjoint_1d2v(n, seed=1, nx=32, k=1.25, alpha=1e-4, beta=1e-4,
                        sigma1=0.02 / math.sqrt(2.0), sigma2=None, degree=3, dt=0.05):
    """The 1D2V Weibel initial condition with positions that are a smooth
    function of alpha, for gradients: x_i = inverse cdf of (i - 1/2)/n
    (a stratified quiet start), v1 ~ N(0, s1^2), v2 ~ N(0, s2^2) (seeded),
    weight = Lx, B0 = beta * P with P = l2projection!(cos(kx), p-1).
    Returns also dx/dalpha per particle and P, so that with the cotangents
    of the initial state, dJ/dbeta = <g_B, P> and
    dJ/dalpha = sum_i (g_x[i] + g_pois[i]) dx_i/dalpha, where g_pois is the
    solve_poisson! pullback of g_E1 (E1 is set by Poisson from x)."""
    if sigma2 is None:
        sigma2 = math.sqrt(12.0) * sigma1
    L = 2.0 * math.pi / k
    rng = np.random.default_rng(seed)
    u = (np.arange(n) + 0.5) / n
    x = newton_cos_cdf(u, L, k, alpha)
    dx_dalpha = -(np.sin(k * x) / k) / (1.0 + alpha * np.cos(k * x))
    v1 = sigma1 * rng.standard_normal(n)
    v2 = sigma2 * rng.standard_normal(n)
    parts = np.asfortranarray(np.vstack([x, v1, v2, np.full(n, L)]))
    p = l2projection_cos(nx, L, degree - 1, 1.0, k)
    return dict(xmin=0.0, xmax=L, nx=nx, degree=degree, particles=parts, b_pattern=p,
                b0=beta * p, dx_dalpha=dx_dalpha, common_weight=1.0 / n, dt=dt, k=k,
                alpha=alpha, beta=beta)


def landau_2d3v(n, seed=1, nx=64, ny=64, k=0.5, alpha=0.01, degree=3):
    """2D3V Landau damping: (x1, x2) in [0, 2pi/k)^2 with density
    1 + alpha cos(k x1) cos(k x2) (rejection sampling), v ~ N(0, I_3),
    weight = the domain's area (the reference sampler's volume weight,
    particle_sampling.jl:279 in 1D), common_weight = 1/N. Rows x1, x2, v1,
    v2, v3, w (ParticleGroup{2,3})."""
    L = 2.0 * math.pi / k
    rng = np.random.default_rng(seed)
    x1 = np.empty(n)
    x2 = np.empty(n)
    filled = 0
    bound = 1.0 + abs(alpha)
    while filled < n:
        m = int((n - filled) * 1.05) + 16
        u, v = rng.random(m) * L, rng.random(m) * L
        acc = rng.random(m) * bound < 1.0 + alpha * np.cos(k * u) * np.cos(k * v)
        take = min(int(acc.sum()), n - filled)
        x1[filled: filled + take] = u[acc][:take]
        x2[filled: filled + take] = v[acc][:take]
        filled += take
    parts = np.empty((6, n), order="F")
    parts[0], parts[1] = x1, x2
    parts[2:5] = rng.standard_normal((3, n))
    parts[5] = L * L
    return dict(xmin=0.0, xmax=L, ymin=0.0, ymax=L, nx=nx, ny=ny, degree=degree,
                particles=parts, common_weight=1.0 / n, dt=0.05)
