"""The oracle's sample!(pg, alpha, k, sigma, mesh) (particle_sampling.jl:236-311)
and the Sobol.jl SobolSeq(2) points it draws (Sobol.jl is an external
dependency, compat "1" in Project.toml, not vendored in the reference).

Pinned by: the published first points of the two-dimensional Sobol
sequence (Bratley & Fox 1988, Gray-code order, zero point skipped), the
(0, m, 2)-net property of every aligned block of 2^m points, newton's
defining equation (the cdf of 1 + alpha cos(k x)), and the moments of the
sampled density and speeds."""
import math

import numpy as np
import pytest

import pyoracle as orc


def test_sobol_first_points():
    want = [(0.5, 0.5), (0.75, 0.25), (0.25, 0.75), (0.375, 0.375), (0.875, 0.875),
            (0.625, 0.125), (0.125, 0.625), (0.1875, 0.3125), (0.6875, 0.8125),
            (0.9375, 0.0625), (0.4375, 0.5625)]
    for k, w in enumerate(want, start=1):
        assert orc.sobol2(k) == w, (k, orc.sobol2(k), w)


@pytest.mark.parametrize("m", [6, 10])
def test_sobol_blocks_are_nets(m):
    """Points 2^m .. 2^(m+1)-1 put exactly one point in every elementary box
    2^-a x 2^-b with a + b = m."""
    pts = np.array([orc.sobol2(k) for k in range(2 ** m, 2 ** (m + 1))])
    for a in range(m + 1):
        b = m - a
        ix = np.floor(pts[:, 0] * 2 ** a).astype(int)
        iy = np.floor(pts[:, 1] * 2 ** b).astype(int)
        cells = ix * 2 ** b + iy
        assert np.array_equal(np.sort(cells), np.arange(2 ** m)), (a, b)


@pytest.mark.parametrize("alpha,k", [(0.5, 0.5), (0.01, 0.5), (1e-4, 1.25)])
def test_newton_inverts_the_cdf(alpha, k):
    L = 2 * math.pi / k
    for r in np.linspace(0.0, 1.0, 41):
        x, it = orc.landau_newton(float(r), alpha, k)
        assert it < 60
        cdf = x + alpha * math.sin(k * x) / k
        assert abs(cdf - r * 2 * math.pi / k) < 1e-10 * L
        assert -1e-12 <= x <= L * (1 + 1e-12)


@pytest.mark.parametrize("dim_v", [1, 2])
def test_sample_landau_moments(dim_v):
    n, alpha, k, sigma = 100000, 0.5, 0.5, 1.0
    L = 2 * math.pi / k
    a, it = orc.sample_landau(dim_v, n, alpha, k, sigma, L)
    assert it < 60
    x = a[0]
    np.testing.assert_array_equal(a[1 + dim_v], L)
    assert x.min() >= 0.0 and x.max() <= L
    assert abs(x.mean() - L / 2) < 2e-3 * L
    # E[cos(k x)] = alpha / 2 for the density (1 + alpha cos(k x)) / L
    assert abs(np.cos(k * x).mean() - alpha / 2) < 2e-3
    i = np.arange(1, n + 1)
    speed = sigma * np.sqrt(-2.0 * np.log((i - 0.5) / n))
    if dim_v == 2:
        np.testing.assert_allclose(np.hypot(a[1], a[2]), speed, rtol=1e-14, atol=1e-15)
        assert abs((a[1] ** 2).mean() - sigma ** 2) < 2e-2
        assert abs((a[2] ** 2).mean() - sigma ** 2) < 2e-2
    else:
        th = np.array([orc.sobol2(int(j))[0] for j in i[:1000]]) * 2 * math.pi
        np.testing.assert_allclose(a[1, :1000], speed[:1000] * np.cos(th), rtol=1e-14, atol=1e-15)
