"""GPU sample!(pg, alpha, k, sigma, mesh) (particle_sampling.jl:266-311)
against the oracle: the Sobol points are exact dyadic rationals on both
sides, so positions and velocities differ only by the last-bit differences
of the device's and the host's sin, cos and log (compared at 1e-13)."""
import math

import numpy as np
import pytest

import gempic_amd as G
from gempic_amd import workloads
import pyoracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim_v", [1, 2])
def test_gpu_sample_landau_matches_oracle(dim_v):
    n, alpha, k, sigma = 100003, 0.5, 0.5, 1.0
    L = 2 * math.pi / k
    mesh = G.OneDGrid(0.0, L, 32)
    pg = G.ParticleGroup(1, dim_v, n)
    G.sample_(pg, alpha, k, sigma, mesh)
    ref, _ = orc.sample_landau(dim_v, n, alpha, k, sigma, L)
    for r in range(ref.shape[0]):
        scale = max(np.abs(ref[r]).max(), 1e-300)
        assert np.abs(pg.array[r] - ref[r]).max() <= 1e-13 * scale, r


def test_gpu_sample_into_handle_and_run():
    n, alpha, k, sigma = 1 << 20, 0.5, 0.5, 1.0
    L = 2 * math.pi / k
    mesh = G.OneDGrid(0.0, L, 32)
    pg = G.ParticleGroup(1, 2, n)
    ks0 = G.ParticleMeshCoupling1D(mesh, n, 3, "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, n, 2, "galerkin")
    mx = G.Maxwell1DFEM(mesh, 3)
    e1, e2, b = np.zeros(32), np.zeros(32), np.zeros(32)
    h = G.HamiltonianSplitting(1, 2, mx, ks0, ks1, pg, [e1, e2], b)
    G.sample_particles_(h, alpha, k, sigma)
    host = G.ParticleGroup(1, 2, n)
    G.sample_(host, alpha, k, sigma, mesh)
    np.testing.assert_array_equal(pg.array, host.array)      # same kernel, caller's order
    rho = np.zeros(32)
    G.solve_poisson_(e1, h, rho)
    assert abs(rho.sum() - 1.0 * L) < 1e-10 * L                # total charge = dimx
    h.to_device()
    G.strang_splitting_(h, 0.05, 5)
    assert np.isfinite(pg.array).all()


def test_gpu_sample_argument_checks():
    pg = G.ParticleGroup(1, 2, 100)
    with pytest.raises(AssertionError):                        # mesh.dimx ≈ 2π / k (:297)
        G.sample_(pg, 0.5, 0.5, 1.0, G.OneDGrid(0.0, 10.0, 16))
    with pytest.raises(ValueError):
        G.sample_(pg, 1.5, 0.5, 1.0, G.OneDGrid(0.0, 4 * math.pi, 16))
