"""The reference's own driver loop (test/test_vm_1d2v.jl:59-117) with the
splitting on the GPU: the splitting is built and driven, and the
diagnostics are called, with the reference's signatures
(solve_poisson!(efield, particle_group, kernel_smoother_0, maxwell_solver,
rho), TimeHistoryDiagnostics(particle_group, maxwell_solver, ks0, ks1),
write_step!(thdiag, time, degree, efield_dofs, bfield_dofs, efield_dofs_n,
efield_poisson)); the diagnostics find the GPU copy of the particle group
through the splitting that holds it. Every write_step! row is compared with
the same loop on the oracle, within the permuted-oracle spread of
tests/test_gpu_parity_full.py (the loop's only difference from the oracle is
the deposits' summation order).
"""
import math

import numpy as np
import pytest

import gempic_amd as G
from gempic_amd import workloads
import pyoracle as orc

from test_gpu_parity_full import WRITE_STEP_COLUMNS, _trajectory_within_spread

pytestmark = pytest.mark.gpu


def _oracle_loop(parts, nx, xmin, xmax, b0, dt, steps, cw):
    e1, e2, b = np.zeros(nx), np.zeros(nx), b0.copy()
    rho, ep = np.zeros(nx), np.zeros(nx)
    osp = orc.Splitting(nx, xmin, xmax, 3, 3, 3, parts, e1, e2, b, common_weight=cw)
    osp.solve_poisson(e1, rho)
    rows = [osp.write_step(0.0, 3, e1, e2, b, e1, e2, ep)]
    for j in range(1, steps + 1):
        osp.strang_splitting(dt, 1)
        osp.solve_poisson(ep, rho)
        rows.append(osp.write_step(j * dt, 3, e1, e2, b, e1, e2, ep))
    return np.array(rows), e1


def test_gpu_reference_driver_loop():
    n, steps = 100000, 20
    w = workloads.weibel_1d2v(n, seed=6)
    nx, xmin, xmax, dt = w["nx"], w["xmin"], w["xmax"], w["dt"]
    print(f"reference driver loop: n={n} steps={steps}", flush=True)

    # --- test_vm_1d2v.jl:42-117, the splitting on the GPU ---
    mesh = G.OneDGrid(xmin, xmax, nx)
    particle_group = G.ParticleGroup(1, 2, n, common_weight=w["common_weight"])
    maxwell_solver = G.Maxwell1DFEM(mesh, 3)
    kernel_smoother1 = G.ParticleMeshCoupling1D(mesh, n, 3, "galerkin")
    kernel_smoother0 = G.ParticleMeshCoupling1D(mesh, n, 3, "galerkin")
    efield_dofs = [np.zeros(nx), np.zeros(nx)]
    bfield_dofs = np.zeros(nx)
    rho = np.zeros(nx)
    efield_poisson = np.zeros(nx)
    propagator = G.HamiltonianSplitting(1, 2, maxwell_solver, kernel_smoother0,
                                        kernel_smoother1, particle_group, efield_dofs,
                                        bfield_dofs)
    efield_dofs_n = propagator.e_dofs
    particle_group.array[:] = w["particles"]          # sample!(particle_group, ...)
    G.solve_poisson_(efield_dofs[0], particle_group, kernel_smoother0, maxwell_solver, rho)
    bfield_dofs[:] = w["b0"]                          # l2projection!(bfield_dofs, ...)
    thdiag = G.TimeHistoryDiagnostics(particle_group, maxwell_solver, kernel_smoother0,
                                      kernel_smoother1)
    G.write_step_(thdiag, 0.0, 3, efield_dofs, bfield_dofs, efield_dofs_n, efield_poisson)
    for j in range(1, steps + 1):
        G.strang_splitting_(propagator, dt, 1)
        G.solve_poisson_(efield_poisson, particle_group, kernel_smoother0, maxwell_solver, rho)
        G.write_step_(thdiag, j * dt, 3, efield_dofs, bfield_dofs, efield_dofs, efield_poisson)
    got = np.array(thdiag.data)

    ref, e1_ref = _oracle_loop(np.asfortranarray(w["particles"].copy()), nx, xmin, xmax,
                               w["b0"], dt, steps, w["common_weight"])
    perm = np.random.default_rng(2).permutation(n)
    alt, e1_alt = _oracle_loop(np.asfortranarray(w["particles"][:, perm].copy()), nx, xmin,
                               xmax, w["b0"], dt, steps, w["common_weight"])
    assert got.shape == (steps + 1, 11)
    np.testing.assert_array_equal(got[:, 0], ref[:, 0])
    for k in range(1, 10):
        _trajectory_within_spread(got[:, k], ref[:, k], alt[:, k], WRITE_STEP_COLUMNS[k])
    # ErrorPoisson subtracts two fields that each carry the deposits'
    # rounding (see tests/test_gpu_parity_full.py): its floor is 32 times the
    # permuted oracle's spread of E1 at the end, and at least 1e-11 of max |E1|
    _trajectory_within_spread(got[:, 10], ref[:, 10], alt[:, 10], "ErrorPoisson",
                              floor_abs=max(32.0 * float(np.max(np.abs(e1_alt - e1_ref))),
                                            1e-11 * float(np.max(np.abs(e1_ref)))))
    # the loop's particles came back to the caller's array in its order
    assert np.all(np.isfinite(particle_group.array))


def test_gpu_diagnostics_need_the_splittings_own_objects():
    w = workloads.weibel_1d2v(1000, seed=2)
    nx = w["nx"]
    mesh = G.OneDGrid(w["xmin"], w["xmax"], nx)
    pg = G.ParticleGroup(1, 2, 1000)
    pg.array[:] = w["particles"]
    ks0 = G.ParticleMeshCoupling1D(mesh, 1000, 3, "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, 1000, 2, "galerkin")
    mx = G.Maxwell1DFEM(mesh, 3)
    h = G.HamiltonianSplitting(1, 2, mx, ks0, ks1, pg, [np.zeros(nx), np.zeros(nx)], np.zeros(nx))
    other = G.ParticleMeshCoupling1D(mesh, 1000, 3, "galerkin")
    with pytest.raises(ValueError):
        G.solve_poisson_(np.zeros(nx), pg, other, mx, np.zeros(nx))
    with pytest.raises(ValueError):
        G.solve_poisson_(np.zeros(nx), G.ParticleGroup(1, 2, 1000), ks0, mx, np.zeros(nx))
    e = np.zeros(nx)
    G.solve_poisson_(e, pg, ks0, mx, np.zeros(nx))
    assert np.any(e != 0)
    del h
