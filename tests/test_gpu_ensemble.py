"""Ensembles of 1D2V splittings stepped together (one launch per pass for
all instances) through the C ABI: every instance's particles and fields
equal those of its own strang_splitting! call bit for bit, across sorts,
for instances of different particle counts, grids (16, 32, 64 cells) and
domains; one small instance is also checked against the oracle."""
import math

import numpy as np
import pytest

import gempic_amd as G
from gempic_amd import workloads
import pyoracle as orc

pytestmark = pytest.mark.gpu


def _instance(n, nx, k, seed, sync=True):
    w = workloads.weibel_1d2v(n, seed=seed, nx=nx, k=k, alpha=1e-2, beta=1e-2,
                              sigma1=0.1, sigma2=0.2)
    mesh = G.OneDGrid(w["xmin"], w["xmax"], nx)
    pg = G.ParticleGroup(1, 2, n, common_weight=w["common_weight"])
    pg.array[:] = w["particles"]
    ks0 = G.ParticleMeshCoupling1D(mesh, n, 3, "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, n, 2, "galerkin")
    rng = np.random.default_rng(seed)
    e1, e2 = 0.01 * rng.normal(size=nx), 0.01 * rng.normal(size=nx)
    h = G.HamiltonianSplitting(1, 2, G.Maxwell1DFEM(mesh, 3), ks0, ks1, pg, [e1, e2],
                               w["b0"].copy(), sync=sync)
    return h, pg, w


SPECS = [(20000, 16, 1.0, 1), (65536, 32, 1.25, 2), (5000, 64, 1.5, 3), (131072, 32, 1.1, 4),
         (777, 16, 1.3, 5)]


def _dt(nx):
    # below the degree-3 FEM Maxwell solver's stability limit at every grid
    return 0.05 if nx <= 32 else 0.02


def test_gpu_ensemble_equals_standalone_runs():
    steps = 25                                # the default sort interval (20) falls inside
    alone = []
    for (n, nx, k, seed) in SPECS:
        h, pg, w = _instance(n, nx, k, seed)
        G.strang_splitting_(h, _dt(nx), steps)
        alone.append((pg.array.copy(), h.e_dofs[0].copy(), h.e_dofs[1].copy(), h.b_dofs.copy()))
    hs = [_instance(n, nx, k, seed) for (n, nx, k, seed) in SPECS]
    ens = G.Ensemble([h for h, _, _ in hs])
    G.ensemble_strang_splitting_(ens, [_dt(nx) for (_, nx, _, _) in SPECS], steps)
    for (h, pg, _), ref in zip(hs, alone):
        np.testing.assert_array_equal(pg.array, ref[0])
        np.testing.assert_array_equal(h.e_dofs[0], ref[1])
        np.testing.assert_array_equal(h.e_dofs[1], ref[2])
        np.testing.assert_array_equal(h.b_dofs, ref[3])


def test_gpu_ensemble_instance_matches_oracle_and_repeated_calls():
    h, pg, w = _instance(3000, 16, 1.25, 7)
    h2, pg2, _ = _instance(50000, 64, 1.0, 8)
    parts = w["particles"].copy(order="F")
    e1, e2, b = h.e_dofs[0].copy(), h.e_dofs[1].copy(), h.b_dofs.copy()
    ens = G.Ensemble([h, h2])
    for steps in (3, 4):
        G.ensemble_strang_splitting_(ens, [0.05, 0.02], steps)
    osp = orc.Splitting(16, w["xmin"], w["xmax"], 3, 3, 2, parts, e1, e2, b,
                        common_weight=w["common_weight"])
    osp.strang_splitting(0.05, 7)
    np.testing.assert_allclose(pg.array, parts, rtol=0, atol=1e-12)
    for got, ref in ((h.e_dofs[0], e1), (h.e_dofs[1], e2), (h.b_dofs, b)):
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(np.max(np.abs(ref)), 1e-300)


def test_gpu_ensemble_rejects_mixed_handles():
    h, _, _ = _instance(1000, 16, 1.25, 1)
    mesh = G.OneDGrid(0.0, 1.0, 16)
    pg = G.ParticleGroup(1, 2, 100)
    pg.array[0] = np.linspace(0.0, 0.9, 100)
    ks0 = G.ParticleMeshCoupling1D(mesh, 100, 2, "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, 100, 1, "galerkin")
    other = G.HamiltonianSplitting(1, 2, G.Maxwell1DFEM(mesh, 2), ks0, ks1, pg,
                                   [np.zeros(16), np.zeros(16)], np.zeros(16))
    with pytest.raises(Exception):
        G.Ensemble([h, other])
    with pytest.raises(Exception):
        G.Ensemble([h, h])


@pytest.mark.parametrize("K,sort_every", [(3, 2), (1, 3)])
def test_gpu_ensemble_adjoint_equals_standalone(K, sort_every):
    """The batched reverse mode of an ensemble (instances of different N, nx,
    k and dt) gives every instance the loss, cotangents and domain
    derivatives of its own strang_splitting_adjoint_, bit for bit, across
    particle sorts."""
    specs = [(20000, 16, 1.0, 1), (8192, 64, 1.5, 3), (65536, 32, 1.25, 2)]
    steps = 8
    alone = []
    for (n, nx, k, seed) in specs:
        h, pg, w = _instance(n, nx, k, seed)
        h.set_sort_interval(sort_every)
        r = G.strang_splitting_adjoint_(h, _dt(nx), steps, checkpoint_every=K)
        alone.append((r, pg.array.copy()))
    hs = []
    for (n, nx, k, seed) in specs:
        h, pg, w = _instance(n, nx, k, seed)
        h.set_sort_interval(sort_every)
        hs.append((h, pg))
    ens = G.Ensemble([h for h, _ in hs])
    res = G.ensemble_strang_splitting_adjoint_(ens, [_dt(nx) for (_, nx, _, _) in specs], steps,
                                               checkpoint_every=K)
    for (h, pg), r, (ra, parts) in zip(hs, res, alone):
        np.testing.assert_array_equal(pg.array, parts)
        assert r.loss == ra.loss
        for a in ("g_particles", "g_e1", "g_e2", "g_b"):
            np.testing.assert_array_equal(getattr(r, a), getattr(ra, a))
        assert r.g_length == ra.g_length and r.g_weight_sum == ra.g_weight_sum


def test_gpu_ensemble_graph_replay_equals_stream_steps():
    """GEMPIC_GRAPH: the regular steps replayed as a captured HIP graph give
    the stream-queued steps' particles and fields bit for bit, call after
    call (the second and third calls replay the graph captured in the
    first, with the host's own work and synchronisation between them and a
    particle sort between replays)."""
    specs = SPECS[:3]
    dts = [_dt(nx) for (_, nx, _, _) in specs]
    eager = [_instance(n, nx, k, seed) for (n, nx, k, seed) in specs]
    graph = [_instance(n, nx, k, seed) for (n, nx, k, seed) in specs]
    ens_e = G.Ensemble([h for h, _, _ in eager])
    ens_g = G.Ensemble([h for h, _, _ in graph])
    for steps in (6, 17, 9):                 # 32 steps: the sort at 20 falls between replays
        G.ensemble_strang_splitting_(ens_e, dts, steps)
        G.ensemble_strang_splitting_(ens_g, dts, steps, graph=True)
        for (he, pe, _), (hg, pgr, _) in zip(eager, graph):
            np.testing.assert_array_equal(pgr.array, pe.array)
            np.testing.assert_array_equal(hg.e_dofs[0], he.e_dofs[0])
            np.testing.assert_array_equal(hg.e_dofs[1], he.e_dofs[1])
            np.testing.assert_array_equal(hg.b_dofs, he.b_dofs)


def test_gpu_ensemble_autotune_restores_state_and_sets_the_choice():
    """gempic_ensemble_autotune: every trial starts from the same state and
    the state after tuning is the one before it (bitwise); the choice is one
    of the candidates."""
    specs = SPECS[:3]
    dts = [_dt(nx) for (_, nx, _, _) in specs]
    hs = [_instance(n, nx, k, seed) for (n, nx, k, seed) in specs]
    ens = G.Ensemble([h for h, _, _ in hs])
    G.ensemble_strang_splitting_(ens, dts, 3)
    before = [(pg.array.copy(), h.e_dofs[0].copy(), h.e_dofs[1].copy(), h.b_dofs.copy())
              for h, pg, _ in hs]
    cands = (1024, 2048, 4096)
    r = G.ensemble_autotune(ens, dts, n_steps=4, particles_per_workgroup=cands)
    assert r["particles_per_workgroup"] in cands
    assert len(r["ms_per_step"]) == 3 and all(v > 0 for v in r["ms_per_step"])
    for (h, pg, _), ref in zip(hs, before):
        h.to_host()
        np.testing.assert_array_equal(pg.array, ref[0])
        np.testing.assert_array_equal(h.e_dofs[0], ref[1])
        np.testing.assert_array_equal(h.e_dofs[1], ref[2])
        np.testing.assert_array_equal(h.b_dofs, ref[3])
    # the ensemble runs on with the chosen launch
    G.ensemble_strang_splitting_(ens, dts, 5)
    for h, pg, _ in hs:
        assert np.all(np.isfinite(pg.array)) and np.all(np.isfinite(h.b_dofs))


def test_gpu_ensemble_in_fused_arithmetic_equals_standalone_runs():
    """Splittings set to the fused arithmetic step as a batch in it: every
    instance equals its own fused strang_splitting! bit for bit across a
    sort; splittings of mixed arithmetics are rejected."""
    specs = SPECS[:3]
    dts = [_dt(nx) for (_, nx, _, _) in specs]
    steps = 25
    alone = []
    for (n, nx, k, seed) in specs:
        h, pg, w = _instance(n, nx, k, seed)
        h.set_arithmetic("fused")
        G.strang_splitting_(h, _dt(nx), steps)
        alone.append((pg.array.copy(), h.e_dofs[0].copy(), h.e_dofs[1].copy(), h.b_dofs.copy()))
    hs = [_instance(n, nx, k, seed) for (n, nx, k, seed) in specs]
    for h, _, _ in hs:
        h.set_arithmetic("fused")
    ens = G.Ensemble([h for h, _, _ in hs])
    G.ensemble_strang_splitting_(ens, dts, steps)
    for (h, pg, _), ref in zip(hs, alone):
        np.testing.assert_array_equal(pg.array, ref[0])
        np.testing.assert_array_equal(h.e_dofs[0], ref[1])
        np.testing.assert_array_equal(h.e_dofs[1], ref[2])
        np.testing.assert_array_equal(h.b_dofs, ref[3])
    hs[1][0].set_arithmetic("reference")
    with pytest.raises(ValueError):
        G.ensemble_strang_splitting_(ens, dts, 2)
