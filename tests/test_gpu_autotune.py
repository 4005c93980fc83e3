"""Launch autotuner (gempic_autotune) through the C ABI.

The tuner runs trial Strang steps on the splitting's own state. It must
leave particles and fields bitwise as it found them, keep the fastest
(blocks, sort interval) pair, and the splitting must then run bitwise like
a fresh one set to that pair by hand. A different pair changes only the
deposit summation order: the final states agree to 1e-12 of their scale.
"""
import math

import numpy as np
import pytest

import gempic_amd as G
from gempic_amd import core

pytestmark = pytest.mark.gpu


def _state(n=65536, nx=32, seed=5):
    rng = np.random.default_rng(seed)
    L = 4 * math.pi
    parts = np.asfortranarray(np.vstack([L * rng.random(n), 0.5 * rng.normal(size=n),
                                         0.5 * rng.normal(size=n), rng.random(n) + 0.5]))
    e1, e2, b = (0.1 * rng.normal(size=nx) for _ in range(3))
    return parts, e1, e2, b, L, nx


def _make(parts, e1, e2, b, L, nx):
    n = parts.shape[1]
    mesh = G.OneDGrid(0.0, L, nx)
    pg = G.ParticleGroup(1, 2, n, common_weight=1.0 / n)
    pg.array[:] = parts
    ks0 = G.ParticleMeshCoupling1D(mesh, n, 3, "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, n, 2, "galerkin")
    h = G.HamiltonianSplitting(1, 2, G.Maxwell1DFEM(mesh, 3), ks0, ks1, pg,
                               [e1.copy(), e2.copy()], b.copy(), sync=False)
    return h, pg


def _snapshot(h, pg):
    h.to_host()
    return (pg.array.copy(), h.e_dofs[0].copy(), h.e_dofs[1].copy(), h.b_dofs.copy())


def test_autotune_restores_state_and_runs_as_the_chosen_pair():
    st = _state()
    dt, steps = 0.05, 12
    blocks, intervals = [0, 64, 256], [3, 7]
    h, pg = _make(*st)
    before = _snapshot(h, pg)
    res = h.autotune(dt, n_steps=8, blocks=blocks, sort_intervals=intervals)
    after = _snapshot(h, pg)
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    assert res["blocks"] in blocks and res["sort_interval"] in intervals
    ms = np.array(res["ms_per_step"])
    assert ms.shape == (3, 2) and np.all(ms > 0)
    i, j = np.unravel_index(np.argmin(ms), ms.shape)
    assert (blocks[i], intervals[j]) == (res["blocks"], res["sort_interval"])

    G.strang_splitting_(h, dt, steps)
    tuned = _snapshot(h, pg)

    h2, pg2 = _make(*st)
    h2.set_launch(res["blocks"])
    h2.set_sort_interval(res["sort_interval"])
    G.strang_splitting_(h2, dt, steps)
    by_hand = _snapshot(h2, pg2)
    for a, b in zip(tuned, by_hand):
        assert np.array_equal(a, b)

    # any other pair: the same trajectory up to the deposit summation order
    h3, pg3 = _make(*st)
    h3.set_launch(256 if res["blocks"] != 256 else 64)
    h3.set_sort_interval(7 if res["sort_interval"] != 7 else 3)
    G.strang_splitting_(h3, dt, steps)
    other = _snapshot(h3, pg3)
    for a, b in zip(tuned, other):
        assert np.max(np.abs(a - b)) <= 1e-12 * max(np.max(np.abs(b)), 1e-300)


def test_autotune_rejects_bad_candidates_and_keeps_settings():
    st = _state(n=4096)
    h, pg = _make(*st)
    before = _snapshot(h, pg)
    with pytest.raises(ValueError):
        h.autotune(0.05, n_steps=4, blocks=[0, 10 ** 7], sort_intervals=[5])
    with pytest.raises(ValueError):
        h.autotune(0.05, n_steps=0, blocks=[0], sort_intervals=[5])
    with pytest.raises(ValueError):
        h.autotune(0.05, n_steps=4, blocks=[0], sort_intervals=[-1])
    for a, b in zip(before, _snapshot(h, pg)):
        assert np.array_equal(a, b)


def test_autotune_cache_file(tmp_path):
    st = _state(n=8192)
    path = str(tmp_path / "tune.json")
    h, pg = _make(*st)
    r1 = h.autotune(0.05, n_steps=6, blocks=[0, 32], sort_intervals=[3], cache=path)
    assert r1["cached"] is False
    h2, _ = _make(*st)
    r2 = h2.autotune(0.05, n_steps=6, blocks=[0, 32], sort_intervals=[3], cache=path)
    assert r2["cached"] is True
    assert (r2["blocks"], r2["sort_interval"]) == (r1["blocks"], r1["sort_interval"])
    key = core.autotune_key(2, 8192, 32, 3, 2, 0.05, 6, [0, 32], [3], 1, "reference",
                            core.device_info())
    assert core.autotune_cache_lookup(path, key)["blocks"] == r1["blocks"]
