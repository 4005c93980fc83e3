"""The oracle's task split, the reference's nthreads() scheme
(hamiltonian_splitting.jl:61-66, hamiltonian_splitting_1d2v.jl:48-164):
particles in n_jobs equal chunks, each chunk's deposit into its own buffer,
the buffers summed in chunk order (Base's reduce over fewer than 1024
vectors is the left fold). Pinned by running each chunk on its own through
the one-task oracle and folding its buffers by hand: the deposits must be
bitwise that fold, the particles bitwise the one-task run's (their
arithmetic does not depend on the split).
"""
import math

import numpy as np
import pytest

import pyoracle as orc


def _state(n=4096, nx=16, seed=5):
    rng = np.random.default_rng(seed)
    L = 2 * math.pi / 1.25
    parts = np.asfortranarray(np.vstack([L * rng.random(n), rng.normal(size=n) * 0.6,
                                         rng.normal(size=n) * 0.4, np.full(n, L)]))
    e1, e2, b = (rng.normal(size=nx) * 0.1 for _ in range(3))
    return parts, e1, e2, b, L, nx


def _split(parts, e1, e2, b, L, nx, cw):
    return orc.Splitting(nx, 0.0, L, 3, 3, 2, parts, e1, e2, b, common_weight=cw)


def _fold(bufs):
    v = bufs[0] + bufs[1]
    for t in bufs[2:]:
        v = v + t
    return v


@pytest.mark.parametrize("tasks", [2, 4, 8])
def test_task_split_deposits_are_the_chunk_fold(tasks):
    parts, e1, e2, b, L, nx = _state()
    n = parts.shape[1]
    cw = 1.0 / n
    c = n // tasks

    # Hp1 (dt 0.7, particles cross cells): j1 = fold of the chunks' currents,
    # j2 = fold of the chunks' rho at the new positions
    pa = parts.copy(order="F")
    s = _split(pa, e1.copy(), e2.copy(), b.copy(), L, nx, cw)
    s.set_tasks(tasks)
    s.operatorHp1(0.7)
    j1, j2 = s.j(1), s.j(2)
    cur, rho, moved = [], [], []
    for t in range(tasks):
        pc = np.asfortranarray(parts[:, t * c:(t + 1) * c].copy())
        sc = _split(pc, e1.copy(), e2.copy(), b.copy(), L, nx, cw)
        sc.operatorHp1(0.7)
        cur.append(sc.j(1))
        rho.append(sc.j(2))
        moved.append(pc)
    np.testing.assert_array_equal(j1, _fold(cur))
    np.testing.assert_array_equal(j2, _fold(rho))
    np.testing.assert_array_equal(pa, np.hstack(moved))

    # Hp2 with dt = 1 (j2 = fold * 1.0 is the fold itself)
    pa = parts.copy(order="F")
    s = _split(pa, e1.copy(), e2.copy(), b.copy(), L, nx, cw)
    s.set_tasks(tasks)
    s.operatorHp2(1.0)
    bufs = []
    for t in range(tasks):
        pc = np.asfortranarray(parts[:, t * c:(t + 1) * c].copy())
        sc = _split(pc, e1.copy(), e2.copy(), b.copy(), L, nx, cw)
        sc.operatorHp2(1.0)
        bufs.append(sc.j(2))
    np.testing.assert_array_equal(s.j(2), _fold(bufs))


def test_task_split_strang_steps_match_one_task():
    parts, e1, e2, b, L, nx = _state(seed=8)
    n = parts.shape[1]
    runs = []
    for tasks in (1, 4):
        pa = parts.copy(order="F")
        f = [e1.copy(), e2.copy(), b.copy()]
        s = _split(pa, *f, L, nx, 1.0 / n)
        s.set_tasks(tasks)
        s.strang_splitting(0.1, 3)
        runs.append((pa, f))
    (p1, f1), (p4, f4) = runs
    # only the deposits' summation order differs
    assert np.max(np.abs(p4 - p1)) <= 1e-12 * np.max(np.abs(p1))
    for a, r in zip(f4, f1):
        assert np.max(np.abs(a - r)) <= 1e-12 * np.max(np.abs(r))


def test_task_split_needs_equal_chunks():
    parts, e1, e2, b, L, nx = _state(n=1000)
    s = _split(parts, e1, e2, b, L, nx, 1e-3)
    with pytest.raises(ValueError):
        s.set_tasks(3)                       # @assert np % n_jobs == 0
    s.set_tasks(8)
