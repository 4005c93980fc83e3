"""GPU parity of the 2D3V Hamiltonian splitting (gempic_splitting2d_*,
split2d.hip) through the C ABI, against the oracle's 2D3V splitting
(oracle/gempic_oracle2d3v.c, pinned by Gauss's law, the second-order energy
error and the reduction to the reference's 1D2V splitting).

Tolerances. Hp1 and Hp2 move the positions with the oracle's operations,
so from the same input the positions are bitwise the oracle's. The kicks
evaluate the shape factors in closed form and sum the gathers with fused
multiply-adds (the oracle: the de Boor recurrence, products and sums
rounded one by one), so they agree to rounding (1e-12 of each row's
largest value). Hp1 and Hp2 also integrate the degree
p-1 splines along the path through their antiderivative (differences of
suffix sums of the degree-p values at the two ends) where the oracle
integrates piece by piece with Gauss-Legendre: the same integrals, rounded
differently, so the magnetic kicks agree to rounding (1e-12 of each row's
largest value). The current deposits add in another order (wave sums,
workgroup grids, timing-dependent between flushes), so the fields that take
a current (E after Hp1, Hp2, Hp3) agree to 1e-12 of their largest value;
HE's B update is the TwoDMaxwell kernel, bitwise equal to the oracle on
equal input. HB's E update and the E updates from the currents apply the
inverse mass matrices as circulant convolutions (kernels from the DFT
eigenvalues in long double) where the oracle solves through the DFT: the
same linear maps, so they agree to rounding. Over several Strang steps the differences feed back:
1e-10 after three steps of a hot state.
"""
import math

import numpy as np
import pytest

import gempic_amd as G
from gempic_amd import workloads
import pyoracle as orc

pytestmark = pytest.mark.gpu

REL = 1e-12


def _hot(n=20000, nx=12, ny=10, L1=4 * math.pi, L2=3.0, seed=1, vscale=0.8, fscale=0.3):
    rng = np.random.default_rng(seed)
    parts = np.asfortranarray(np.vstack([L1 * rng.random(n), L2 * rng.random(n),
                                         rng.normal(size=(3, n)) * vscale, 1.0 + rng.random(n)]))
    e = [fscale * rng.normal(size=nx * ny) for _ in range(3)]
    b = [fscale * rng.normal(size=nx * ny) for _ in range(3)]
    return dict(particles=parts, e=e, b=b, nx=nx, ny=ny, xmin=0.0, xmax=L1, ymin=0.0, ymax=L2,
                common_weight=1.0 / n)


def _pair(w, degree=3, sync=True):
    n = w["particles"].shape[1]
    nx, ny = w["nx"], w["ny"]
    mesh = G.TwoDGrid(w["xmin"], w["xmax"], nx, w["ymin"], w["ymax"], ny)
    pg = G.ParticleGroup(2, 3, n, common_weight=w["common_weight"])
    pg.array[:] = w["particles"]
    ks0 = G.ParticleMeshCoupling2D(pg, mesh, degree, "galerkin")
    ks1 = G.ParticleMeshCoupling2D(pg, mesh, degree - 1, "galerkin")
    mx = G.TwoDMaxwell(mesh, degree)
    e = [a.copy() for a in w["e"]]
    b = [a.copy() for a in w["b"]]
    h = G.HamiltonianSplitting2D3V(mx, ks0, ks1, pg, e, b, sync=sync)
    o_parts = np.asfortranarray(w["particles"].copy())
    o_e = [a.copy() for a in w["e"]]
    o_b = [a.copy() for a in w["b"]]
    osp = orc.Splitting2D3V(nx, ny, w["xmin"], w["xmax"], w["ymin"], w["ymax"], degree, o_parts,
                            o_e, o_b, common_weight=w["common_weight"])
    return h, pg, osp, o_parts, o_e, o_b


def _close(got, ref, rel, what):
    scale = max(np.max(np.abs(ref)), 1e-300)
    err = np.max(np.abs(got - ref))
    assert err <= rel * scale, f"{what}: {err:.3e} > {rel} * {scale:.3e}"


@pytest.mark.parametrize("degree, nx, ny", [(2, 12, 10), (3, 12, 10), (3, 9, 7), (3, 16, 11)])
def test_gpu_2d3v_operators_match_oracle(degree, nx, ny):
    """Grid sizes: the inverse-mass convolutions take 4 outputs per work item
    along a line of a length divisible by 4, 2 for an even length, else 1."""
    print(f"2d3v operators: degree {degree}, {nx} x {ny}", flush=True)
    w = _hot(nx=nx, ny=ny)
    h, pg, osp, o_parts, o_e, o_b = _pair(w, degree)
    dt = 0.3
    for k, name in enumerate(("operatorHp1", "operatorHp2", "operatorHp3", "operatorHE",
                              "operatorHB", "operatorHp1", "operatorHp2")):
        getattr(h, name)(dt)
        getattr(osp, name)(dt)
        if k == 0:
            # Hp1 from the common input: the positions bitwise, the kicks to rounding
            np.testing.assert_array_equal(pg.array[[0, 1, 2, 5]], o_parts[[0, 1, 2, 5]],
                                          err_msg=name)
        for r in range(5):
            _close(pg.array[r], o_parts[r], REL, f"{name} row {r}")
        for c in range(3):
            _close(h.e_dofs[c], o_e[c], REL, f"{name} E{c + 1}")
            _close(h.b_dofs[c], o_b[c], REL, f"{name} B{c + 1}")
        if name.startswith("operatorHp"):
            c = int(name[-1])
            _close(h.j_dofs[c - 1], osp.j(c), REL, f"{name} j{c}")
        # the particles stay inside the domain
        assert np.all(pg.array[0] >= 0) and np.all(pg.array[0] < w["xmax"])
        assert np.all(pg.array[1] >= 0) and np.all(pg.array[1] < w["ymax"])


def test_gpu_2d3v_strang_steps_and_gauss_law():
    print("2d3v strang steps", flush=True)
    w = _hot(seed=4)
    h, pg, osp, o_parts, o_e, o_b = _pair(w, 3, sync=False)
    g0 = h.rho() - h.rho_from_e()
    h.set_sort_interval(2)
    h.strang_splitting(0.1, 3)
    osp.strang_splitting(0.1, 3)
    h.to_host()
    _close(pg.array, o_parts, 1e-10, "particles")
    for c in range(3):
        _close(h.e_dofs[c], o_e[c], 1e-10, f"E{c + 1}")
        _close(h.b_dofs[c], o_b[c], 1e-10, f"B{c + 1}")
    # Gauss's law on the GPU's own state (the deposits conserve charge)
    g1 = h.rho() - h.rho_from_e()
    assert np.max(np.abs(g1 - g0)) <= 1e-13 * np.max(np.abs(g0))


def test_gpu_2d3v_landau_4m_substeps_match_oracle():
    """The 2D3V Landau configuration's geometry and distribution (64 x 64,
    degree 3, [0, 4 pi)^2, v ~ N(0, I)) at 4,194,304 particles: the nine
    sub-steps of one Strang step against the oracle."""
    n = 4194304
    print(f"2d3v landau sub-steps: n={n}", flush=True)
    w = workloads.landau_2d3v(n, seed=3)
    nn = w["nx"] * w["ny"]
    # E from Poisson of the deposited charge (TwoDPoisson), B = 0
    mx = orc.Maxwell2D(w["xmin"], w["xmax"], w["nx"], w["ymin"], w["ymax"], w["ny"], 3)
    pm = orc.PMC2D(w["xmin"], w["xmax"], w["nx"], w["ymin"], w["ymax"], w["ny"], 3, orc.GALERKIN)
    rho = np.zeros(nn)
    pm.add_charge_n(rho, w["particles"][0], w["particles"][1],
                    w["particles"][5] * w["common_weight"])
    e = [np.zeros(nn), np.zeros(nn), np.zeros(nn)]
    mx.compute_e_from_rho(e, rho)
    w["e"] = e
    w["b"] = [np.zeros(nn) for _ in range(3)]
    h, pg, osp, o_parts, o_e, o_b = _pair(w, 3)
    dt = w["dt"]
    for name, f in (("operatorHB", 0.5), ("operatorHE", 0.5), ("operatorHp3", 0.5),
                    ("operatorHp2", 0.5), ("operatorHp1", 1.0), ("operatorHp2", 0.5),
                    ("operatorHp3", 0.5), ("operatorHE", 0.5), ("operatorHB", 0.5)):
        getattr(h, name)(dt * f)
        getattr(osp, name)(dt * f)
        for r in range(5):
            _close(pg.array[r], o_parts[r], REL, f"{name} row {r}")
        for c in range(3):
            _close(h.e_dofs[c], o_e[c], REL, f"{name} E{c + 1}")
            _close(h.b_dofs[c], o_b[c], REL, f"{name} B{c + 1}")


def test_gpu_2d3v_errors_and_order():
    w = _hot(n=1000)
    h, pg, *_ = _pair(w, 3)
    h.operatorHE(0.0)       # no change; the caller's order comes back
    np.testing.assert_array_equal(pg.array, w["particles"])
    pg.array[0, 17] = float("nan")
    with pytest.raises(G.InexactError):
        h.operatorHE(0.1)
    with pytest.raises(ValueError):
        mesh = G.TwoDGrid(0.0, 1.0, 128, 0.0, 1.0, 64)
        pg2 = G.ParticleGroup(2, 3, 10)
        G.HamiltonianSplitting2D3V(G.TwoDMaxwell(mesh, 3), G.ParticleMeshCoupling2D(pg2, mesh, 3, "galerkin"),
                                   G.ParticleMeshCoupling2D(pg2, mesh, 2, "galerkin"), pg2,
                                   [np.zeros(128 * 64)] * 3, [np.zeros(128 * 64)] * 3)
