"""GPU parity at the benchmark's own sizes, through the C ABI, against the
oracle on the same seeded input:

* the 1D2V Weibel run (4,194,304 particles, nx = 32, degrees (3, 2),
  dt = 0.05, its real initial state): every one of the seven sub-steps of
  one strang_splitting! step (hamiltonian_splitting.jl:98-108), then fused
  steps, then write_step! (diagnostics.jl:186-250) after each of 20 steps;
* the 1D1V strong Landau damping run at its full size (100,000 particles,
  500 steps): the sub-steps of the first step, then rho, the E DOFs, the
  particles and the electric energy after every step;
* the batched ParticleMeshCoupling1D verbs with a nonzero B field, degrees
  1 to 5, both smoothings, ragged particle counts.

Tolerances. Per-particle arithmetic is the reference's, rounded the same
way, so the only difference between the GPU and the oracle is the order in
which deposits (rho, j) are summed. Where a quantity is compared at 1e-12
that is the per-sub-step parity bound, normwise: |a - b|_inf <= 1e-12 |b|_inf.
Over many steps the summation-order differences feed back into the
particles and fields, and some diagnostics cancel (rho is a constant plus a
1e-4 perturbation, so the Poisson field carries 1e4 times the relative
rounding of rho). For those comparisons the tolerance is measured, not
guessed: the oracle is run a second time on the same particles stored in a
permuted order, which changes exactly its deposit summation order and
nothing else, and the GPU must stay within 16 times that spread (plus a
floor of 1e-14 of the value, for quantities whose spread happens to be
zero). |gpu - ref| <= |gpu - exact| + |ref - exact|, and both errors are of
the size the permutation shows, so the factor 16 is a margin, not a fit.
Along a trajectory the spread at one step can be small by chance while the
accumulated divergence is not, so the spread used at step s is the largest
relative spread seen up to s; each step's magnitude is floored at 1e-6 of
the largest magnitude so far, so that a quantity passing through zero does
not turn its relative spread into noise.
"""
import numpy as np
import pytest

import gempic_amd as G
from gempic_amd import workloads
import pyoracle as orc

from test_gpu_parity import REL, _setup_pair, assert_rows_close

pytestmark = pytest.mark.gpu

SPREAD_FACTOR = 16.0
SPREAD_FLOOR = 1e-14


def _field_close(got, ref, rel=REL, what=""):
    scale = max(np.max(np.abs(ref)), 1e-300)
    err = np.max(np.abs(got - ref))
    assert err <= rel * scale, f"{what}: {err:.3e} > {rel} * {scale:.3e}"


def _oracle(w, perm=None, degree0=3, degree1=2):
    """An oracle splitting on the workload's particles (optionally stored in
    the order `perm`), E1 from Poisson."""
    nx = w["nx"]
    parts = w["particles"] if perm is None else w["particles"][:, perm]
    parts = np.asfortranarray(parts.copy())
    e1, e2 = np.zeros(nx), np.zeros(nx)
    b = w.get("b0", np.zeros(nx)).copy()
    osp = orc.Splitting(nx, w["xmin"], w["xmax"], 3, degree0, degree1, parts, e1, e2, b,
                        common_weight=w["common_weight"])
    osp.solve_poisson(e1, np.zeros(nx))
    return osp, parts, (e1, e2, b)


def _trajectory_within_spread(got, ref, alt, what, floor_abs=0.0):
    """got, ref, alt: per-step values (a sequence of scalars or arrays);
    the envelope form of _within_spread (module docstring), plus an absolute
    floor for quantities that are themselves differences of nearly equal
    fields."""
    env, run = 0.0, 0.0
    for s, (g, r, a) in enumerate(zip(got, ref, alt)):
        g, r, a = (np.atleast_1d(np.asarray(v, dtype=float)) for v in (g, r, a))
        mag = float(np.max(np.abs(r)))
        run = max(run, mag)
        scale = max(mag, 1e-6 * run, 1e-300)
        env = max(env, float(np.max(np.abs(a - r))) / scale)
        tol = (SPREAD_FACTOR * env + SPREAD_FLOOR) * scale + floor_abs
        err = float(np.max(np.abs(g - r)))
        assert err <= tol, (f"{what} at step {s}: {err:.3e} > {tol:.3e} "
                            f"(permuted-oracle relative spread {env:.3e})")


def _within_spread(got, ref, perm_val, what):
    spread = np.max(np.abs(np.asarray(perm_val) - np.asarray(ref)))
    tol = SPREAD_FACTOR * spread + SPREAD_FLOOR * np.max(np.abs(ref))
    err = np.max(np.abs(np.asarray(got) - np.asarray(ref)))
    assert err <= tol, f"{what}: {err:.3e} > {tol:.3e} (permuted-oracle spread {spread:.3e})"


# ---------------------------------------------------------------------------
# 1D2V Weibel, 4,194,304 particles

STRANG = (("operatorHB", 0.5), ("operatorHE", 0.5), ("operatorHp2", 0.5), ("operatorHp1", 1.0),
          ("operatorHp2", 0.5), ("operatorHE", 0.5), ("operatorHB", 0.5))


def test_gpu_weibel_4m_every_substep_matches_oracle():
    """The seven sub-steps of one Strang step on the real 4M-particle Weibel
    state, each compared at 1e-12 (particles, E1, E2, B, j1, j2), then two
    fused Strang steps from the state both sides reached."""
    n = 4194304
    print(f"weibel 4M sub-steps: n={n} seed=1", flush=True)
    w = workloads.weibel_1d2v(n, seed=1)
    h, pg, osp, o_parts, (o_e1, o_e2, o_b) = _setup_pair(w)
    dt = w["dt"]
    for name, frac in STRANG:
        getattr(G, name)(h, dt * frac)
        getattr(osp, name)(dt * frac)
        assert_rows_close(pg.array, o_parts, REL)
        _field_close(h.e_dofs[0], o_e1, what=f"{name} E1")
        _field_close(h.e_dofs[1], o_e2, what=f"{name} E2")
        _field_close(h.b_dofs, o_b, what=f"{name} B")
        if name in ("operatorHp1", "operatorHp2"):
            for c in (1, 2):
                jr = osp.j(c)
                if np.any(jr):
                    _field_close(h.j_dofs[c - 1], jr, what=f"{name} j{c}")
    G.strang_splitting_(h, dt, 2)
    osp.strang_splitting(dt, 2)
    assert_rows_close(pg.array, o_parts, REL)
    _field_close(h.e_dofs[0], o_e1, what="fused E1")
    _field_close(h.e_dofs[1], o_e2, what="fused E2")
    _field_close(h.b_dofs, o_b, what="fused B")


WRITE_STEP_COLUMNS = ("Time", "KineticEnergy", "Momentum1", "Momentum2", "PotentialEnergyE1",
                      "PotentialEnergyE2", "PotentialEnergyB3", "Transfer", "VVB", "Poynting",
                      "ErrorPoisson")


def test_gpu_weibel_4m_write_step_every_step():
    """write_step! (all 11 columns, ErrorPoisson from solve_poisson!) after
    each of 20 fused Strang steps of the 4M-particle Weibel run, against the
    oracle's; tolerance from the permuted-oracle spread (module docstring)."""
    n = 4194304
    steps = 20
    print(f"weibel 4M write_step trajectory: n={n} seed=1 steps={steps}", flush=True)
    w = workloads.weibel_1d2v(n, seed=1)
    nx = w["nx"]
    h, pg, osp, o_parts, (o_e1, o_e2, o_b) = _setup_pair(w, sync=False)
    h.to_device()
    perm = np.random.default_rng(12).permutation(n)
    psp, _, (p_e1, p_e2, p_b) = _oracle(w, perm)
    th = G.TimeHistoryDiagnostics(h)
    dt = w["dt"]
    got, ref, alt = [], [], []
    field_spread = 0.0
    for s in range(steps + 1):
        t = s * dt
        ep = np.zeros(nx)
        G.solve_poisson_(ep, h, np.zeros(nx))
        got.append(G.write_step_(th, t, 3, efield_poisson=ep))
        pois = []
        for sp, (a1, a2, ab), out in ((osp, (o_e1, o_e2, o_b), ref), (psp, (p_e1, p_e2, p_b), alt)):
            epo = np.zeros(nx)
            sp.solve_poisson(epo, np.zeros(nx))
            pois.append(epo)
            out.append(sp.write_step(t, 3, a1, a2, ab, a1, a2, epo))
        # the summation order's spread of the two fields ErrorPoisson subtracts
        field_spread = max(field_spread, float(np.max(np.abs(p_e1 - o_e1))),
                           float(np.max(np.abs(pois[1] - pois[0]))))
        if s < steps:
            G.strang_splitting_(h, dt, 1)
            osp.strang_splitting(dt, 1)
            psp.strang_splitting(dt, 1)
    got, ref, alt = np.array(got), np.array(ref), np.array(alt)
    np.testing.assert_array_equal(got[:, 0], ref[:, 0])
    for k in range(1, 10):
        _trajectory_within_spread(got[:, k], ref[:, k], alt[:, k], WRITE_STEP_COLUMNS[k])
    # ErrorPoisson = max |E1 - E_Poisson| is a difference of two fields, each
    # carrying its deposit's summation-order rounding (the Poisson field of a
    # charge that is a constant plus a 1e-4 perturbation: 1e4 times the
    # rounding of rho): its floor is 16 times the permuted oracle's spread of
    # those two fields
    _trajectory_within_spread(got[:, 10], ref[:, 10], alt[:, 10], "ErrorPoisson",
                              floor_abs=SPREAD_FACTOR * 2.0 * field_spread)
    # the energy of the Strang scheme: the same drift on both sides
    energy = lambda d: 0.5 * (d[:, 1] + d[:, 4] + d[:, 5] + d[:, 6])
    _trajectory_within_spread(energy(got), energy(ref), energy(alt), "total energy")


# ---------------------------------------------------------------------------
# 1D1V strong Landau damping at its full size

def test_gpu_landau_1d1v_full_run_matches_oracle():
    """100,000 particles, 500 steps of dt = 0.05: the sub-steps of the first
    step at 1e-12, then rho, E1, the particles and the electric energy
    1/2 l2norm_squared(E1, p-1) after every step, within the permuted-oracle
    spread (module docstring)."""
    n, steps = 100000, 500
    print(f"landau 1d1v full: n={n} seed=1 steps={steps}", flush=True)
    w = workloads.landau_1d1v(n, seed=1)
    nx, dt = w["nx"], w["dt"]
    h, pg, osp, o_parts, (o_e1, _, _) = _setup_pair(w, 3, 2)
    for name in ("operatorHB", "operatorHp1", "operatorHB"):
        getattr(G, name)(h, dt / 2 if name == "operatorHB" else dt)
        getattr(osp, name)(dt / 2 if name == "operatorHB" else dt)
        assert_rows_close(pg.array, o_parts, REL)
        _field_close(h.e_dofs[0], o_e1, what=f"{name} E1")
    # the rest of the run, step by step, with a permuted oracle beside it
    perm = np.random.default_rng(5).permutation(n)
    inv = np.argsort(perm)
    psp, p_parts, (p_e1, _, _) = _oracle(w, perm)
    for name in ("operatorHB", "operatorHp1", "operatorHB"):
        getattr(psp, name)(dt / 2 if name == "operatorHB" else dt)
    om = orc.Maxwell1D(w["xmin"], w["xmax"], nx, 3)
    worst = {}
    traj = {}
    for s in range(1, steps):
        G.strang_splitting_(h, dt, 1)
        osp.strang_splitting(dt, 1)
        psp.strang_splitting(dt, 1)
        rho, rho_o, rho_p = np.zeros(nx), np.zeros(nx), np.zeros(nx)
        G.solve_poisson_(np.zeros(nx), h, rho)
        osp.solve_poisson(np.zeros(nx), rho_o)
        psp.solve_poisson(np.zeros(nx), rho_p)
        en = 0.5 * G.l2norm_squared(h.maxwell_solver, h.e_dofs[0], 2)
        en_o = 0.5 * om.l2norm_squared(o_e1, 2)
        en_p = 0.5 * om.l2norm_squared(p_e1, 2)
        checks = (("rho", rho, rho_o, rho_p), ("E1", h.e_dofs[0], o_e1, p_e1),
                  ("electric energy", en, en_o, en_p))
        if s % 25 == 0 or s == steps - 1:
            checks += tuple((f"particle row {r}", pg.array[r], o_parts[r], p_parts[r][inv])
                            for r in range(2))
        for what, a, b, c in checks:
            t = traj.setdefault(what, ([], [], []))
            for lst, v in zip(t, (a, b, c)):
                lst.append(np.array(v, dtype=float, copy=True))
            worst[what] = max(worst.get(what, 0.0), float(np.max(np.abs(np.asarray(a) - b))))
    print("landau 1d1v worst |gpu - oracle|:", worst, flush=True)
    for what, (g, r, a) in traj.items():
        _trajectory_within_spread(g, r, a, what)


# ---------------------------------------------------------------------------
# ParticleMeshCoupling1D batched verbs with a field

@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("smoothing", ["galerkin", "collocation"])
def test_gpu_pmc1d_verbs_with_b_field(degree, smoothing):
    """add_charge!, add_current_update_v! with a nonzero B (the v update is
    per particle: bitwise), the form without B, evaluate, and the _pp forms,
    on ragged particle counts, against the oracle's per-particle functions
    called in the reference's loop order. Deposits (rho, j) at 1e-12 of the
    largest grid value (summation order); values and velocities bitwise."""
    nx, xmin, xmax = 24, -1.0, 3.5
    L = xmax - xmin
    sm = orc.GALERKIN if smoothing == "galerkin" else orc.COLLOCATION
    mesh = G.OneDGrid(xmin, xmax, nx)
    rng = np.random.default_rng(100 * degree + sm)
    for n in (1, 63, 65, 1000, 4099):
        print(f"pmc1d verbs: degree={degree} {smoothing} n={n}", flush=True)
        p = G.ParticleMeshCoupling1D(mesh, n, degree, smoothing)
        op = orc.PMC1D(xmin, xmax, nx, degree, sm)
        x_old = xmin + L * rng.random(n)
        # moves of up to three cells either way, wrapped afterwards by the caller
        x_new = x_old + (L / nx) * rng.uniform(-3.0, 3.0, n)
        q = rng.uniform(0.5, 1.5, n)
        v = rng.normal(size=n)
        b = rng.normal(size=nx) + 0.5
        qm = -1.7
        # add_current_update_v! with B
        j, j_o = np.zeros(nx), np.zeros(nx)
        v_gpu = G.add_current_update_v_(j, p, x_old, x_new, q, qm, b, v.copy())
        v_o = np.array([op.add_current_update_v(j_o, x_old[i], x_new[i], q[i], qm, b, v[i])
                        for i in range(n)])
        np.testing.assert_array_equal(v_gpu, v_o)
        assert not np.array_equal(v_o, v)
        _field_close(j, j_o, what="j with B")
        if degree >= 1:
            j_pp, j_pp_o = np.zeros(nx), np.zeros(nx)
            v_pp = G.add_current_update_v_pp_(j_pp, p, x_old, x_new, q, qm, b, v.copy())
            v_pp_o = np.array([op.add_current_update_v_pp(j_pp_o, x_old[i], x_new[i], q[i], qm, b,
                                                          v[i]) for i in range(n)])
            np.testing.assert_array_equal(v_pp, v_pp_o)
            _field_close(j_pp, j_pp_o, what="j with B, pp")
        # without B (1d1v form)
        j1, j1_o = np.zeros(nx), np.zeros(nx)
        G.add_current_update_v_(j1, p, x_old, x_new, q, qm, None, v.copy())
        for i in range(n):
            op.add_current_update_v_nob(j1_o, x_old[i], x_new[i], q[i], qm, v[i])
        _field_close(j1, j1_o, what="j without B")
        # add_charge! and evaluate
        rho, rho_o = np.zeros(nx), np.zeros(nx)
        G.add_charge_(rho, p, x_old, q)
        for i in range(n):
            op.add_charge(rho_o, x_old[i], q[i])
        _field_close(rho, rho_o, what="rho")
        vals = G.evaluate(p, x_old, b)
        np.testing.assert_array_equal(vals, np.array([op.evaluate(x_old[i], b) for i in range(n)]))
        if degree >= 1:
            rho_pp, rho_pp_o = np.zeros(nx), np.zeros(nx)
            G.add_charge_pp_(rho_pp, p, x_old, q)
            for i in range(n):
                op.add_charge_pp(rho_pp_o, x_old[i], q[i])
            _field_close(rho_pp, rho_pp_o, what="rho pp")
            bpp = orc.b_to_pp(degree, b)
            vals_pp = G.evaluate_pp(p, x_old, G.b_to_pp(p, nx, b))
            np.testing.assert_array_equal(
                vals_pp, np.array([op.evaluate_pp(x_old[i], bpp) for i in range(n)]))


# ---------------------------------------------------------------------------
# Fused arithmetic (HamiltonianSplitting.set_arithmetic("fused")): the same
# sub-steps with closed-form bases and fma, compared with the oracle at the
# 1e-12 per-sub-step bound. Per-particle results differ from the reference's by a few
# ulp per operation instead of being bitwise equal.

@pytest.mark.parametrize("pair", [(3, 2), (3, 3), (2, 1), (2, 2), (1, 0), (1, 1)])
def test_gpu_fused_arithmetic_substeps_match_oracle(pair):
    """The hot Weibel state (particles cross cells, leave through both ends
    and skip cells: multi-segment line integrals) through every operator, for
    every smoother pair, at 1e-12."""
    from test_gpu_parity import _hot_weibel
    print(f"fused arithmetic sub-steps: pair {pair}", flush=True)
    w = _hot_weibel(20000, seed=11)
    h, pg, osp, o_parts, (o_e1, o_e2, o_b) = _setup_pair(w, *pair)
    h.set_arithmetic("fused")
    dt = 0.4
    for name in ("operatorHp1", "operatorHp2", "operatorHE", "operatorHB", "operatorHp1"):
        getattr(G, name)(h, dt)
        getattr(osp, name)(dt)
        assert_rows_close(pg.array, o_parts, REL)
        scale = max(np.max(np.abs(o_e1)), np.max(np.abs(o_e2)), 1.0)
        assert np.max(np.abs(h.e_dofs[0] - o_e1)) <= REL * scale, name
        assert np.max(np.abs(h.e_dofs[1] - o_e2)) <= REL * scale, name
        _field_close(h.b_dofs, o_b, what=f"{name} B")
        for c in (1, 2):
            jr = osp.j(c)
            if np.any(jr):
                _field_close(h.j_dofs[c - 1], jr, what=f"{name} j{c}")


def test_gpu_fused_arithmetic_weibel_4m_substeps_and_steps():
    """The 4M-particle Weibel state in fused arithmetic: the seven sub-steps
    of one Strang step at 1e-12 against the oracle, then fused Strang steps
    (the benchmark's path) at 1e-12 after one and after three steps."""
    n = 4194304
    print(f"fused arithmetic weibel 4M: n={n} seed=1", flush=True)
    w = workloads.weibel_1d2v(n, seed=1)
    h, pg, osp, o_parts, (o_e1, o_e2, o_b) = _setup_pair(w)
    h.set_arithmetic("fused")
    dt = w["dt"]
    for name, frac in STRANG:
        getattr(G, name)(h, dt * frac)
        getattr(osp, name)(dt * frac)
        assert_rows_close(pg.array, o_parts, REL)
        _field_close(h.e_dofs[0], o_e1, what=f"{name} E1")
        _field_close(h.e_dofs[1], o_e2, what=f"{name} E2")
        _field_close(h.b_dofs, o_b, what=f"{name} B")
    for steps in (1, 2):
        G.strang_splitting_(h, dt, steps)
        osp.strang_splitting(dt, steps)
        assert_rows_close(pg.array, o_parts, REL)
        _field_close(h.e_dofs[0], o_e1, what="fused E1")
        _field_close(h.e_dofs[1], o_e2, what="fused E2")
        _field_close(h.b_dofs, o_b, what="fused B")


def test_gpu_fused_arithmetic_settings():
    """Fused arithmetic needs 1d2v and the one-kernel step; the reference
    arithmetic stays the default and is restored bitwise."""
    from test_gpu_parity import _hot_weibel
    w = _hot_weibel(4096, seed=2)
    a, pga, *_ = _setup_pair(w)
    b, pgb, *_ = _setup_pair(w)
    b.set_arithmetic("fused")
    b.set_arithmetic("reference")
    G.strang_splitting_(a, 0.1, 3)
    G.strang_splitting_(b, 0.1, 3)
    np.testing.assert_array_equal(pga.array, pgb.array)
    with pytest.raises(ValueError):
        b.set_arithmetic("fused")
        b.set_fusion(2)
    b.set_arithmetic("reference")
    b.set_fusion(2)
    with pytest.raises(ValueError):
        b.set_arithmetic("fused")
    wl = workloads.landau_1d1v(1000, seed=2)
    c, *_ = _setup_pair(wl, 3, 2)
    with pytest.raises(ValueError):
        c.set_arithmetic("fused")


# ---------------------------------------------------------------------------
# Particle counts that do not fill the launch evenly. With 2 or more
# workgroups per CU the pass splits the particles unevenly between the
# workgroup generations (oldest first, in whole waves of 64); a count that is
# not a multiple of 64 x CUs leaves the last workgroups short or empty.

@pytest.mark.parametrize("arith", ["reference", "fused"])
@pytest.mark.parametrize("n", [3000017, 1048577, 700001])
def test_gpu_ragged_counts_fused_steps_match_oracle(n, arith):
    print(f"ragged counts: n={n} arithmetic={arith}", flush=True)
    w = workloads.weibel_1d2v(n, seed=4)
    h, pg, osp, o_parts, (o_e1, o_e2, o_b) = _setup_pair(w)
    h.set_arithmetic(arith)
    dt = w["dt"]
    G.strang_splitting_(h, dt, 3)
    osp.strang_splitting(dt, 3)
    assert_rows_close(pg.array, o_parts, REL)
    _field_close(h.e_dofs[0], o_e1, what="E1")
    _field_close(h.e_dofs[1], o_e2, what="E2")
    _field_close(h.b_dofs, o_b, what="B")
    # the deposits' summation order is a function of the particle order
    # only: the same run again gives the same bits
    h2, pg2, _, _, _ = _setup_pair(w)
    h2.set_arithmetic(arith)
    G.strang_splitting_(h2, dt, 3)
    np.testing.assert_array_equal(pg2.array, pg.array)
    np.testing.assert_array_equal(h2.e_dofs[0], h.e_dofs[0])
    np.testing.assert_array_equal(h2.b_dofs, h.b_dofs)
