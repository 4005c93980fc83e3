#!/usr/bin/env python3
"""Benchmark: 1D2V Weibel instability, 4,194,304 particles, FP64, on one GPU.

A step is one Strang step (HB HE Hp2 Hp1 Hp2 HE HB) of all particles, run by
the fused path of the HIP library (one particle pass per step plus the field
solves in its last workgroup). Inputs are resident in GPU memory before the
timed region starts.

Prints one JSON line: particle-steps/s over the timed steps, the roofline of
the fused pass (HIP-event kernel times on the library's stream, algorithmic
56 bytes per particle-step: read x, v1, v2, w; write x, v1, v2), and a CPU
baseline (the C oracle, one thread, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gempic.jl_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
BYTES_FUSED = 56          # per particle-step
BYTES_SUBSTEP = 208       # per particle-step, sub-step kernels (HE 40x2, Hp2 40x2, Hp1 48)


def make_splitting(G, w, sync=False):
    n = w["particles"].shape[1]
    nx = w["nx"]
    mesh = G.OneDGrid(w["xmin"], w["xmax"], nx)
    pg = G.ParticleGroup(1, 2, n, common_weight=w["common_weight"])
    pg.array[:] = w["particles"]
    ks0 = G.ParticleMeshCoupling1D(mesh, n, w["degree"], "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, n, w["degree"] - 1, "galerkin")
    mx = G.Maxwell1DFEM(mesh, w["degree"])
    e1, e2, b = np.zeros(nx), np.zeros(nx), w["b0"].copy()
    h = G.HamiltonianSplitting(1, 2, mx, ks0, ks1, pg, [e1, e2], b, sync=False)
    rho = np.zeros(nx)
    G.solve_poisson_(e1, h, rho)          # E1 from Poisson (diagnostics.jl:15-31)
    h.to_device(particles=False)
    return h


def arith_run(G, _lib, w, arith, warmup, steps, blocks, interval, sort_s):
    """A fused-path run of `steps` Strang steps of w in the given arithmetic:
    particle-steps/s (the sort charged per step as in the headline when
    sort_s is given, else measured in the window), the pass's HIP-event time
    and its roofline at 56 B per particle-step."""
    L = _lib.load()
    n = w["particles"].shape[1]
    h = make_splitting(G, w)
    h.set_arithmetic(arith)
    _lib.check(L.gempic_set_launch(h._h, blocks))
    _lib.check(L.gempic_set_sort_interval(h._h, interval))
    if warmup:
        _lib.check(L.gempic_strang_splitting(h._h, w["dt"], warmup, _lib.STRANG_FUSED))
    _lib.check(L.gempic_sort_particles(h._h))
    _lib.check(L.gempic_profile_enable(h._h, 10))
    _lib.check(L.gempic_profile_reset(h._h))
    t0 = time.perf_counter()
    _lib.check(L.gempic_strang_splitting(h._h, w["dt"], steps, _lib.STRANG_FUSED))
    _lib.check(L.gempic_synchronize(h._h))
    el = time.perf_counter() - t0
    if sort_s is not None and interval > 0:
        el += (steps / interval - (steps - 1) // interval) * sort_s
    ms, cnt = h.profile_read(0)
    epi_ms, epi_cnt = h.profile_read(9)
    h.profile(False)
    pass_s = 1e-3 * ms / max(cnt, 1)
    ach = BYTES_FUSED * n / pass_s / 1e9
    del h
    return {"arithmetic": arith, "particles": n, "steps": steps,
            "value": n * steps / el, "ms_per_step": 1e3 * el / steps,
            "pass_us": 1e6 * pass_s, "field_part_us": 1e3 * epi_ms / max(epi_cnt, 1),
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "bytes_per_particle_step": BYTES_FUSED}}


BYTES_2D3V = 472          # per particle-step of the 2D3V splitting's passes: HE 64 x 2,
                          # Hp3 64 x 2, Hp2 72 x 2, Hp1 72 (the state read: 40 B by HE,
                          # 48 B with the weights by the others; the pushed components
                          # written)


def split2d_run(G, n, steps, cpu_n, cpu_steps):
    """The 2D3V Landau configuration (16,777,216 particles, 64 x 64, degree
    3, dt = 0.05) through the 2D3V splitting: E from Poisson of the charge,
    B = 0, `steps` Strang steps timed on the GPU; the oracle's 2D3V splitting
    on a bounded sample of the same workload on one core."""
    from gempic_amd import workloads
    print(f"[bench] 2D3V Landau, {n} particles", file=sys.stderr, flush=True)
    w = workloads.landau_2d3v(n, seed=3)
    nx, ny = w["nx"], w["ny"]
    mesh = G.TwoDGrid(w["xmin"], w["xmax"], nx, w["ymin"], w["ymax"], ny)
    pg = G.ParticleGroup(2, 3, n, common_weight=w["common_weight"])
    pg.array[:] = w["particles"]
    mx = G.TwoDMaxwell(mesh, 3)
    ks0 = G.ParticleMeshCoupling2D(pg, mesh, 3, "galerkin")
    ks1 = G.ParticleMeshCoupling2D(pg, mesh, 2, "galerkin")
    e = [np.zeros(nx * ny) for _ in range(3)]
    b = [np.zeros(nx * ny) for _ in range(3)]
    h = G.HamiltonianSplitting2D3V(mx, ks0, ks1, pg, e, b, sync=False)
    rho = h.rho()
    G.compute_e_from_rho_(e, mx, rho)
    h.to_device(particles=False)
    h.strang_splitting(w["dt"], 2)           # warmup
    h.synchronize()
    t0 = time.perf_counter()
    h.strang_splitting(w["dt"], steps)
    h.synchronize()
    el = time.perf_counter() - t0
    rate = n * steps / el
    ach = BYTES_2D3V * rate / 1e9
    out = {"workload": f"2D3V Landau damping, {n} particles, 64x64 on [0,4pi)^2, degree 3, "
                       f"dt 0.05, {steps} steps",
           "value": rate, "unit": "particle-steps/s", "ms_per_step": 1e3 * el / steps,
           "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "bytes_per_particle_step": BYTES_2D3V,
                        "note": "whole-step time (all passes, field solves and sorts)"}}
    del h
    if cpu_n:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc
        parts = np.asfortranarray(w["particles"][:, :cpu_n].copy())
        parts[5] = w["particles"][5, :cpu_n]
        ec = [a.copy() for a in e]
        bc = [np.zeros(nx * ny) for _ in range(3)]
        osp = orc.Splitting2D3V(nx, ny, w["xmin"], w["xmax"], w["ymin"], w["ymax"], 3, parts,
                                ec, bc, common_weight=1.0 / cpu_n)
        t0 = time.perf_counter()
        osp.strang_splitting(w["dt"], cpu_steps)
        secs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cpu_n * cpu_steps / secs, "unit": "particle-steps/s",
                               "cores": 1, "kind": "port",
                               "sample": f"{cpu_n} particles of the same sample x {cpu_steps} "
                                         f"steps ({secs:.1f} s), the oracle's 2D3V splitting, "
                                         "one thread"}
    return out


def cpu_threads():
    """The host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS where that is set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(n, steps, seed=1, tasks=1):
    """The oracle's 1D2V Strang splitting on the Weibel sample: one thread,
    or `tasks` threads with the reference's nthreads() task split (equal
    particle chunks, a buffer per chunk, buffers summed in chunk order)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc
    from gempic_amd import workloads
    w = workloads.weibel_1d2v(n, seed=seed)
    nx = w["nx"]
    parts = np.asfortranarray(w["particles"])
    e1, e2, b = np.zeros(nx), np.zeros(nx), w["b0"].copy()
    osp = orc.Splitting(nx, w["xmin"], w["xmax"], 3, 3, 2, parts, e1, e2, b,
                        common_weight=w["common_weight"])
    rho = np.zeros(nx)
    osp.solve_poisson(e1, rho)
    if tasks > 1:
        osp.set_tasks(tasks)
    t0 = time.perf_counter()
    osp.strang_splitting(w["dt"], steps)
    dt = time.perf_counter() - t0
    return n * steps / dt, dt


def pieces_2d(n, reps, cpu_n):
    """The 2D3V configuration's particle-mesh pieces on one GPU: 64 x 64 grid
    on [0, 4 pi)^2, degree-3 tensor splines, n particles sampled from the
    density 1 + 0.01 cos(0.5 x) cos(0.5 y) (seeded, rejection sampling), kept
    in GPU memory. Times the device-resident add_charge! (fast and
    deterministic, particles in sampled order and sorted by cell), evaluate
    and evaluate_multiple with HIP events on the stream the kernels run on.
    Algorithmic bytes per particle: deposit 24 (x, y, wp read), evaluate 24
    (x, y read, value written), evaluate_multiple 32."""
    import gempic_amd as G
    from gempic_amd.device import DeviceArray, Event, Stream
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    nx = ny = 64
    L = 4.0 * np.pi
    rng = np.random.default_rng(3)
    xs, ys, have = [], [], 0
    while have < n:
        m = int((n - have) * 1.02) + 1024
        x, y, u = L * rng.random(m), L * rng.random(m), rng.random(m)
        keep = u * 1.01 < 1.0 + 0.01 * np.cos(0.5 * x) * np.cos(0.5 * y)
        xs.append(x[keep])
        ys.append(y[keep])
        have += int(keep.sum())
    x = np.concatenate(xs)[:n]
    y = np.concatenate(ys)[:n]
    w = np.full(n, 1.0 / n)
    cell = np.floor(x / (L / nx)).astype(np.int64) + nx * np.floor(y / (L / ny)).astype(np.int64)
    order = np.argsort(cell, kind="stable")
    D = DeviceArray.from_numpy
    dx_, dy_, dw = D(x), D(y), D(w)
    dxs, dys = D(x[order]), D(y[order])
    grid = G.TwoDGrid(0.0, L, nx, 0.0, L, ny)
    pm = G.ParticleMeshCoupling2D(G.ParticleGroup(2, 2, n), grid, 3, "galerkin")
    rho = DeviceArray.zeros(nx * ny)
    f1, f2 = D(rng.normal(size=nx * ny)), D(rng.normal(size=nx * ny))
    o1, o2 = DeviceArray(n), DeviceArray(n)
    st = Stream()

    def timed(fn, nbytes):
        for _ in range(3):
            fn()
        G.pmc2d_sync(st)
        e0, e1 = Event(), Event()
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        G.pmc2d_sync(st)
        us = 1e3 * e0.elapsed_ms(e1) / reps
        gbs = nbytes * n / (us * 1e-6) / 1e9
        return {"us_per_call": us, "particles_per_s": n / (us * 1e-6), "achieved_GBs": gbs,
                "roofline_frac": gbs / HBM_PEAK_GBS, "bytes_per_particle": nbytes}

    out = {"workload": f"2D3V Landau pieces: {n} particles, 64x64 grid on [0,4pi)^2, degree 3",
           "bound": "hbm", "peak_GBs": HBM_PEAK_GBS}
    kw = dict(async_=True, stream=st)
    out["add_charge_fast"] = timed(lambda: G.add_charge_device_(rho, pm, dx_, dy_, dw, **kw), 24)
    out["add_charge_fast_sorted"] = timed(
        lambda: G.add_charge_device_(rho, pm, dxs, dys, dw, **kw), 24)
    out["add_charge_deterministic"] = timed(
        lambda: G.add_charge_device_(rho, pm, dx_, dy_, dw, deterministic=True, **kw), 24)
    out["add_charge_deterministic_sorted"] = timed(
        lambda: G.add_charge_device_(rho, pm, dxs, dys, dw, deterministic=True, **kw), 24)
    out["evaluate"] = timed(lambda: G.evaluate_device(pm, dx_, dy_, f1, o1, **kw), 24)
    out["evaluate_sorted"] = timed(lambda: G.evaluate_device(pm, dxs, dys, f1, o1, **kw), 24)
    out["evaluate_multiple"] = timed(
        lambda: G.evaluate_multiple_device(pm, (dx_, dy_), (f1, f2), (o1, o2), **kw), 32)
    # TwoDMaxwell at 64 x 64 through the host-array verbs (copies included)
    m = G.TwoDMaxwell(grid, 3)
    e = [np.random.default_rng(1).normal(size=nx * ny) for _ in range(3)]
    b = [np.random.default_rng(2).normal(size=nx * ny) for _ in range(3)]
    t0 = time.perf_counter()
    for _ in range(20):
        G.compute_e_from_b_(e, m, 0.05, b)
    out["maxwell2d_compute_e_from_b_host_call_us"] = 1e6 * (time.perf_counter() - t0) / 20
    if cpu_n:
        import pyoracle as orc
        p = orc.PMC2D(0.0, L, nx, 0.0, L, ny, 3, orc.GALERKIN)
        r = np.zeros(nx * ny)
        t0 = time.perf_counter()
        p.add_charge_n(r, x[:cpu_n], y[:cpu_n], w[:cpu_n])
        secs = time.perf_counter() - t0
        out["cpu_baseline_add_charge"] = {
            "value": cpu_n / secs, "unit": "particles/s", "cores": 1, "kind": "port",
            "sample": f"{cpu_n} particles of the same sample ({secs:.2f} s), C restatement, one thread"}
    return out


def boris_run(G, _lib, w, steps, warmup):
    """HamiltonianSplittingBoris on the same 1D2V Weibel particles (E1 from
    Poisson): staggering!, then strang_splitting! steps, one fused particle
    pass per step (56 algorithmic bytes per particle-step, as the Strang
    pass), with HIP-event kernel times on the library's stream."""
    n = w["particles"].shape[1]
    nx = w["nx"]
    mesh = G.OneDGrid(w["xmin"], w["xmax"], nx)
    pg = G.ParticleGroup(1, 2, n, common_weight=w["common_weight"])
    pg.array[:] = w["particles"]
    ks0 = G.ParticleMeshCoupling1D(mesh, n, w["degree"], "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, n, w["degree"] - 1, "galerkin")
    mx = G.Maxwell1DFEM(mesh, w["degree"])
    e1, e2, b = np.zeros(nx), np.zeros(nx), w["b0"].copy()
    h = G.HamiltonianSplittingBoris(mx, ks0, ks1, pg, [e1, e2], b, sync=False)
    G.solve_poisson_(e1, h, np.zeros(nx))
    h.to_device(particles=False)
    L = _lib.load()
    dt = w["dt"]
    _lib.check(L.gempic_boris_staggering(h._h, dt))
    if warmup:
        _lib.check(L.gempic_boris_strang_splitting(h._h, dt, warmup, _lib.ASYNC))
    _lib.check(L.gempic_synchronize(h._h))
    _lib.check(L.gempic_profile_enable(h._h, 10))
    _lib.check(L.gempic_profile_reset(h._h))
    t0 = time.perf_counter()
    _lib.check(L.gempic_boris_strang_splitting(h._h, dt, steps, _lib.ASYNC))
    _lib.check(L.gempic_synchronize(h._h))
    elapsed = time.perf_counter() - t0
    pass_ms, pass_cnt = h.profile_read(0)
    epi_ms, epi_cnt = h.profile_read(9)
    h.profile(False)
    pass_s = pass_ms / pass_cnt * 1e-3 if pass_cnt else float("nan")
    achieved = BYTES_FUSED * n / pass_s / 1e9
    return {"workload": f"1D2V Weibel, {n} particles, HamiltonianSplittingBoris",
            "value": n * steps / elapsed, "unit": "particle-steps/s", "steps": steps,
            "ms_per_step": 1e3 * elapsed / steps,
            "pass_us": pass_s * 1e6,
            "field_part_us": (epi_ms / epi_cnt * 1e3) if epi_cnt else None,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "bytes_per_particle_step": BYTES_FUSED,
                         "kernel": "k_boris1d fused Boris step"}}


def sampling_run(G, n, cpu_n):
    """sample!(pg, 0.5, 0.5, 1.0, mesh) (the strong-Landau initial condition)
    of n particles straight into a splitting's GPU particles, including the
    initial sort; wall time of the call, and the oracle on cpu_n particles."""
    import math
    L = 4 * math.pi
    mesh = G.OneDGrid(0.0, L, 32)
    pg = G.ParticleGroup(1, 2, n)
    ks0 = G.ParticleMeshCoupling1D(mesh, n, 3, "galerkin")
    ks1 = G.ParticleMeshCoupling1D(mesh, n, 2, "galerkin")
    h = G.HamiltonianSplitting(1, 2, G.Maxwell1DFEM(mesh, 3), ks0, ks1, pg,
                               [np.zeros(32), np.zeros(32)], np.zeros(32), sync=False)
    G.sample_particles_(h, 0.5, 0.5, 1.0)               # warm-up (kernel load)
    t0 = time.perf_counter()
    G.sample_particles_(h, 0.5, 0.5, 1.0)
    secs = time.perf_counter() - t0
    out = {"workload": f"sample!(pg, 0.5, 0.5, 1.0, mesh), {n} particles, 1D2V, into GPU memory",
           "ms": secs * 1e3, "particles_per_s": n / secs}
    if cpu_n:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc
        t0 = time.perf_counter()
        orc.sample_landau(2, cpu_n, 0.5, 0.5, 1.0, L)
        cs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cpu_n / cs, "unit": "particles/s", "cores": 1, "kind": "port",
                               "sample": f"{cpu_n} particles ({cs:.2f} s), C restatement, one thread"}
    return out


BYTES_REV = 80            # reverse pass: read x, v1, v2, w and the cotangents, write the cotangents
BYTES_RECOMPUTE = 56      # forward re-run of a checkpoint interval: read x, v1, v2, w; write x, v1, v2


def adjoint_run(G, _lib, n, steps, every, cpu_n, cpu_steps):
    """strang_splitting! forward, then forward + reverse mode (the gradient
    of the time-integrated magnetic energy with respect to the initial
    particles and fields, checkpoints every `every` steps), on the Weibel
    initial condition whose positions are a smooth function of alpha; wall
    times of both calls, HIP-event times of the reverse kernels, and the
    gradients with respect to alpha and beta."""
    from gempic_amd import workloads
    L = _lib.load()
    w = workloads.weibel_adjoint_1d2v(n, seed=4)
    h = make_splitting(G, w)
    dt = w["dt"]
    G.strang_splitting_adjoint_(h, dt, 4, checkpoint_every=2, fetch=False)   # kernel loads
    h.to_device()
    _lib.check(L.gempic_synchronize(h._h))
    t0 = time.perf_counter()
    _lib.check(L.gempic_strang_splitting(h._h, dt, steps, _lib.STRANG_FUSED))
    _lib.check(L.gempic_synchronize(h._h))
    t_fwd = time.perf_counter() - t0
    h.to_device()
    _lib.check(L.gempic_profile_enable(h._h, 10))
    _lib.check(L.gempic_profile_reset(h._h))
    t0 = time.perf_counter()
    G.strang_splitting_adjoint_(h, dt, steps, checkpoint_every=every, fetch=False)
    t_adj = time.perf_counter() - t0
    kern = {}
    for kid, name in ((0, "fused_pass"), (9, "deposit_sum_and_solves"), (10, "reverse_pass"),
                      (11, "recompute_pass"), (12, "reverse_field_part")):
        ms, cnt = h.profile_read(kid)
        if cnt:
            kern[name] = {"avg_us": 1e3 * ms / cnt, "count": cnt}
    h.profile(False)
    r = G.get_adjoint(h)
    h.to_device()                                  # the initial positions again
    gpois = G.solve_poisson_adjoint(h, r.g_e1)
    # k moves Lx = 2 pi / k: positions x = Lx xi, weights w = Lx, E1 (Poisson) ~ Lx
    L, k = w["xmax"], w["k"]
    dj_dl = (r.g_length + r.g_weight_sum + float(r.g_particles[0] @ w["particles"][0]) / L +
             float(r.g_e1 @ h.e_dofs[0]) / L)
    rev_s = kern["reverse_pass"]["avg_us"] * 1e-6 if "reverse_pass" in kern else float("nan")
    achieved = BYTES_REV * n / rev_s / 1e9
    out = {"workload": f"1D2V Weibel adjoint, {n} particles, {steps} steps, checkpoints every "
                       f"{every} steps, loss = dt * sum of 1/2 l2norm_squared(B) per step",
           "particles": n, "steps": steps, "checkpoint_every": every,
           "fwd_particle_steps_per_s": n * steps / t_fwd,
           "fwd_adj_particle_steps_per_s": n * steps / t_adj,
           "fwd_adj_over_fwd": t_adj / t_fwd,
           "loss": r.loss,
           "dloss_dbeta": float(r.g_b @ w["b_pattern"]),
           "dloss_dalpha": float((r.g_particles[0] + gpois) @ w["dx_dalpha"]),
           "dloss_dk": -(L / k) * dj_dl,
           "kernels": kern,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "bytes_per_particle_step": BYTES_REV,
                        "kernel": "k_adj1d reverse pass"}}
    if cpu_n:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc
        wc = workloads.weibel_adjoint_1d2v(cpu_n, seed=4)
        nx = wc["nx"]
        parts = np.asfortranarray(wc["particles"])
        e1, e2, b = np.zeros(nx), np.zeros(nx), wc["b0"].copy()
        osp = orc.Splitting(nx, wc["xmin"], wc["xmax"], 3, 3, 2, parts, e1, e2, b,
                            common_weight=wc["common_weight"])
        osp.solve_poisson(e1, np.zeros(nx))
        t0 = time.perf_counter()
        osp.strang_adjoint(dt, cpu_steps, dt)
        cs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cpu_n * cpu_steps / cs, "unit": "particle-steps/s (fwd+adj)",
                               "cores": 1, "kind": "port",
                               "sample": f"{cpu_n} particles x {cpu_steps} steps ({cs:.2f} s), the "
                                         "oracle's reverse mode (every sub-step state kept), one thread"}
    return out


def ensemble_run(G, _lib, n_inst, steps, seq_steps, adj_steps=500, adj_every=25, seed=5):
    """Independent 1D2V Weibel instances: N uniform in [65536, 1048576],
    nx in {16, 32, 64}, k uniform in [1.0, 1.5], alpha and beta log-uniform
    in [1e-5, 1e-3] (seeded). Particles come from sample!(pg, alpha, k,
    sigma1, mesh) on the GPU, sigma1 = 0.02 / sqrt(2), with v2 then scaled
    to sigma2 = sqrt(12) sigma1 (the Weibel anisotropy); E1 from Poisson,
    B0 = l2projection!(beta cos(kx)). dt = min(0.05, dx/2): the degree-3
    FEM Maxwell solver is unstable for dt > 0.64 dx, which 0.05 exceeds on
    the finest grids. Times `steps` batched Strang steps of the whole
    ensemble against `seq_steps` steps run instance by instance, then
    `adj_steps` steps forward against forward + reverse mode of every
    instance (checkpoints every `adj_every` steps; for 146M particles x 500
    steps at 25 the reverse mode holds about 195 GB: 20 checkpoints, a
    24-state re-run buffer and 20 sort epochs, 1.33 KB per particle)."""
    import math
    from gempic_amd import workloads
    L_ = _lib.load()
    rng = np.random.default_rng(seed)
    hs, dts, total = [], [], 0
    for i in range(n_inst):
        n = int(rng.integers(65536, 1048577))
        nx = int(rng.choice([16, 32, 64]))
        k = float(rng.uniform(1.0, 1.5))
        alpha = float(np.exp(rng.uniform(math.log(1e-5), math.log(1e-3))))
        beta = float(np.exp(rng.uniform(math.log(1e-5), math.log(1e-3))))
        L = 2 * math.pi / k
        mesh = G.OneDGrid(0.0, L, nx)
        pg = G.ParticleGroup(1, 2, n)
        ks0 = G.ParticleMeshCoupling1D(mesh, n, 3, "galerkin")
        ks1 = G.ParticleMeshCoupling1D(mesh, n, 2, "galerkin")
        e1, e2 = np.zeros(nx), np.zeros(nx)
        b = workloads.l2projection_cos(nx, L, 2, beta, k)
        h = G.HamiltonianSplitting(1, 2, G.Maxwell1DFEM(mesh, 3), ks0, ks1, pg, [e1, e2], b,
                                   sync=False)
        G.sample_particles_(h, alpha, k, 0.02 / math.sqrt(2.0))
        h.to_host(fields=False)
        pg.array[2] *= math.sqrt(12.0)
        h.to_device(fields=False)
        G.solve_poisson_(e1, h, np.zeros(nx))
        h.to_device(particles=False)
        hs.append(h)
        dts.append(min(0.05, 0.5 * L / nx))
        total += n
    ens = G.Ensemble(hs)
    # the forward in the reference's arithmetic, then in the fused one (the
    # reverse mode below runs in the reference's arithmetic)
    G.ensemble_strang_splitting_(ens, dts, 2)                # kernel loads
    t0 = time.perf_counter()
    G.ensemble_strang_splitting_(ens, dts, steps)
    t_ref = time.perf_counter() - t0
    for h in hs:
        h.set_arithmetic("fused")
    G.ensemble_strang_splitting_(ens, dts, 2)
    # particles per workgroup of every instance's pass, tuned on this state
    tune = G.ensemble_autotune(ens, dts, 10, particles_per_workgroup=(2048, 4096, 8192, 16384))
    t0 = time.perf_counter()
    G.ensemble_strang_splitting_(ens, dts, steps)
    t_ens = time.perf_counter() - t0
    # the same steps with the regular step replayed as a HIP graph
    G.ensemble_strang_splitting_(ens, dts, 3, graph=True)    # capture
    t0 = time.perf_counter()
    G.ensemble_strang_splitting_(ens, dts, steps, graph=True)
    t_graph = time.perf_counter() - t0
    t0 = time.perf_counter()
    for h, dt in zip(hs, dts):
        _lib.check(L_.gempic_strang_splitting(h._h, dt, seq_steps, _lib.STRANG_FUSED | _lib.ASYNC))
    for h in hs:
        _lib.check(L_.gempic_synchronize(h._h))
    t_seq = time.perf_counter() - t0
    adj = None
    if adj_steps:
        # forward and forward + reverse mode of every instance (loss: dt times the
        # magnetic energy summed over the steps), same number of steps, both in
        # the reference's arithmetic
        for h in hs:
            h.set_arithmetic("reference")
        t0 = time.perf_counter()
        G.ensemble_strang_splitting_(ens, dts, adj_steps)
        t_f = time.perf_counter() - t0
        t0 = time.perf_counter()
        G.ensemble_strang_splitting_adjoint_(ens, dts, adj_steps, checkpoint_every=adj_every,
                                             fetch=False)
        t_a = time.perf_counter() - t0
        adj = {"steps": adj_steps, "checkpoint_every": adj_every,
               "fwd_particle_steps_per_s": total * adj_steps / t_f,
               "fwd_adj_particle_steps_per_s": total * adj_steps / t_a,
               "fwd_adj_over_fwd": t_a / t_f}
    return {"workload": f"{n_inst} independent 1D2V Weibel instances, {total} particles in all, "
                        f"N in [65536, 1048576], nx in {{16, 32, 64}}, k in [1, 1.5], "
                        f"dt = min(0.05, dx/2)",
            "instances": n_inst, "particles": total, "steps": steps,
            "arithmetic": "fused",
            "particle_steps_per_s": total * steps / t_ens,
            "graph_particle_steps_per_s": total * steps / t_graph,
            "reference_arithmetic_particle_steps_per_s": total * steps / t_ref,
            "one_by_one_particle_steps_per_s": total * seq_steps / t_seq,
            "one_by_one_steps": seq_steps, "autotune": tune, "adjoint": adj}


def copy_calibration(mib):
    """Run a plain device copy of `mib` MiB five times (reads and writes mib MiB each).

    Used only under `rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` to scale those
    counters to bytes: the counters of a copy of known size give the factor.
    torch provides the device memory and the elementwise copy kernel.
    """
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)        # one elementwise kernel: n*8 B read, n*8 B written
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "launches": 5,
                      "bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--particles", type=int, default=4194304)
    # sort interval of the fused-path run: 40 steps measured fastest for this
    # workload (1000 steps: 4.07e10 at 10, 4.39e10 at 20, 4.52e10 at 30,
    # 4.59e10 at 40, 4.55e10 at 60); the library's default is 20
    ap.add_argument("--sort-interval", type=int, default=40)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep --blocks and --sort-interval instead of autotuning them "
                         "(HamiltonianSplitting.autotune on the same state, before the warmup)")
    ap.add_argument("--cpu-particles", type=int, default=262144)
    ap.add_argument("--cpu-steps", type=int, default=300)      # about 10 s on one core
    ap.add_argument("--cpu-steps-mt", type=int, default=100,
                    help="steps of the all-threads CPU baseline on the benchmark's particles")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--substep", action="store_true", help="also time the sub-step path")
    ap.add_argument("--blocks", type=int, default=0, help="workgroups per particle pass (0 = auto)")
    ap.add_argument("--fusion", type=int, default=1, help="particle kernels per fused step (1 or 2)")
    ap.add_argument("--arith", choices=("fused", "reference"), default="fused",
                    help="arithmetic of the particle pass: fused (closed-form bases and fma, "
                         "parity with the reference to 1e-12 per sub-step) or reference (the "
                         "reference's rounding, bitwise per particle)")
    ap.add_argument("--hbm-particles", type=int, default=16777216,
                    help="particles of the state-larger-than-the-caches point (0 = skip it)")
    ap.add_argument("--hbm-steps", type=int, default=200)
    ap.add_argument("--particles-2d3v", type=int, default=16777216,
                    help="particles of the 2D3V Landau measurement (0 = skip it)")
    ap.add_argument("--steps-2d3v", type=int, default=200)
    ap.add_argument("--no-sampling", action="store_true", help="skip the sample! measurement")
    ap.add_argument("--boris-steps", type=int, default=200,
                    help="steps of the Boris splitting measurement (0 = skip it)")
    ap.add_argument("--adjoint-steps", type=int, default=1000,
                    help="steps of the forward + reverse-mode measurements (0 = skip them)")
    ap.add_argument("--adjoint-every", type=int, default=50, help="checkpoint interval in steps")
    ap.add_argument("--ensemble", type=int, default=256,
                    help="instances of the ensemble measurement (0 = skip it)")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--particles-2d", type=int, default=16777216,
                    help="particles of the 2D pieces measurement (0 = skip it)")
    ap.add_argument("--reps-2d", type=int, default=20)
    ap.add_argument("--only-2d", action="store_true",
                    help="print only the 2D pieces measurement (for profiling its kernels)")
    ap.add_argument("--only-2d3v", action="store_true",
                    help="print only the 2D3V Landau measurement (for profiling its kernels)")
    ap.add_argument("--cpu-particles-2d", type=int, default=4194304)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(particles would shard with one all-reduce of the nx-sized deposits per "
              "step). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.only_2d:
        print(json.dumps(pieces_2d(args.particles_2d, args.reps_2d,
                                   0 if args.no_cpu else args.cpu_particles_2d)))
        return

    import gempic_amd as G
    from gempic_amd import workloads
    from gempic_amd import _lib

    if args.only_2d3v:
        print(json.dumps(split2d_run(G, args.particles_2d3v, args.steps_2d3v,
                                     0 if args.no_cpu else 131072, 4)))
        return

    n = args.particles
    print(f"[bench] generating 1D2V Weibel, {n} particles", file=sys.stderr, flush=True)
    w = workloads.weibel_1d2v(n, seed=1)
    h = make_splitting(G, w)
    L = _lib.load()
    _lib.check(L.gempic_set_sort_interval(h._h, args.sort_interval))
    _lib.check(L.gempic_set_launch(h._h, args.blocks))
    _lib.check(L.gempic_set_fusion(h._h, args.fusion))
    h.set_arithmetic(args.arith)
    dt = w["dt"]
    tune = None
    blocks, interval = args.blocks, args.sort_interval
    if not args.no_autotune:
        # every pair runs 120 steps twice from the same state, which is
        # restored after each trial; the fastest pair stays set
        print("[bench] autotuning workgroups per pass and sort interval", file=sys.stderr,
              flush=True)
        t0 = time.perf_counter()
        tune = h.autotune(dt, n_steps=120, blocks=(0, 768, 1024, 1536),
                          sort_intervals=(20, 30, 40, 60))
        tune["seconds"] = time.perf_counter() - t0
        blocks, interval = tune["blocks"], tune["sort_interval"]
        print(f"[bench] autotune: blocks {blocks}, sort interval {interval} "
              f"({tune['seconds']:.2f} s)", file=sys.stderr, flush=True)
    print(f"[bench] warmup {args.warmup} steps", file=sys.stderr, flush=True)
    if args.warmup:
        _lib.check(L.gempic_strang_splitting(h._h, dt, args.warmup, _lib.STRANG_FUSED))
    _lib.check(L.gempic_synchronize(h._h))
    # the particle sort, timed where it runs: HIP events on the library's
    # stream around the sorts of three sort intervals of steps, each sorting
    # particles that drifted for a whole interval; then one more sort, so
    # that the timed window starts right after a sort
    _lib.check(L.gempic_profile_enable(h._h, 1))
    _lib.check(L.gempic_profile_reset(h._h))
    if interval > 0:
        _lib.check(L.gempic_strang_splitting(h._h, dt, 3 * interval + 1, _lib.STRANG_FUSED))
    sort_ms_tot, sort_cnt = h.profile_read(6)
    if not sort_cnt:
        for _ in range(3):
            _lib.check(L.gempic_sort_particles(h._h))
        sort_ms_tot, sort_cnt = h.profile_read(6)
    sort_s = 1e-3 * sort_ms_tot / max(sort_cnt, 1)
    _lib.check(L.gempic_sort_particles(h._h))
    _lib.check(L.gempic_synchronize(h._h))
    # The timed window starts right after a sort: steps s = interval,
    # 2 interval, ... inside it sort, floor((K - 1) / interval) of them. The
    # figures below charge every step its share of a sort instead
    # (sort time / interval), whether or not one falls in the window.
    sorts_in_window = (args.steps - 1) // interval if interval > 0 else 0
    sort_share_s = (args.steps / interval - sorts_in_window) * sort_s if interval > 0 else 0.0
    # HIP events around every 10th launch of each kernel (for runs of fewer
    # than 50 steps, every (steps / 5)-th, so that at least 5 launches are
    # timed): the kernel times come from the timed run itself, at the cost of
    # a few microseconds on the sampled steps only
    _lib.check(L.gempic_profile_enable(h._h, max(1, min(10, args.steps // 5))))
    _lib.check(L.gempic_profile_reset(h._h))
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    _lib.check(L.gempic_strang_splitting(h._h, dt, args.steps, _lib.STRANG_FUSED))
    _lib.check(L.gempic_synchronize(h._h))
    elapsed_window = time.perf_counter() - t0
    elapsed = elapsed_window + sort_share_s
    fused_ms, fused_cnt = h.profile_read(0)
    kernels = {}
    for kid, name in enumerate(["fused_pass", "trailing_pass", "he", "hp2", "hp1", "grid",
                                "sort", "other", "fused_pass_a", "deposit_sum_and_solves"]):
        ms, cnt = h.profile_read(kid)
        if cnt:
            kernels[name] = {"avg_us": 1e3 * ms / cnt, "count": cnt}
    h.profile(False)
    value = n * args.steps / elapsed
    avg_s = (fused_ms / fused_cnt) * 1e-3 if fused_cnt else float("nan")
    achieved = BYTES_FUSED * n / avg_s / 1e9
    print(f"[bench] {value:.4e} particle-steps/s, fused pass {avg_s * 1e6:.2f} us",
          file=sys.stderr, flush=True)

    # the same timed run in the reference's arithmetic, on a fresh copy of the state
    ref_arith = None
    if args.arith != "reference":
        print("[bench] the same run in the reference's arithmetic", file=sys.stderr, flush=True)
        ref_arith = arith_run(G, _lib, w, "reference", args.warmup, args.steps, blocks, interval,
                              sort_s)

    hbm = None
    if args.hbm_particles:
        print(f"[bench] {args.hbm_particles} particles (state larger than the caches)",
              file=sys.stderr, flush=True)
        wh = workloads.weibel_1d2v(args.hbm_particles, seed=1)
        hbm = arith_run(G, _lib, wh, args.arith, min(args.warmup, 20), args.hbm_steps, 0,
                        interval, None)
        hbm["workload"] = (f"1D2V Weibel, {args.hbm_particles} particles (the particle state, "
                           f"{args.hbm_particles * 40 / 2**30:.2f} GiB with the sort buffers' "
                           f"halves, exceeds the 256 MB last-level cache)")
        del wh

    substep = None
    if args.substep:
        _lib.check(L.gempic_synchronize(h._h))
        t0 = time.perf_counter()
        ks = max(args.steps // 10, 1)
        _lib.check(L.gempic_strang_splitting(h._h, dt, ks, _lib.STRANG_SUBSTEP))
        _lib.check(L.gempic_synchronize(h._h))
        substep = n * ks / (time.perf_counter() - t0)

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "fused_pass_weibel_4m_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        print(f"[bench] CPU baseline: oracle, {args.cpu_particles} particles x "
              f"{args.cpu_steps} steps", file=sys.stderr, flush=True)
        rate1, secs1 = cpu_baseline(args.cpu_particles, args.cpu_steps)
        one = {"value": rate1, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"1D2V Weibel, {args.cpu_particles} particles x {args.cpu_steps} Strang "
                         f"steps ({secs1:.1f} s), C restatement of the reference, one thread"}
        tasks = cpu_threads()
        while tasks > 1 and n % tasks:
            tasks -= 1
        if tasks > 1:
            print(f"[bench] CPU baseline: oracle on {tasks} threads, {n} particles x "
                  f"{args.cpu_steps_mt} steps", file=sys.stderr, flush=True)
            rate, secs = cpu_baseline(n, args.cpu_steps_mt, tasks=tasks)
            cpu = {"value": rate, "unit": "particle-steps/s", "cores": tasks, "kind": "port",
                   "sample": f"1D2V Weibel, the benchmark's {n} particles x {args.cpu_steps_mt} "
                             f"Strang steps ({secs:.1f} s), C restatement of the reference with "
                             f"its nthreads() task split on {tasks} threads",
                   "one_thread": one}
        else:
            cpu = one

    boris = None
    if args.boris_steps:
        print(f"[bench] Boris splitting, {args.boris_steps} steps", file=sys.stderr, flush=True)
        boris = boris_run(G, _lib, w, args.boris_steps, 10)

    adj = None
    if args.adjoint_steps:
        # the benchmark's particle count with every step's state kept (one-step
        # checkpoint intervals, ~100 GB at 4M particles x 1000 steps) and with
        # checkpoints every `--adjoint-every` steps; the 1M-particle case with
        # checkpoints every `--adjoint-every` steps
        adj = {}
        for na, every in ((n, 1), (n, args.adjoint_every), (1048576, args.adjoint_every)):
            print(f"[bench] forward and reverse mode, {na} particles x {args.adjoint_steps} steps, "
                  f"checkpoints every {every}", file=sys.stderr, flush=True)
            adj[f"{na}_every{every}"] = adjoint_run(
                G, _lib, na, args.adjoint_steps, every,
                0 if (args.no_cpu or na != n or every != 1) else 32768, 50)

    ens = None
    if args.ensemble:
        print(f"[bench] ensemble of {args.ensemble} instances, {args.ensemble_steps} steps",
              file=sys.stderr, flush=True)
        ens = ensemble_run(G, _lib, args.ensemble, args.ensemble_steps,
                           max(1, args.ensemble_steps // 5))

    samp = None
    if not args.no_sampling:
        print("[bench] sample! into GPU memory", file=sys.stderr, flush=True)
        samp = sampling_run(G, n, 0 if args.no_cpu else 262144)

    s2d = None
    if args.particles_2d3v:
        s2d = split2d_run(G, args.particles_2d3v, args.steps_2d3v,
                          0 if args.no_cpu else 131072, 4)

    p2d = None
    if args.particles_2d:
        print(f"[bench] 2D pieces, {args.particles_2d} particles", file=sys.stderr, flush=True)
        p2d = pieces_2d(args.particles_2d, args.reps_2d, 0 if args.no_cpu else args.cpu_particles_2d)

    line = {
        "metric": "particle-steps/s (1D2V, 4M particles, FP64) fwd & fwd+adjoint; % HBM roofline",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Weibel initial condition)",
        "config": {"workload": "1D2V Weibel instability", "particles": n, "nx": w["nx"],
                   "spline_degree": w["degree"], "dt": dt, "sort_interval": interval,
                   "blocks": blocks, "kernels_per_step": args.fusion,
                   "arithmetic": args.arith, "path": "fused Strang step"},
        "sort": {"us_per_sort": 1e6 * sort_s, "interval": interval,
                 "sorts_in_timed_window": sorts_in_window,
                 "share_added_us_per_step": 1e6 * sort_share_s / args.steps,
                 "window_ms_per_step": 1e3 * elapsed_window / args.steps,
                 "timed_on": f"{sort_cnt} sorts of particles that drifted for {interval} steps",
                 "note": "value and ms_per_step charge each step sort time / interval"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_pass1d fused Strang pass",
                     "bytes_per_particle_step": BYTES_FUSED},
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    if tune is not None:
        line["autotune"] = tune
    if boris is not None:
        line["boris_1d2v"] = boris
    if adj is not None:
        line["adjoint_1d2v"] = adj
    if ens is not None:
        line["ensemble_1d2v"] = ens
    if samp is not None:
        line["sampling"] = samp
    if p2d is not None:
        line["pieces_2d"] = p2d
    if substep is not None:
        line["substep_path_value"] = substep
    if ref_arith is not None:
        line["reference_arithmetic"] = ref_arith
    if hbm is not None:
        line["hbm_16m"] = hbm
    if s2d is not None:
        line["landau_2d3v"] = s2d
    print(json.dumps(line))


if __name__ == "__main__":
    main()
