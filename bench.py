"""Benchmark: FP64 cell-updates/s of the HIP shallow-water step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the 4096x4096 Cartesian grid (dx = 10 m, seeded beach + 64 bumps +
noise, west Flather M2 tide, Manning 0.025, f = 1e-4, CFL 0.45), state resident
in HBM, self-timed dt (maxDt fused into every step), one step = both SSP-RK2
stages + friction + CFL reduction, replayed from a HIP graph.  Prints one JSON
line (rank 0).  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "water-simulator_amd"))

METRIC = "FP64 cell-updates/s + % HBM peak; batched vs plain; adjoint/forward time"
HBM_PEAK_GBS = 8000.0
# FP64 VALU issue: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per 4 cycles
VALU_PEAK_WAVE_INSTR_S = 256 * 4 * 2.4e9 / 4
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(seconds=12.0):
    """The CPU oracle (single-threaded C restatement of the reference step) on a
    256x256 sample of the same workload recipe, timed on this host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import swe_amd
    import workloads
    st, p = workloads.large_grid(nx=256, ny=256, dx=p_dx())
    rows = swe_amd.geometry_cartesian(256, 256, p["dx"], coriolis=p["coriolis"])
    o = oracle.OracleSim(256, 256, rows, manning=p["manning"], coriolis=p["coriolis"], cfl=p["cfl"],
                         dx=p["dx"], dy=p["dx"])
    o.load_state(*st)
    o.set_bc(p["bc"])
    ext = [1.0, 0, 0, 0, 0, 0, 0, 0]
    o.step(None, ext)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        o.run(5, ext)
        n += 5
    el = time.perf_counter() - t0
    out = dict(value=256 * 256 * n / el, unit="cell-updates/s", cores=1, kind="port",
               sample=f"oracle C restatement, 256x256 sample of the 4096x4096 recipe, {n} self-timed steps in {el:.1f} s")
    ref = committed_figure("reference_js_cell_updates_per_s", "large_1024", "profiles/reference_js_large_grid.md")
    if ref:
        out["reference_js"] = {"value": ref, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                               "sample": "src/swe.mjs under Node v12.22.9, 1024x1024 sample of the same recipe, "
                                         "5 self-timed steps, best of 3",
                               "host": "build container, Intel Xeon @ 2.70 GHz (not the GPU host)",
                               "source": "profiles/reference_js_large_grid.md (committed, not measured in this run)"}
    return out


PROFILE = "profiles/k_sweep_large_grid_4096.md"


def committed_figure(name, kernel, path=PROFILE):
    """A figure of a committed rocprofv3 (or timing) summary, or None: for the
    4096x4096 workload at 94 rows per chunk traffic_bytes_per_launch (FETCH_SIZE and
    WRITE_SIZE, each scaled by a plain-copy calibration),
    valu_wave_instructions_per_launch (SQ_INSTS_VALU) and
    fp64_wave_instructions_per_launch (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64)."""
    path = os.path.join(ROOT, path)
    if not os.path.exists(path):
        return None
    key = f"{name} {kernel}:"
    with open(path) as f:
        for line in f:
            if line.startswith(key):
                return float(line.split(":", 1)[1])
    return None


def committed_traffic(kernel):
    return committed_figure("traffic_bytes_per_launch", kernel)


def p_dx():
    return 10.0


def batched_measure(steps, warmup, seq_steps):
    """Config 3: 512 uneven basins in one batch vs the same basins as sequential
    plain runs on the same GPU (device time of the stepping only).  The batch is
    timed over `steps` steps (by default the configuration's 3000: the forced
    resonators fill and friction engages in more cells as the run goes on, so a
    short window from rest reads about 3% high); the sequential runs over a
    sample of seq_steps steps per basin from rest."""
    import swe_amd
    import workloads
    specs = workloads.ensemble(nb=512)
    cells = sum(s["nx"] * s["ny"] for s in specs)
    B = swe_amd.Batch([dict(nx=s["nx"], ny=s["ny"], manning=s["manning"], cfl=s["cfl"],
                            rows=swe_amd.geometry_cartesian(s["nx"], s["ny"], s["dx"])) for s in specs])
    for i, s in enumerate(specs):
        B.load_state(i, *s["state"])
        B.set_bc(i, s["bc"])
        B.set_harmonic(i, 0, amp=[s["amp"]], omega=[s["omega"]], phase=[0.0],
                       uamp=[s["uamp"]], uomega=[s["omega"]], uphase=[0.0])
    rows, waves = B.autotune(steps=2)
    log(f"[bench] batch autotune: {rows} rows per task, {waves} persistent wavefronts")
    B.advance(warmup)
    t0 = time.perf_counter()
    n = B.advance(steps)
    el = time.perf_counter() - t0
    batched = cells * steps / el
    log(f"[bench] batched: {batched:.4e} cell-updates/s over {len(specs)} basins ({cells / 1e6:.1f} M cells)")
    B.close()
    seq_ms = 0.0
    for s in specs:
        g = swe_amd.Solver(s["nx"], s["ny"], swe_amd.geometry_cartesian(s["nx"], s["ny"], s["dx"]),
                           manning=s["manning"], cfl=s["cfl"])
        g.load_state(*s["state"])
        g.set_bc(s["bc"])
        g.set_harmonic(0, amp=[s["amp"]], omega=[s["omega"]], phase=[0.0],
                       uamp=[s["uamp"]], uomega=[s["omega"]], uphase=[0.0])
        g.use_graph(True)
        g.advance(2)
        g.advance(seq_steps)
        seq_ms += g.last_advance_ms()
        g.close()
    seq = cells * seq_steps / (seq_ms * 1e-3)
    log(f"[bench] sequential plain: {seq:.4e} cell-updates/s")
    return {"value": batched, "unit": "cell-updates/s", "basins": len(specs), "cells": cells, "steps": steps,
            "task_rows": rows, "persistent_wavefronts": waves, "sequential_plain": seq,
            "sequential_plain_steps": seq_steps, "batched_vs_plain": batched / seq}


def adjoint_measure(nsteps, every=10, fd_dirs=0):
    """Config 4: gradient of a gauge misfit w.r.t. the initial surface and the per-cell
    Manning n on the 1024x1024 shelf (32 gauges every 10 steps), with sqrt(N)
    checkpointing.  Observations come from a truth run (8 surface humps, Manning
    0.025 x (1 + 0.3 sin)); the gradient is taken at half the truth surface and the
    baseline n = 0.025.  Reports device ms of the gradient (recompute + reverse)
    against one plain forward of the same steps."""
    import numpy as np
    import swe_amd
    import workloads
    b, ht, p, gauges = workloads.inversion()
    nx, ny, dx = p["nx"], p["ny"], p["dx"]
    W, H = nx + 4, ny + 4
    rows = swe_amd.geometry_cartesian(nx, ny, dx, coriolis=p["coriolis"])
    s = swe_amd.Solver(nx, ny, rows, manning=p["manning"], coriolis=p["coriolis"], cfl=p["cfl"], dx=dx, dy=dx)
    z = np.zeros_like(ht)
    # first guess: half the truth surface plus 5 cm of smooth seeded noise (the
    # noise keeps the surface differences away from the limiters' kinks, which
    # finite differences need)
    h_eval = workloads.inversion_first_guess(b, ht, nx, ny)
    X, Y = np.meshgrid(np.arange(W), np.arange(H))
    n_truth = (0.025 * (1 + 0.3 * np.sin(X / 97.0) * np.cos(Y / 131.0))).ravel()
    n0 = np.full(W * H, 0.025)
    s.load_state(b, h_eval, z, z)
    s.set_bc(p["bc"])
    s.set_manning_field(n0)
    chunk = s.autotune()
    s.use_graph(True)
    s.advance(nsteps)
    plain_ms = s.last_advance_ms()
    adj = swe_amd.Adjoint(s, nsteps)
    nsamp = nsteps // every
    adj.set_gauges(gauges, every, np.zeros((nsamp, len(gauges))))
    adj.forward(ht, n_truth)
    adj.set_gauges(gauges, every, adj.series())
    J = adj.forward(h_eval, n0)
    gh, gn = adj.backward()
    f_ms, b_ms = adj.timing()
    log(f"[bench] adjoint: plain forward {plain_ms:.1f} ms, recorded forward {f_ms:.1f} ms, gradient {b_ms:.1f} ms "
        f"({b_ms / plain_ms:.2f}x), J = {J:.4e}")
    out = {"workload": "inversion_1024x1024_dx50_shelf_seamounts_wetdry_east", "steps": nsteps,
           "gauges": int(len(gauges)), "sample_every": every, "checkpoint_interval": int(round(math.sqrt(nsteps))),
           "chunk_rows": chunk, "loss": J, "plain_forward_ms": plain_ms, "recorded_forward_ms": f_ms,
           "gradient_ms": b_ms, "gradient_over_forward": b_ms / plain_ms,
           "grad_h0_max": float(np.abs(gh).max()), "grad_n_max": float(np.abs(gn).max())}
    if fd_dirs > 0:
        # seeded random directions of the control space (surface humps of 1 km to
        # 5 km centred within 2 km of a gauge, and low Fourier modes of n, on the
        # wet cells), central differences at steps of 1e-3 to 1e-5 of them; a
        # direction agrees when two successive steps' differences are within 1e-6
        # of each other (step-converged) and the smaller step's is within 1e-6 of
        # the adjoint
        wet = h_eval > 1.0
        steps = (1e-3, 1e-4, 1e-5)
        rel = {e: [] for e in steps}
        agree = 0
        for q, (dh, dn) in enumerate(workloads.inversion_directions(fd_dirs, nx, ny, dx, wet, seed=46, near=gauges)):
            ad = float(np.dot(gh, dh) + np.dot(gn, dn))
            fd = [(adj.loss(h_eval + e * dh, n0 + e * dn) - adj.loss(h_eval - e * dh, n0 - e * dn)) / (2 * e)
                  for e in steps]
            for e, f in zip(steps, fd):
                rel[e].append(abs(f - ad) / abs(ad))
            agree += any(abs(fd[k] - fd[k + 1]) <= 1e-6 * abs(fd[k + 1]) and abs(fd[k + 1] - ad) <= 1e-6 * abs(ad)
                         for k in range(len(steps) - 1))
            if (q + 1) % 4 == 0:
                log(f"[bench] adjoint FD: {q + 1} of {fd_dirs} directions, {agree} agree so far")
        out["fd_directions"] = fd_dirs
        out["fd_steps"] = list(steps)
        out["fd_rel_err_median"] = {f"{e:g}": float(np.median(rel[e])) for e in steps}
        out["fd_rel_err_max"] = {f"{e:g}": float(max(rel[e])) for e in steps}
        out["fd_agree_1e-6"] = int(agree)
        log(f"[bench] adjoint FD over {fd_dirs} directions: {agree} agree to 1e-6 (step-converged); median rel err "
            + ", ".join(f"{e:g}: {np.median(rel[e]):.2e}" for e in steps))
    adj.close()
    s.close()
    return out


def inversion_measure(iters, nsteps=4000, every=10):
    """Config 4 as an inversion: scipy's L-BFGS-B on the initial surface of the wet
    cells, from a flat sea plus 1 cm of noise, driven by the adjoint's forward/backward pair (one
    recorded forward for the misfit and one backward for dJ/d eta0 per
    evaluation).  Reports the misfit and the distance to the hidden truth."""
    import numpy as np
    from scipy.optimize import minimize
    import swe_amd
    import workloads
    b, ht, p, gauges = workloads.inversion()
    nx, ny, dx = p["nx"], p["ny"], p["dx"]
    W, H = nx + 4, ny + 4
    rows = swe_amd.geometry_cartesian(nx, ny, dx, coriolis=p["coriolis"])
    s = swe_amd.Solver(nx, ny, rows, manning=p["manning"], coriolis=p["coriolis"], cfl=p["cfl"], dx=dx, dy=dx)
    z = np.zeros_like(ht)
    s.load_state(b, ht, z, z)
    s.set_bc(p["bc"])
    s.autotune()
    n0 = np.full(W * H, 0.025)
    adj = swe_amd.Adjoint(s, nsteps)
    adj.set_gauges(gauges, every, np.zeros((nsteps // every, len(gauges))))
    adj.forward(ht, n0)
    adj.set_gauges(gauges, every, adj.series())
    inner = np.zeros((H, W), bool)
    inner[2:-2, 2:-2] = True
    wet = inner.ravel() & (b < -1.0)
    eta_truth = (b + ht)[wet]

    def h_of(e):
        h = np.maximum(0.0, -b)
        h[wet] = np.maximum(0.0, e - b[wet])
        return h

    hist = []
    t0 = time.perf_counter()

    def f(e):
        J = adj.forward(h_of(e), n0)
        gh, _ = adj.backward()
        hist.append(J)
        log(f"[bench] inversion evaluation {len(hist)}: J = {J:.4e} ({time.perf_counter() - t0:.0f} s)")
        return J, gh[wet]

    # a flat sea puts every limiter argument exactly on its kink (all surface
    # differences zero), where the subgradient is no descent direction: start
    # 1 cm of seeded noise away from it
    e0 = 0.01 * workloads.SplitMix64(47).normals(int(wet.sum()))
    r = minimize(f, e0, jac=True, method="L-BFGS-B", options=dict(maxiter=iters))
    el = time.perf_counter() - t0
    out = {"optimizer": "L-BFGS-B", "iterations": int(r.nit), "evaluations": len(hist), "steps": nsteps,
           "misfit_start": hist[0], "misfit_end": float(r.fun),
           "truth_error_start": float(np.linalg.norm(e0 - eta_truth)),
           "truth_error_end": float(np.linalg.norm(r.x - eta_truth)), "seconds": el}
    adj.close()
    s.close()
    return out


def sphere_measure(steps, adj_steps, every=10):
    """Config 5: 16-member global-ocean ensemble on the 2160x1080 sphere with the
    M2/S2/K1 equilibrium tide, one batched launch per phase; throughput over a
    fixed step count, the 5-day extrapolation, the same member run plain, and the
    gradient of a gauge misfit (64 open-ocean points) w.r.t. the per-cell Manning n
    of member 0 over adj_steps steps."""
    import numpy as np
    import swe_amd
    import workloads
    b, h, p, mann = workloads.sphere_tides()
    nx, ny = p["nx"], p["ny"]
    W, H = nx + 4, ny + 4
    rows = swe_amd.geometry_spherical(nx, ny, R=p["R"], lat0=p["lat0"], lat1=p["lat1"], omega=p["omega"])
    lat, lon = swe_amd.sphere_lonlat(nx, ny, p["lat0"], p["lat1"])
    z = np.zeros_like(h)
    B = swe_amd.Batch([dict(nx=nx, ny=ny, rows=rows, sphere=True, R=p["R"], manning=n, cfl=p["cfl"]) for n in mann])
    for q in range(len(mann)):
        B.load_state(q, b, h, z, z)
        B.set_tide(q, p["tides"], lat, lon)
    task_rows, waves = B.autotune(steps=2)
    B.advance(2)
    t0 = [B.get_scalar(q, swe_amd.S_T) for q in range(len(mann))]
    B.advance(steps)
    ms = B.last_advance_ms()
    dts = [(B.get_scalar(q, swe_amd.S_T) - t0[q]) / steps for q in range(len(mann))]
    finite = all(B.get_scalar(q, swe_amd.S_NONFINITE) == 0 for q in range(len(mann)))
    B.close()
    cells = len(mann) * nx * ny
    value = cells * steps / (ms * 1e-3)
    steps_5d = p["days"] * 86400.0 / min(dts)
    log(f"[bench] sphere ensemble: {value:.4e} cell-updates/s ({len(mann)} x {nx}x{ny}), dt ~ {min(dts):.4f} s, "
        f"5 days = {steps_5d:.3e} steps = {steps_5d * ms / steps / 3.6e6:.2f} h")
    # the same member as a plain run
    s = swe_amd.Solver(nx, ny, rows, sphere=True, R=p["R"], manning=mann[0], cfl=p["cfl"])
    s.load_state(b, h, z, z)
    s.set_tide(p["tides"], lat, lon)
    s.autotune()
    s.use_graph(True)
    s.advance(2)
    s.advance(steps)
    plain_ms = s.last_advance_ms()
    plain = nx * ny * steps / (plain_ms * 1e-3)
    out = {"workload": "sphere_2160x1080_global_ocean_k-2_bathymetry_M2_S2_K1_equilibrium_tide", "members": len(mann),
           "steps": steps, "value": value, "unit": "cell-updates/s", "task_rows": task_rows,
           "persistent_wavefronts": waves, "dt_min": min(dts), "steps_for_5_days": steps_5d,
           "hours_for_5_days_extrapolated": steps_5d * ms / steps / 3.6e6, "plain_member": plain,
           "finite": finite}
    if adj_steps > 0:
        rng = workloads.SplitMix64(51)
        gauges = []
        LA = np.repeat(lat, W)
        while len(gauges) < 64:
            k = (2 + int(rng.uniform(0, ny))) * W + 2 + int(rng.uniform(0, nx))
            if h[k] > 1000.0 and abs(LA[k]) < 1.2 and k not in gauges:
                gauges.append(k)
        gauges = np.array(gauges, dtype=np.int32)
        # a surface disturbance to invert for: a 1 m large-scale pattern (truth)
        # observed from a first guess of half of it, with n = 0.025 (truth)
        # against member 0's n
        LO = np.tile(lon, H)
        eta_t = 1.0 * np.cos(LA) ** 2 * np.cos(3 * LO) * (h > 0)
        h_t = np.where(h > 0, h + eta_t, 0.0)
        h_g = np.where(h > 0, h + 0.5 * eta_t, 0.0)
        n_truth = np.full(W * H, 0.025)
        n0 = np.full(W * H, mann[0])
        adj = swe_amd.Adjoint(s, adj_steps)
        ns = adj_steps // every
        adj.set_gauges(gauges, every, np.zeros((ns, len(gauges))))
        adj.forward(h_t, n_truth)
        adj.set_gauges(gauges, every, adj.series())
        J = adj.forward(h_g, n0)
        gh, gn = adj.backward()
        f_ms, b_ms = adj.timing()
        s.load_state(b, h, z, z)
        s.set_manning_field(n0)
        s.advance(adj_steps)
        pf_ms = s.last_advance_ms()
        out["adjoint"] = {"steps": adj_steps, "gauges": 64, "sample_every": every, "loss": J,
                          "plain_forward_ms": pf_ms, "gradient_ms": b_ms, "gradient_over_forward": b_ms / pf_ms,
                          "grad_h0_max": float(np.abs(gh).max()), "grad_n_max": float(np.abs(gn).max())}
        log(f"[bench] sphere adjoint (member 0, {adj_steps} steps): gradient {b_ms:.1f} ms = {b_ms / pf_ms:.2f}x forward")
        # the 24-hour misfit of one member: memory of one- and two-level checkpointing
        n24 = int(math.ceil(86400.0 / out["dt_min"]))
        ncell = (nx + 4) * (ny + 4)
        k1 = max(1, round(math.sqrt(n24)))
        k2 = max(1, round(n24 ** (1.0 / 3.0)))
        out["adjoint"]["plan_24h"] = {
            "steps": n24, "one_level": {"inner": k1, "outer": k1,
                                        "GB": swe_amd.adjoint_plan_bytes(n24, ncell, k1, k1) / 1e9},
            "two_levels": {"inner": k2, "outer": k2 * k2,
                           "GB": swe_amd.adjoint_plan_bytes(n24, ncell, k2, k2 * k2) / 1e9}}
        adj.close()
    s.close()
    return out


def sphere_adjoint_long(hours, fd_dirs, every=10, steps=(1e-1, 1e-2)):
    """Config 5's adjoint over a horizon where friction acts: member 0 of the
    2160x1080 global ocean under the M2/S2/K1 equilibrium tide, started from a
    1 m large-scale surface disturbance, the misfit at 64 open-ocean gauges
    sampled every 10 steps over `hours` of simulated time.  Observations come
    from a truth run with n = 0.025 everywhere, the gradient is taken at member
    0's own n, from the same initial state: a friction-only calibration.  dJ/dn
    by the adjoint (checkpointing chosen by the library: two levels at this
    length), checked by central differences along fd_dirs seeded smooth
    directions of n (1e-3 times 0.5 + 0.5 cos(k lon + a) cos(l lat) on the
    ocean cells) at the given steps of them; the gradient's cost against one
    plain forward of the same steps."""
    import numpy as np
    import swe_amd
    import workloads
    b, h, p, mann = workloads.sphere_tides()
    nx, ny = p["nx"], p["ny"]
    W, H = nx + 4, ny + 4
    rows = swe_amd.geometry_spherical(nx, ny, R=p["R"], lat0=p["lat0"], lat1=p["lat1"], omega=p["omega"])
    lat, lon = swe_amd.sphere_lonlat(nx, ny, p["lat0"], p["lat1"])
    z = np.zeros_like(h)
    s = swe_amd.Solver(nx, ny, rows, sphere=True, R=p["R"], manning=mann[0], cfl=p["cfl"])
    s.set_tide(p["tides"], lat, lon)
    LA, LO = np.repeat(lat, W), np.tile(lon, H)
    h_t = np.where(h > 0, h + 1.0 * np.cos(LA) ** 2 * np.cos(3 * LO), 0.0)
    s.load_state(b, h_t, z, z)
    s.autotune()
    s.advance(20)
    dt = s.t / 20
    nsteps = int(math.ceil(hours * 3600.0 / dt))
    rng = workloads.SplitMix64(51)
    gauges = []
    while len(gauges) < 64:
        k = (2 + int(rng.uniform(0, ny))) * W + 2 + int(rng.uniform(0, nx))
        if h[k] > 1000.0 and abs(LA[k]) < 1.2 and k not in gauges:
            gauges.append(k)
    gauges = np.array(gauges, dtype=np.int32)
    n_truth = np.full(W * H, 0.025)
    n0 = np.full(W * H, mann[0])
    log(f"[bench] sphere adjoint over {hours:g} h: {nsteps} steps of ~{dt:.4f} s")
    adj = swe_amd.Adjoint(s, nsteps)
    ns = nsteps // every
    adj.set_gauges(gauges, every, np.zeros((ns, len(gauges))))
    adj.forward(h_t, n_truth)
    log("[bench] sphere adjoint: truth run done")
    adj.set_gauges(gauges, every, adj.series())
    J = adj.forward(h_t, n0)
    log(f"[bench] sphere adjoint: recorded forward done, J = {J:.6e}")
    gh, gn = adj.backward()
    f_ms, b_ms = adj.timing()
    log(f"[bench] sphere adjoint: J = {J:.6e}, |dJ/dn|max = {np.abs(gn).max():.3e}, gradient {b_ms / 1e3:.1f} s")
    s.load_state(b, h_t, z, z)
    s.set_manning_field(n0)
    s.use_graph(True)
    s.advance(nsteps)
    pf_ms = s.last_advance_ms()
    out = {"workload": "sphere_2160x1080_member0_friction_calibration_M2_S2_K1", "hours": hours, "steps": nsteps,
           "gauges": len(gauges), "sample_every": every, "loss": J, "plain_forward_ms": pf_ms,
           "recorded_forward_ms": f_ms, "gradient_ms": b_ms, "gradient_over_forward": b_ms / pf_ms,
           "grad_n_max": float(np.abs(gn).max()), "grad_n_l2": float(np.linalg.norm(gn))}
    if fd_dirs > 0:
        ocean = (h > 0).reshape(H, W)
        ocean[:2] = ocean[-2:] = False
        ocean[:, :2] = ocean[:, -2:] = False
        r2 = workloads.SplitMix64(52)
        rows_out = []
        for q in range(fd_dirs):
            k, l, a = int(r2.uniform(1, 7)), int(r2.uniform(1, 5)), r2.uniform(0, 2 * math.pi)
            dn = (1e-3 * (0.5 + 0.5 * np.cos(k * lon[None, :] + a) * np.cos(l * lat[:, None])) * ocean).ravel()
            ad = float(np.dot(gn, dn))
            fd = [(adj.loss(h_t, n0 + e * dn) - adj.loss(h_t, n0 - e * dn)) / (2 * e) for e in steps]
            rel = [abs(f - ad) / abs(ad) for f in fd]
            row = {"k": k, "l": l, "adjoint": ad, "fd": fd, "rel_err": rel}
            if len(fd) > 1:
                row["step_converged"] = abs(fd[0] - fd[1]) <= 1e-6 * abs(fd[1])
            rows_out.append(row)
            log(f"[bench] sphere adjoint FD {q + 1}/{fd_dirs}: adjoint {ad:+.6e}, rel err "
                + ", ".join(f"{e:g}: {r:.2e}" for e, r in zip(steps, rel)))
        out["fd_steps"] = list(steps)
        out["fd"] = rows_out
        out["fd_within_1e-6"] = int(sum(min(r["rel_err"]) <= 1e-6 for r in rows_out))
    adj.close()
    s.close()
    return out


def diagnostics_measure(sim, swe_amd, nx, ny, reps=20):
    """Device diagnostics of the 4096^2 state (kernel time from HIP events, 32 B per
    interior cell read) against a host copy of b, h, hu, hv; and the gauge recorder's
    cost per step (1024 gauges sampled every step) against unrecorded steps."""
    log("[bench] device diagnostics and gauge recorder")
    for _ in range(3):
        sim.diagnostics(0.0)
    kern, t0 = [], time.perf_counter()
    for _ in range(reps):
        d = sim.diagnostics(0.0)
        kern.append(sim.get_scalar(swe_amd.S_DIAG_MS))
    wall = (time.perf_counter() - t0) / reps
    k_ms = sum(kern) / reps
    gbs = 32.0 * nx * ny / (k_ms * 1e-3) / 1e9
    t0 = time.perf_counter()
    sim.state()
    copy_s = time.perf_counter() - t0
    steps = 100
    t = sim.t
    sim.advance(steps)
    plain = sim.last_advance_ms() / steps
    W = nx + 4
    cells = [(2 + (q // 32) * (ny // 32)) * W + 2 + (q % 32) * (nx // 32) for q in range(1024)]
    sim.set_recorder(cells, every=1, capacity=steps)
    sim.advance(steps)
    recorded = sim.last_advance_ms() / steps
    _, eta = sim.record()
    sim.set_recorder([], every=1)
    return {"kernel_us": 1e3 * k_ms, "call_us": 1e6 * wall, "bytes_per_cell": 32, "achieved_GBs": gbs,
            "frac_hbm": gbs / HBM_PEAK_GBS, "host_copy_ms": 1e3 * copy_s, "values": d,
            "recorder": {"gauges": len(cells), "every": 1, "ms_per_step_plain": plain,
                         "ms_per_step_recorded": recorded, "samples": int(eta.shape[0]), "t_from": t}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--batched-steps", type=int, default=3000,
                    help="timed steps of the batched ensemble (default: the configuration's 3000); 0 skips it")
    # 94 rows is what the autotuner picks at 4096x4096 in most runs (48-96 rows time
    # within 1% of each other); pinning it keeps the profiled per-launch traffic and
    # instruction counts, taken at 94 rows, valid for the kernels timed here
    ap.add_argument("--chunk", type=int, default=94,
                    help="rows per wavefront chunk (default 94); -1 autotune, 0 the library default")
    ap.add_argument("--adjoint-steps", type=int, default=4000, help="0 skips the adjoint measurement")
    ap.add_argument("--adjoint-fd", type=int, default=16, help="finite-difference directions checked on the adjoint")
    ap.add_argument("--inversion-iters", type=int, default=0, help="L-BFGS iterations of the config 4 inversion")
    ap.add_argument("--sphere-steps", type=int, default=100, help="0 skips the spherical ensemble")
    ap.add_argument("--sphere-adjoint-steps", type=int, default=100)
    ap.add_argument("--sphere-adjoint-hours", type=float, default=0.0,
                    help="simulated hours of the friction-calibration adjoint on sphere member 0 (0 skips it)")
    ap.add_argument("--sphere-adjoint-fd", type=int, default=8, help="its finite-difference directions")
    ap.add_argument("--sphere-adjoint-fd-steps", default="0.1,0.01", help="their steps, comma-separated")
    ap.add_argument("--only-sphere-adjoint", action="store_true",
                    help="run only the --sphere-adjoint-hours leg and print its JSON")
    ap.add_argument("--no-diagnostics", action="store_true", help="skip the device diagnostics / recorder measurement")
    ap.add_argument("--globe-deg", type=float, default=1.0,
                    help="grid spacing of the n = 2 mode-period measurement (globe harness); 0 skips it")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    import swe_amd
    import workloads

    if args.only_sphere_adjoint:
        steps = tuple(float(x) for x in args.sphere_adjoint_fd_steps.split(","))
        print(json.dumps({"sphere_adjoint_long": sphere_adjoint_long(args.sphere_adjoint_hours or 0.5,
                                                                     args.sphere_adjoint_fd, steps=steps)}),
              flush=True)
        return
    log("[bench] building 4096x4096 workload")
    st, p = workloads.large_grid()
    nx, ny = p["nx"], p["ny"]
    rows = swe_amd.geometry_cartesian(nx, ny, p["dx"], coriolis=p["coriolis"])
    sim = swe_amd.Solver(nx, ny, rows, manning=p["manning"], coriolis=p["coriolis"], cfl=p["cfl"],
                         dx=p["dx"], dy=p["dx"])
    sim.load_state(*st)
    del st
    sim.set_bc(p["bc"])
    sim.set_harmonic(0, amp=[p["tide_amp"]], omega=[p["tide_omega"]], phase=[0.0])
    if args.chunk > 0:
        sim.set_chunk_rows(args.chunk)
    elif args.chunk < 0:
        log(f"[bench] autotuned chunk rows: {sim.autotune(reps=10)}")
    sim.use_graph(True)
    log(f"[bench] warmup {args.warmup} steps")
    sim.advance(args.warmup)
    sim.sync()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    n = sim.advance(args.steps)
    sim.sync()
    el = time.perf_counter() - t0
    ev_ms = sim.last_advance_ms()
    if n != args.steps:
        raise RuntimeError(f"only {n} of {args.steps} steps ran")
    value = nx * ny * n / el
    log(f"[bench] {value:.4e} cell-updates/s ({1e3 * el / n:.3f} ms/step; events {ev_ms / n:.3f} ms/step)")
    ms = sim.profile_step(args.profile_steps)
    W, H = nx + 4, ny + 4
    cells = nx * ny
    chunk = sim.chunk_rows()
    # HBM roof on the algorithmic bytes: 56 B per cell-update (read b, h, hu, hv;
    # write h, hu, hv), whole step; and for the dominant kernel, stage 2, its share
    # of a step's cell-updates (the same 56 B) over its own HIP-event time, and its
    # own compulsory 80 B per cell (b, A and B read, C written) beside it
    alg = 56.0 * value / 1e9
    k2 = 56.0 * cells / (ms[1] * 1e-3) / 1e9
    k2_own = 80.0 * W * H / (ms[1] * 1e-3) / 1e9
    profiled = chunk == 94
    roofline = {
        "bound": "valu", "achieved": alg, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": alg / HBM_PEAK_GBS,
        "basis": "56 B per cell-update x value (whole step, algorithmic bytes)",
        "kernel": "k_sweep<stage2>", "kernel_ms": ms[1],
        "kernel_achieved_56B": k2, "kernel_frac_56B": k2 / HBM_PEAK_GBS,
        "kernel_achieved_own_80B": k2_own, "kernel_frac_own_80B": k2_own / HBM_PEAK_GBS,
        "traffic": committed_traffic("k_sweep<2>") if profiled else None,
        "traffic_source": (f"{PROFILE}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per launch of k_sweep<2> at 94-row "
                           "chunks, each scaled by a plain-copy calibration; not measured in this run")
        if profiled else None,
    }
    # what bounds the stage kernels: VALU issue.  Instruction counts per launch
    # come from the committed profile (94-row chunks); the times are this run's.
    i1, i2 = committed_figure("valu_wave_instructions_per_launch", "k_sweep<1>"), \
        committed_figure("valu_wave_instructions_per_launch", "k_sweep<2>")
    d1, d2 = committed_figure("fp64_wave_instructions_per_launch", "k_sweep<1>"), \
        committed_figure("fp64_wave_instructions_per_launch", "k_sweep<2>")
    if i1 and i2 and profiled:
        t = (ms[0] + ms[1]) * 1e-3
        rate = (i1 + i2) / t
        roofline["valu"] = {
            "kernels": "k_sweep<stage1> + k_sweep<stage2>", "achieved": rate, "peak": VALU_PEAK_WAVE_INSTR_S,
            "unit": "wave64 VALU instr/s", "frac": rate / VALU_PEAK_WAVE_INSTR_S,
            "peak_note": ("approximate: every VALU instruction counted at one wave64 issue per SIMD per 4 cycles "
                          "(256 CU x 4 SIMD x 2.4 GHz / 4), the FP64 add/mul/FMA rate; FP64 transcendentals issue "
                          "slower, 32-bit moves and integer ops no faster"),
            "instructions_per_step": i1 + i2, "instructions_source": f"{PROFILE}: SQ_INSTS_VALU per launch",
            "cell_update_ceiling": VALU_PEAK_WAVE_INSTR_S * cells / (i1 + i2),
        }
        if d1 and d2:
            roofline["valu"]["fp64_instructions_per_step"] = d1 + d2
            roofline["valu"]["fp64_share"] = (d1 + d2) / (i1 + i2)
    # the whole step's HBM fraction on the algorithmic bytes, under its old name
    finite = sim.get_scalar(swe_amd.S_NONFINITE) == 0
    diag = None if args.no_diagnostics else diagnostics_measure(sim, swe_amd, nx, ny)
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": n,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / n, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded SplitMix64 bed), state resident in HBM",
        "config": {"workload": "large_grid_4096x4096_dx10_flather_tide_manning0.025_f1e-4_cfl0.45",
                   "nx": nx, "ny": ny, "dt": "self-timed (maxDt each step)", "graph": True,
                   "chunk_rows": sim.chunk_rows()},
        "hbm_pct_algorithmic_56B": 100.0 * 56.0 * value / (HBM_PEAK_GBS * 1e9),
        "kernel_ms": {"stage1": ms[0], "stage2": ms[1], "finalize": ms[2]},
        "roofline": roofline,
        "finite": finite,
        "diagnostics": diag,
    }
    sim.close()
    if args.batched_steps > 0:
        out["batched"] = batched_measure(args.batched_steps, 5, 20)
    if args.adjoint_steps > 0:
        out["adjoint"] = adjoint_measure(args.adjoint_steps, fd_dirs=args.adjoint_fd)
    if args.inversion_iters > 0:
        out["inversion"] = inversion_measure(args.inversion_iters)
    if args.sphere_steps > 0:
        out["sphere"] = sphere_measure(args.sphere_steps, args.sphere_adjoint_steps)
    if args.sphere_adjoint_hours > 0:
        out["sphere_adjoint_long"] = sphere_adjoint_long(
            args.sphere_adjoint_hours, args.sphere_adjoint_fd,
            steps=tuple(float(x) for x in args.sphere_adjoint_fd_steps.split(",")))
    if args.globe_deg > 0:
        import globe
        log(f"[bench] globe: n = 2 mode period at {args.globe_deg} deg")
        r = globe.measure_mode_period(n=2, deg=args.globe_deg, g_ref=9.80665)
        out["globe"] = {"workload": f"mode_period_n2_{args.globe_deg:g}deg_4000m_aquaplanet",
                        "nx": r["nx"], "ny": r["ny"], "steps": r["steps"], "seconds": r["gpu_ms"] / 1e3,
                        "cell_updates_per_s": r["cell_updates"] / (r["gpu_ms"] * 1e-3),
                        "measured_h": r["measured"] / 3600, "closed_h": r["closed"] / 3600, "rel": r["rel"],
                        "polar_tax": r["polar_tax"]["ratio"]}
    if not args.no_cpu:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
