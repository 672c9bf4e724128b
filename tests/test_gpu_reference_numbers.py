"""Numbers the reference records for its own runs, reproduced on the GPU
through the C ABI by restating the runs that produced them.

* tools/verify.mjs:405-449, the closed-basin seiche (Merian): L = 5 km, 20 m
  deep, 500 x 1 cells, eta = 0.05 cos(pi x / L), reflecting ends, sampled once
  a second by sub-stepping min(maxDt, 1 - acc), 8.4 closed-form periods; the
  reference's comment (:426-427) records 714.0411 s (against Merian's
  714.0435 s) and 9 downward crossings.
* tools/verify-tide.mjs:492-570, quarter-wave resonance: a 91 km, 40 m bay
  behind a 39 km, 400 m approach (200 x 1 cells), Manning 0.010, CFL 0.4,
  forced at the west by Flather with a 0.5 m cosine faded in over 3 periods
  and u = sqrt(g / 400) eta, recorded at the head 240 times a period over
  periods 8-12, at eleven forcing periods; the parabola through the peak
  gain and its neighbours puts resonance at T / T_res = 1.0220 (:681, "fit
  1.02200" in its mutation table), and the gains either side of the peak
  read 2.7820 (0.975) and 2.7866 (1.025) (:548-552).  Run device-resident:
  the forcing with swe_set_harmonic_ramp, the samples with the recorder on
  waves.mjs's schedule, which is this loop's.
* tools/verify-tide.mjs:1131-1195 and README.md:284-287, a real tide over a
  drying flat: 30 km, 150 cells, bed -6 + 7.4 (x/L)^1.5, Manning 0.025, CFL
  0.4, Flather with Tide({M2: [1.6, 0], S2: [0.5, 30]}) and u = sqrt(g/6) eta,
  sampled every 120 s over three M2 periods (the last one and a half for the
  gauges); the README records the wetted fraction swinging 86.7% to 100.0%,
  the range falling from 6.53 m at the mouth to 4.84 m in the inner basin,
  mean flood 2.83 h against mean ebb 9.47 h, an asymmetry of 107.9%.
"""
import math

import numpy as np
import pytest

import swe_amd

pytestmark = pytest.mark.gpu
G_REF = 9.80665


def test_merian_seiche_period_is_the_recorded_714_0411():
    L, h0, nx = 5000.0, 20.0, 500
    dx = L / nx
    rows = swe_amd.geometry_cartesian(nx, 1, dx)
    s = swe_amd.Solver(nx, 1, rows, manning=0.0, dx=dx, dy=dx)
    W, H = nx + 4, 5
    x = (np.arange(W) - 2 + 0.5) * dx
    eta = np.array([0.05 * math.cos(math.pi * xi / L) for xi in x])
    b = np.full(W * H, -h0)
    h = np.maximum(0.0, np.tile(eta, H) - b)
    z = np.zeros_like(h)
    s.load_state(b, h, z, z)
    s.set_bc(("reflect", "reflect", "periodic", "periodic"))
    want = 2 * L / math.sqrt(G_REF * h0)
    k = 2 * W + 4                                   # sim.idx(2, 0)
    hist = []
    while s.t < 8.4 * want:
        acc = 0.0
        while acc < 1.0 - 1e-12:
            acc += s.step(min(s.max_dt(), 1.0 - acc))
        hist.append((s.t, float(s.b[k] + s.h[k])))
    zc = []
    for n in range(1, len(hist)):
        if hist[n - 1][1] > 0 and hist[n][1] <= 0:
            f = hist[n - 1][1] / (hist[n - 1][1] - hist[n][1])
            zc.append(hist[n - 1][0] + f * (hist[n][0] - hist[n - 1][0]))
    measured = (zc[-1] - zc[0]) / (len(zc) - 1)
    print(f"  Merian seiche: {measured!r} s, {len(zc)} crossings (recorded 714.0411 s, 9)", flush=True)
    assert len(zc) == 9
    assert abs(measured - 714.0411) < 5e-5
    assert abs(measured - want) / want < 1e-5


def test_resonance_sweep_peaks_at_the_recorded_1_0220():
    Lbay, hBay, Lapp, hApp = 91000.0, 40.0, 39000.0, 400.0
    t_res = 4 * Lbay / math.sqrt(G_REF * hBay)
    L, nx = Lapp + Lbay, 200
    dx = L / nx
    rows = swe_amd.geometry_cartesian(nx, 1, dx)
    W, H = nx + 4, 5
    x = (np.arange(W) - 2 + 0.5) * dx
    b = np.tile(np.where(x < Lapp, -hApp, -hBay), H)
    h = np.maximum(0.0, 0.0 - b)
    z = np.zeros_like(h)
    head = 2 * W + (nx - 2 + 2)                     # sim.eta(nx - 2, 0)
    ratios = [0.55, 0.80, 0.90, 0.95, 0.975, 1.00, 1.025, 1.05, 1.10, 1.25, 1.70]
    gains = []
    for r in ratios:
        period = t_res * r
        s = swe_amd.Solver(nx, 1, rows, manning=0.010, cfl=0.4, dx=dx, dy=dx)
        s.load_state(b, h, z, z)
        s.set_bc(("flather", "reflect", "periodic", "periodic"))
        s.set_harmonic(0, amp=[0.5], omega=[2 * math.pi / period], phase=[0.0], ramp=3 * period,
                       u_scale=math.sqrt(G_REF / hApp))
        s.set_recorder([head], every=0, t0=8 * period, dt_rec=period / 240, capacity=2000)
        s.autotune(reps=2)
        s.use_graph(True)
        s.advance(10_000_000, swe_amd.DT_SAMPLED, t_end=12 * period)
        _, eta = s.record()
        s.close()
        gains.append(0.5 * (eta.max() - eta.min()) / 0.5)
        print(f"  T/T_res = {r:.3f}: gain {gains[-1]:.5f} ({eta.shape[0]} samples)", flush=True)
    i = int(np.argmax(gains))

    def vertex(ks):
        """The vertex of the parabola through three (ratio, gain) points, and its curvature."""
        (p0, g0), (p1, g1), (p2, g2) = [(ratios[k], gains[k]) for k in ks]
        d1, d2, d3 = (p0 - p1) * (p0 - p2), (p1 - p0) * (p1 - p2), (p2 - p0) * (p2 - p1)
        A2 = g0 / d1 + g1 / d2 + g2 / d3
        B2 = -(g0 * (p1 + p2) / d1 + g1 * (p0 + p2) / d2 + g2 * (p0 + p1) / d3)
        return -B2 / (2 * A2), A2

    refined, A2 = vertex((i - 1, i, i + 1))
    print(f"  fitted resonance T/T_res = {refined:.6f} (recorded 1.02200)", flush=True)
    assert A2 < 0 and 0 < i < len(ratios) - 1
    assert abs(refined - 1.02200) < 1e-5
    assert abs(gains[ratios.index(0.975)] - 2.7820) < 1e-4
    assert abs(gains[ratios.index(1.025)] - 2.7866) < 1e-4
    # the whole gain table and the wider stencils, as the suite prints them
    # (verify-tide.mjs:596-640: "0.55:2.032 ... 1.7:2.418", "wider stencils give
    # 1.01653 (0.975, 1.025, 1.1), 1.02700 (0.95, 1.025, 1.25)")
    assert " ".join(f"{g:.3f}" for g in gains) == \
        "2.032 2.603 2.742 2.774 2.782 2.784 2.787 2.783 2.768 2.683 2.418"
    wide = [vertex([ratios.index(r) for r in rs])[0] for rs in ((0.975, 1.025, 1.10), (0.95, 1.025, 1.25))]
    assert [f"{w:.5f}" for w in wide] == ["1.01653", "1.02700"]


def test_drying_flat_tide_matches_the_recorded_figures():
    L, nx = 30000.0, 150
    dx = L / nx
    rows = swe_amd.geometry_cartesian(nx, 1, dx)
    W, H = nx + 4, 5
    x = (np.arange(W) - 2 + 0.5) * dx
    # the west ghost columns (x < 0) get NaN, as the reference's constructor gives
    # them; the boundary fill before the first stage replaces them
    with np.errstate(invalid="ignore"):
        brow = np.array([-6 + 7.4 * (xi / L) ** 1.5 for xi in x])
    b = np.tile(brow, H)
    h = np.maximum(0.0, 0.0 - b)
    z = np.zeros_like(h)
    s = swe_amd.Solver(nx, 1, rows, manning=0.025, cfl=0.4, dx=dx, dy=dx)
    s.load_state(b, h, z, z)
    s.set_bc(("flather", "reflect", "periodic", "periodic"))
    # Tide.eta: msl + sum amp cos(omega t - phase pi / 180), largest amplitude first (tide.mjs:87-107)
    m2, s2 = 12.4206012 * 3600, 12.0 * 3600
    s.set_harmonic(0, amp=[1.6, 0.5], omega=[2 * math.pi / m2, 2 * math.pi / s2],
                   phase=[0 * math.pi / 180, 30 * math.pi / 180], u_scale=math.sqrt(G_REF / 6))
    cells = [2 * W + (i + 2) for i in range(nx)]
    T, T0 = 3 * m2, 1.5 * m2
    s.set_recorder(cells, every=0, t0=0.0, dt_rec=120.0, capacity=2000)
    s.use_graph(True)
    s.advance(10_000_000, swe_amd.DT_SAMPLED, t_end=T)
    times, eta = s.record()
    assert s.finite()
    wet = ((eta - brow[2:-2][None, :]) > 1e-3).sum(axis=1) / nx
    late = times >= T0
    mouth, inner, ti = eta[late, 4], eta[late, nx - 30], times[late]
    ex = []
    for n in range(1, len(inner) - 1):
        a, c, d = inner[n - 1], inner[n], inner[n + 1]
        if c > a and c >= d:
            ex.append((ti[n], "H"))
        if c < a and c <= d:
            ex.append((ti[n], "L"))
    flood = [ex[n + 1][0] - ex[n][0] for n in range(len(ex) - 1) if ex[n][1] == "L" and ex[n + 1][1] == "H"]
    ebb = [ex[n + 1][0] - ex[n][0] for n in range(len(ex) - 1) if ex[n][1] == "H" and ex[n + 1][1] == "L"]
    fm, em = sum(flood) / len(flood), sum(ebb) / len(ebb)
    asym = abs(fm - em) / ((fm + em) / 2)
    rng_mouth, rng_inner = mouth.max() - mouth.min(), inner.max() - inner.min()
    print(f"  wetted {100 * wet.min():.1f}% to {100 * wet.max():.1f}%, range {rng_mouth:.2f} m -> {rng_inner:.2f} m, "
          f"flood {fm / 3600:.2f} h, ebb {em / 3600:.2f} h, asymmetry {100 * asym:.1f}%", flush=True)
    assert f"{100 * wet.min():.1f}" == "86.7" and f"{100 * wet.max():.1f}" == "100.0"
    assert f"{rng_mouth:.2f}" == "6.53" and f"{rng_inner:.2f}" == "4.84"
    assert f"{fm / 3600:.2f}" == "2.83" and f"{em / 3600:.2f}" == "9.47"
    assert f"{100 * asym:.1f}" == "107.9"
