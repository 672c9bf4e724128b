"""Quarter-wave resonance of a tidal basin, run as one batched period sweep.

A channel closed at its east end and forced through a west Flather boundary
(external surface a cos(2 pi t / T), external velocity 0) settles into a standing
wave whose antinode sits at the wall: eta(x) = A cos(k (L - x)), k = 2 pi / (c T),
c = sqrt(g h).  Near T_res = 4 L / c the mouth becomes a node -- the quarter-wave
resonance that the reference's tide suite takes as its headline check
(tools/verify-tide.mjs, section 2).  Nothing in the solver knows about it: the
response emerges from the step (swe.mjs:754-814) with the Flather boundary
(swe.mjs:1084-1116) and the closed wall (:1026-1039).

Checked, frictionless, against linear theory for every period of the sweep: the
ratio of the mouth cell's amplitude to the closed-end cell's is
|cos(k (L - dx/2))| / cos(k dx/2) to 0.006, the smallest ratio (the node) falls
at T = T_res, and the wall's amplitude equals the forcing amplitude to 0.5%
(the radiating boundary lets the reflected wave out)."""
import math

import numpy as np
import pytest

import swe_amd

pytestmark = pytest.mark.gpu

G = 9.80665


def test_quarter_wave_resonance_period_sweep():
    nx, ny, dx, depth, amp = 80, 4, 500.0, 10.0, 0.05
    L = nx * dx
    c = math.sqrt(G * depth)
    t_res = 4 * L / c
    ratios = [0.7, 0.85, 0.95, 1.0, 1.05, 1.2, 1.5]
    rows = swe_amd.geometry_cartesian(nx, ny, dx)
    W, H = nx + 4, ny + 4
    B = swe_amd.Batch([dict(nx=nx, ny=ny, rows=rows, manning=0.0) for _ in ratios])
    b = np.full(W * H, -depth)
    h = np.full(W * H, depth)
    z = np.zeros(W * H)
    for q, r in enumerate(ratios):
        B.load_state(q, b, h, z, z)
        B.set_bc(q, ("flather", "reflect", "reflect", "reflect"))
        B.set_harmonic(q, 0, amp=[amp], omega=[2 * math.pi / (r * t_res)], phase=[0.0])
    B.use_graph(True)
    jm = 2 + ny // 2
    k_end, k_mouth = jm * W + nx + 1, jm * W + 2
    end = np.zeros(len(ratios))
    mouth = np.zeros(len(ratios))
    # spin up for 6 forcing periods, then record the amplitudes over 3 more
    while True:
        B.advance(20)
        done = True
        for q, r in enumerate(ratios):
            T = r * t_res
            t = B.get_scalar(q, swe_amd.S_T)
            if t < 9 * T:
                done = False
            if 6 * T <= t < 9 * T:
                hh = B.state(q)[1]
                end[q] = max(end[q], abs(hh[k_end] - depth))
                mouth[q] = max(mouth[q], abs(hh[k_mouth] - depth))
        if done:
            break
    for q, r in enumerate(ratios):
        k = 2 * math.pi / (c * r * t_res)
        theory = abs(math.cos(k * (L - dx / 2))) / math.cos(k * dx / 2)
        print(f"  T/T_res {r:.2f}: wall {end[q] / amp:.4f} x forcing, mouth/wall {mouth[q] / end[q]:.4f} "
              f"(linear theory {theory:.4f})")
        assert abs(mouth[q] / end[q] - theory) <= 0.006, (r, mouth[q] / end[q], theory)
        assert abs(end[q] / amp - 1.0) <= 0.005, (r, end[q] / amp)
    assert ratios[int(np.argmin(mouth / end))] == 1.0
