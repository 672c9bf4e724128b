"""Boundary forcing evaluated on the device, against the host.

* The Flather values a device-resident run evaluates (swe_set_harmonic: the sums
  of cosines Tide.eta computes, src/tide.mjs:103-107, and regularWave's ramped
  wave with u = c (eta - msl), src/waves.mjs:163-175) against the same sums
  evaluated on the host with the C library's cos, step by step: most are equal,
  and every one is within 4 eps (|c0| + sum |a_k|), the reach of a cos that is
  off by at most one unit in its last place in each term (a sum that nearly
  cancels can show that as several ulps of the sum itself).
* A host step (swe_step with the caller's values) leaves the harmonics armed: a
  later swe_advance evaluates them again.
* A batch of nine uneven basins with nonzero Flather harmonics on one or two
  sides against the oracle (the reference's step restated on the CPU), basin by
  basin: fed the device's own values the oracle reproduces every step bit for
  bit; fed the host's cos it stays within 1e-12 of the device with the same
  wet/dry mask.
"""
import math

import numpy as np
import pytest

import oracle
import swe_amd
from gpu_util import random_basin

pytestmark = pytest.mark.gpu


EPS = 2.0 ** -52


def within(got, want, scale):
    """|got - want| in units of eps * scale, scale = |c0| + sum |a_k| of the sum."""
    return abs(got - want) / (EPS * scale)


def harm(t, msl, terms, ramp=0.0):
    """msl + sum a cos(w t - p) in the device's order, a faded in by min(1, t/ramp)."""
    e = msl
    m = min(1.0, t / ramp) if ramp > 0 else 1.0
    for a, w, p in terms:
        e += (m * a if ramp > 0 else a) * math.cos(w * t - p)
    return e


def test_device_harmonics_within_one_ulp_of_host_cos():
    nx, ny = 60, 40
    rows = swe_amd.geometry_cartesian(nx, ny, 5.0)
    s = swe_amd.Solver(nx, ny, rows, manning=0.02, dx=5.0, dy=5.0)
    s.load_state(*random_basin(nx, ny, 11))
    s.set_bc(("flather", "flather", "reflect", "reflect"))
    west = [(0.2, 2 * math.pi / 37.0, 0.4), (0.05, 2 * math.pi / 11.0, 1.3)]
    east = [(0.1, 2 * math.pi / 23.0, -0.7)]
    uw = [(0.03, 2 * math.pi / 37.0, 0.9)]
    s.set_harmonic(0, amp=[a for a, _, _ in west], omega=[w for _, w, _ in west], phase=[p for _, _, p in west],
                   msl=0.05, uamp=[a for a, _, _ in uw], uomega=[w for _, w, _ in uw], uphase=[p for _, _, p in uw],
                   u0=0.01)
    s.set_harmonic(1, amp=[a for a, _, _ in east], omega=[w for _, w, _ in east], phase=[p for _, _, p in east],
                   msl=-0.02, ramp=15.0, u_scale=0.7)
    worst, exact, n = 0, 0, 0
    for step in range(300):
        s.advance(1)
        t = s.t
        ext = s.ext()
        ew = harm(t, 0.05, west)
        ee = harm(t, -0.02, east, ramp=15.0)
        want = [ew, harm(t, 0.01, uw), ee, 0.7 * (ee - (-0.02))]
        scale = [0.05 + 0.25, 0.01 + 0.03, 0.02 + 0.1, 0.7 * (0.02 + 0.1 + 0.02)]
        got = [ext[0], ext[1], ext[2], ext[3]]
        for g_, w_, sc in zip(got, want, scale):
            d = within(g_, w_, sc)
            worst = max(worst, d)
            exact += g_ == w_
            n += 1
    print(f"device harmonics vs host cos: {exact} of {n} equal, worst {worst:.2f} eps x scale", flush=True)
    assert exact >= 0.9 * n
    assert worst <= 4.0


def _solver(nx, ny, seed, bc):
    rows = swe_amd.geometry_cartesian(nx, ny, 5.0)
    s = swe_amd.Solver(nx, ny, rows, manning=0.02, dx=5.0, dy=5.0)
    s.load_state(*random_basin(nx, ny, seed))
    s.set_bc(bc)
    s.set_harmonic(0, amp=[0.15, 0.04], omega=[2 * math.pi / 31.0, 2 * math.pi / 9.0], phase=[0.2, 0.0],
                   msl=0.02, uamp=[0.02], uomega=[2 * math.pi / 31.0], uphase=[0.5])
    return s


def test_host_step_keeps_the_harmonics_for_advance():
    """set_harmonic; advance(1); step() with the device's own next values; advance(n)
    equals set_harmonic; advance(1); advance(1); advance(n) bit for bit: the host step
    does not disarm the device forcing."""
    nx, ny = 70, 44
    bc = ("flather", "reflect", "periodic", "periodic")
    a, b = _solver(nx, ny, 21, bc), _solver(nx, ny, 21, bc)
    a.advance(1)
    b.advance(1)
    ext = a.ext()
    assert ext[0] != 0.0
    a.step(None, ext)
    b.advance(1)
    a.advance(40)
    b.advance(40)
    assert a.t == b.t
    for f in (swe_amd.F_H, swe_amd.F_HU, swe_amd.F_HV):
        assert np.array_equal(a.get_field(f).view(np.uint64), b.get_field(f).view(np.uint64))
    assert a.ext()[0] == b.ext()[0]


BASINS = [(64, 48), (97, 33), (130, 71), (40, 150), (200, 45), (33, 33), (77, 129), (150, 90), (58, 58)]


def test_batch_flather_harmonics_against_the_oracle():
    rng = np.random.default_rng(5)
    specs = []
    for q, (nx, ny) in enumerate(BASINS):
        sides = ["flather", "reflect", "reflect", "reflect"]
        if q % 3 == 1:
            sides[1] = "flather"
        if q % 3 == 2:
            sides[2], sides[3] = "periodic", "periodic"
        forc = {}
        for side in (0, 1):
            if sides[side] != "flather":
                continue
            P = rng.uniform(15, 60)
            forc[side] = dict(msl=float(rng.uniform(-0.05, 0.05)), amp=[float(rng.uniform(0.05, 0.25))],
                              omega=[2 * math.pi / P], phase=[float(rng.uniform(0, 2 * math.pi))],
                              uamp=[float(rng.uniform(0.0, 0.05))], uomega=[2 * math.pi / P],
                              uphase=[float(rng.uniform(0, 2 * math.pi))], u0=float(rng.uniform(-0.01, 0.01)))
        specs.append(dict(nx=nx, ny=ny, bc=tuple(sides), forc=forc, state=random_basin(nx, ny, 100 + q)))
    B = swe_amd.Batch([dict(nx=s["nx"], ny=s["ny"], manning=0.02, rows=swe_amd.geometry_cartesian(s["nx"], s["ny"], 5.0))
                       for s in specs])
    dev_fed, host_fed = [], []
    for q, s in enumerate(specs):
        B.load_state(q, *s["state"])
        B.set_bc(q, s["bc"])
        for side, f in s["forc"].items():
            B.set_harmonic(q, side, **f)
        rows = swe_amd.geometry_cartesian(s["nx"], s["ny"], 5.0)
        for lst in (dev_fed, host_fed):
            o = oracle.OracleSim(s["nx"], s["ny"], rows, manning=0.02, dx=5.0, dy=5.0)
            o.load_state(*s["state"])
            o.set_bc(s["bc"])
            lst.append(o)
    B.set_chunk_rows(16)
    B.advance(0)             # evaluates the forcing at t = 0
    worst, exact, n = 0, 0, 0
    for step in range(150):
        for q, s in enumerate(specs):
            ext_dev = B.ext(q)

            def host_ext(t):
                e = np.zeros(8)
                for side, f in s["forc"].items():
                    e[2 * side] = harm(t, f["msl"], list(zip(f["amp"], f["omega"], f["phase"])))
                    e[2 * side + 1] = harm(t, f["u0"], list(zip(f["uamp"], f["uomega"], f["uphase"])))
                return e
            # the device's values against the host's at the device's own t
            at_dev_t = host_ext(dev_fed[q].t)
            for side, f in s["forc"].items():
                for k, sc in ((2 * side, abs(f["msl"]) + sum(map(abs, f["amp"]))),
                              (2 * side + 1, abs(f["u0"]) + sum(map(abs, f["uamp"])))):
                    d = within(ext_dev[k], at_dev_t[k], sc)
                    worst = max(worst, d)
                    exact += ext_dev[k] == at_dev_t[k]
                    n += 1
            dev_fed[q].step(None, ext_dev)
            host_fed[q].step(None, host_ext(host_fed[q].t))
        B.advance(1)
        for q, s in enumerate(specs):
            b_, h, hu, hv = B.state(q)
            o = dev_fed[q]
            for name, x, y in (("h", h, o.h), ("hu", hu, o.hu), ("hv", hv, o.hv)):
                assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), (step, q, name)
            assert B.get_scalar(q, swe_amd.S_T) == o.t, (step, q)
    print(f"batch Flather values vs host cos: {exact} of {n} equal, worst {worst:.2f} eps x scale", flush=True)
    assert exact >= 0.9 * n
    assert worst <= 4.0
    for q, s in enumerate(specs):
        _, h, hu, hv = B.state(q)
        o = host_fed[q]
        for x, y in ((h, o.h), (hu, o.hu), (hv, o.hv)):
            scale = max(np.abs(y).max(), 1e-300)
            assert np.abs(x - y).max() <= 1e-12 * scale, q
        assert np.array_equal(h > 1e-3, o.h > 1e-3), q
