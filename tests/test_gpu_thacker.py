"""Config 1: Thacker's radially symmetric bowl, 256x256, two analytic periods.

* Parity: the reference solver's own run (golden fixture thacker256, g = 9.80665
  as the reference hard-codes it) -- the first 60 steps stepped one by one
  through the C ABI with every fingerprint compared, then the whole run to two
  periods device-resident (dt = min(maxDt, 2T - t) on the device) with the final
  fingerprint, step count and t compared bit for bit.
* Closed form: the same run at g = 9.81 against Thacker's solution after two
  periods (mean depth error over the wet disc, shoreline band excluded).
"""
import math

import numpy as np
import pytest

import goldens
import swe_amd
import workloads

pytestmark = pytest.mark.gpu


def gpu_factory(nx, ny, rows, **kw):
    kw.pop("use_manning_field", None)
    return swe_amd.Solver(nx, ny, rows, **kw)


def test_thacker256_bit_exact_against_reference():
    print("[gpu] thacker256 parity", flush=True)
    fx = goldens.load("thacker256")
    sim = goldens.make_sim(gpu_factory, fx)
    steps = fx["steps"]
    for n, st in enumerate(steps[:60], start=1):
        dt = sim.step(goldens.hexf(st["dtArg"]))
        assert dt == goldens.hexf(st["dt"]), n
        assert goldens.fingerprint(sim) == int(st["hash"], 16), n
    # the whole run device-resident, self-timed against the end time
    sim2 = goldens.make_sim(gpu_factory, fx)
    T2 = goldens.hexf(fx["final"]["t"])
    taken = sim2.advance(len(steps), swe_amd.DT_UNTIL, 0.0, T2)
    assert taken == len(steps) == fx["final"]["steps"]
    assert sim2.t == T2
    assert goldens.fingerprint(sim2) == int(steps[-1]["hash"], 16)
    # at t = T2 every further step is refused (dt = 0): like the reference's
    # step(), it refreshes the ghost ring and leaves the interior and t alone
    inner = [sim2.interior(a).copy() for a in (sim2.h, sim2.hu, sim2.hv)]
    assert sim2.advance(5, swe_amd.DT_UNTIL, 0.0, T2) == 0
    assert sim2.t == T2
    for a, b in zip(inner, (sim2.h, sim2.hu, sim2.hv)):
        assert np.array_equal(a, sim2.interior(b))


def test_thacker_closed_form_g981():
    print("[gpu] thacker closed form, g = 9.81", flush=True)
    st, p, exact = workloads.thacker_bowl(256, g=9.81)
    rows = swe_amd.geometry_cartesian(256, 256, p["dx"])
    sim = swe_amd.Solver(256, 256, rows, g=9.81, manning=0.0, dx=p["dx"], dy=p["dx"])
    sim.load_state(*st)
    sim.set_bc(p["bc"])
    T2 = 2 * p["period"]
    sim.advance(20000, swe_amd.DT_UNTIL, 0.0, T2)
    assert sim.t == T2
    h = sim.interior(sim.h)
    he = sim.interior(exact(T2))
    wet = he > 0.05
    err = np.abs(h - he)[wet].mean()
    print(f"  mean |h - exact| over {int(wet.sum())} wet cells after 2 periods: {err:.3e} m (h0 = 1 m)")
    assert err < 0.02
    # the volume is conserved exactly by the walls
    assert abs(h.sum() - sim.interior(st[1]).sum()) <= 1e-11 * st[1].sum()
