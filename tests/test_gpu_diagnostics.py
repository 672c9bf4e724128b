"""Device diagnostics (swe_diagnostics) against the reference and the oracle.

Bar: energy, maxSurfaceDeviation, maxSpeed, finite and the spherical volume
equal the reference's bit for bit -- the reference's own final values recorded
in the golden fixtures where the GPU state is the reference's, and the oracle's
on the same state everywhere else.  The Cartesian volume is the exactly
rounded sum of h times dx times dy (math.fsum is the check), and lies within
the reference's own sequential rounding error of the reference's value.
"""
import math
import sys

import numpy as np
import pytest

import goldens
import oracle
import swe_amd
import workloads
from gpu_util import pair

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -52


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_factory(nx, ny, rows, **kw):
    kw.pop("use_manning_field", None)
    return swe_amd.Solver(nx, ny, rows, **kw)


def same(a, b):
    return (math.isnan(a) and math.isnan(b)) or a == b


def exact_volume(g, dx, dy):
    _, h, _, _ = g.state()
    return math.fsum(g.interior(h).ravel().tolist()) * dx * dy


def check_against_oracle(g, o, refs=(0.0,), dxdy=None):
    """Every diagnostic of g's device state against the oracle loaded with it."""
    o.load_state(*g.state())
    for ref in refs:
        d = g.diagnostics(ref)
        assert same(d["maxSurfaceDeviation"], o.max_surface_deviation(ref)), (ref, d)
    assert same(d["energy"], o.energy()), (d["energy"], o.energy())
    assert same(d["maxSpeed"], o.max_speed())
    assert d["finite"] == o.finite()
    vo = o.volume()
    if g.sphere:
        assert same(d["volume"], vo)
    elif math.isfinite(vo):
        dx, dy = dxdy
        assert d["volume"] == exact_volume(g, dx, dy)
        _, h, _, _ = g.state()
        n = g.nx * g.ny
        bound = n * EPS * float(np.abs(g.interior(h)).sum()) * dx * dy
        assert abs(d["volume"] - vo) <= bound, (d["volume"], vo, bound)
    else:
        assert same(d["volume"], vo)
    return d


@pytest.mark.parametrize("name", goldens.STEP_SCENARIOS)
def test_device_diagnostics_equal_reference(name):
    say(f"[gpu] diagnostics {name}")
    fx = goldens.load(name)
    g = goldens.make_sim(gpu_factory, fx)
    o = goldens.make_sim(oracle.OracleSim, fx)
    p = goldens.params(fx)
    fin = fx["final"]
    d0 = g.diagnostics(0.0)
    v0 = goldens.hexf(fin["volume0"])
    if p["sphere"]:
        assert d0["volume"] == v0
    else:
        assert d0["volume"] == exact_volume(g, p["dx"], p["dy"])
    exact = not goldens.uses_friction(fx)
    problems = goldens.replay(g, fx, exact=exact, rtol=1e-13)
    assert not problems, problems
    d = check_against_oracle(g, o, refs=(0.0, 0.25, -1.5), dxdy=(p["dx"], p["dy"]))
    if exact and fin["finite"]:
        # the GPU state is the reference's final one (a blow-up's replay stops at
        # its first non-finite step): its recorded diagnostics, bit for bit
        assert same(d["energy"], goldens.hexf(fin["energy"]))
        assert same(d["maxSpeed"], goldens.hexf(fin["maxSpeed"]))
        assert same(g.diagnostics(0.0)["maxSurfaceDeviation"], goldens.hexf(fin["maxSurfaceDeviation0"]))
        assert d["finite"] == fin["finite"]
        if p["sphere"]:
            assert same(d["volume"], goldens.hexf(fin["volume"]))


SHAPES = [(2, 1), (5, 3), (127, 9), (128, 8), (129, 17), (300, 203), (58, 64)]


@pytest.mark.parametrize("shape", SHAPES)
def test_device_diagnostics_tile_edges(shape):
    nx, ny = shape
    say(f"[gpu] diagnostics {nx}x{ny}")
    g, o = pair(nx, ny, 77 + nx, bc=("reflect", "outflow", "periodic", "periodic"), manning=0.02, coriolis=1e-3)
    for _ in range(6 if min(nx, ny) >= 8 else 0):    # the tiniest grids: the loaded state only
        g.step(None)
    check_against_oracle(g, o, refs=(0.0, 0.1, -3.0), dxdy=(5.0, 5.0))


def test_device_diagnostics_see_non_finite_cells():
    say("[gpu] diagnostics non-finite")
    nx, ny = 70, 41
    W = nx + 4
    for where, field, val in (("ghost-corner", 1, math.nan), ("ghost-east", 3, math.inf),
                              ("interior", 2, -math.inf), ("interior-h", 1, math.nan)):
        g, o = pair(nx, ny, 5)
        b, h, hu, hv = (a.copy() for a in g.state())
        k = {"ghost-corner": 0, "ghost-east": 20 * W + W - 1, "interior": 30 * W + 40,
             "interior-h": 12 * W + 9}[where]
        [b, h, hu, hv][field][k] = val
        g.load_state(b, h, hu, hv)
        d = check_against_oracle(g, o, dxdy=(5.0, 5.0))
        assert d["finite"] is False, where


def test_device_diagnostics_full_sizes():
    """The 4096x4096 benchmark workload after 5 steps and the 2160x1080 global ocean
    after 3: every diagnostic against the oracle on the same state."""
    (b, h, hu, hv), p = workloads.large_grid()
    nx = ny = 4096
    rows = swe_amd.geometry_cartesian(nx, ny, p["dx"], coriolis=p["coriolis"])
    g = swe_amd.Solver(nx, ny, rows, manning=p["manning"], coriolis=p["coriolis"], cfl=p["cfl"], dx=p["dx"], dy=p["dx"])
    g.load_state(b, h, hu, hv)
    g.set_bc(p["bc"])
    g.set_harmonic(0, amp=[p["tide_amp"]], omega=[p["tide_omega"]], phase=[0.0])
    assert g.advance(5) == 5
    o = oracle.OracleSim(nx, ny, rows, manning=p["manning"], coriolis=p["coriolis"], cfl=p["cfl"],
                         dx=p["dx"], dy=p["dx"])
    d = check_against_oracle(g, o, refs=(0.0, 1.0), dxdy=(p["dx"], p["dx"]))
    say(f"  4096^2: {d}, kernels {g.get_scalar(swe_amd.S_DIAG_MS) * 1e3:.1f} us")
    g.close()
    del o

    b, h, ps, manning = workloads.sphere_tides()
    rows = swe_amd.geometry_spherical(ps["nx"], ps["ny"], R=ps["R"], lat0=ps["lat0"], lat1=ps["lat1"],
                                      omega=ps["omega"])
    kw = dict(sphere=True, manning=manning[0], cfl=ps["cfl"], R=ps["R"])
    g = swe_amd.Solver(ps["nx"], ps["ny"], rows, **kw)
    z = np.zeros_like(h)
    g.load_state(b, h, z, z)
    g.set_bc(("periodic", "periodic", "reflect", "reflect"))
    assert g.advance(3) == 3
    o = oracle.OracleSim(ps["nx"], ps["ny"], rows, **kw)
    d = check_against_oracle(g, o, refs=(0.0,))
    say(f"  sphere 2160x1080: {d}, kernels {g.get_scalar(swe_amd.S_DIAG_MS) * 1e3:.1f} us")
