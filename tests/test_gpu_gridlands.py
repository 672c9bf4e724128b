"""The gridlands-terrain import driving the HIP solver (the acceptance gates of
tools/verify-gridlands.mjs on this build's synthetic export).

Bar: the GPU equals the oracle bit for bit over 1000 fixed-dt steps of the
imported map, as imported (two perched lakes spilling) and settled; the settled
map stays at rest within the reference's derived tolerance
4 eps g h_max dt N / dx (max speed from the device diagnostics) with the
surface within 1e-9 m of where it started, and both runs conserve volume to
1e-13 relative.
"""
import gzip
import json
import os
import sys

import numpy as np
import pytest

import gridlands as gl
import oracle
import swe_amd

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "gridlands_synth.terrain.json.gz")
G, EPS = 9.80665, 2.0 ** -52


@pytest.mark.parametrize("settle", [False, True])
def test_imported_map_gpu_equals_oracle_and_rests_when_settled(settle):
    print(f"[gpu] gridlands import settle={settle}", flush=True)
    with gzip.open(FIXTURE, "rt") as f:
        t = gl.gridlands_terrain(json.load(f), settle=settle)
    nx, ny, dx = t["W"], t["H"], t["cellMeters"]
    rows = swe_amd.geometry_cartesian(nx, ny, dx)
    kw = dict(manning=0.0, cfl=0.45, dx=dx, dy=dx)
    g = swe_amd.Solver(nx, ny, rows, **kw)
    o = oracle.OracleSim(nx, ny, rows, **kw)
    b, h = gl.sample_state(t)
    z = np.zeros_like(h)
    for s in (g, o):
        s.load_state(b, h, z, z)
        s.set_bc(("reflect",) * 4)
    inner = np.zeros((ny + 4, nx + 4), bool)
    inner[2:-2, 2:-2] = True
    inner = inner.ravel()
    h_max = h[inner].max()
    v0 = g.diagnostics()["volume"]
    dt = g.max_dt()
    assert dt == o.max_dt()
    for n in range(1000):
        assert g.step(dt) == o.step(dt) == dt
    _, gh, ghu, ghv = g.state()
    for a, w in ((gh, o.h), (ghu, o.hu), (ghv, o.hv)):
        assert np.array_equal(a.view(np.uint64), w.view(np.uint64))
    d = g.diagnostics()
    assert d["finite"]
    assert abs(d["volume"] - v0) / v0 < 1e-13
    u_tol = 4 * EPS * G * h_max * dt * 1000 / dx
    wet = inner & (gh > 1e-3)
    d_eta = np.abs((b + gh) - (b + h))[wet].max()
    print(f"  max speed {d['maxSpeed']:.3e} m/s (tolerance {u_tol:.3e}), max |eta - eta0| {d_eta:.3e} m")
    if settle:
        assert d["maxSpeed"] < u_tol and d_eta < 1e-9
    else:
        assert d["maxSpeed"] > 1e3 * u_tol            # the perched lakes spill
