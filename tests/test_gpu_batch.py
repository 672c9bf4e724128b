"""Batched ensembles on the GPU: every basin of a batch must be bit-identical to
a plain single-basin run of the same basin (same dt sequence, same state), with
and without graph replay, whatever the task shape the autotuner picks."""
import sys

import numpy as np
import pytest

import swe_amd
import workloads
from gpu_util import bits

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def build(specs, kind):
    if kind == "batch":
        B = swe_amd.Batch([dict(nx=s["nx"], ny=s["ny"], manning=s["manning"], cfl=s["cfl"],
                                rows=swe_amd.geometry_cartesian(s["nx"], s["ny"], s["dx"])) for s in specs])
        for i, s in enumerate(specs):
            B.load_state(i, *s["state"])
            B.set_bc(i, s["bc"])
            B.set_harmonic(i, 0, amp=[s["amp"]], omega=[s["omega"]], phase=[0.0],
                           uamp=[s["uamp"]], uomega=[s["omega"]], uphase=[0.0])
        return B
    sims = []
    for s in specs:
        g = swe_amd.Solver(s["nx"], s["ny"], swe_amd.geometry_cartesian(s["nx"], s["ny"], s["dx"]),
                           manning=s["manning"], cfl=s["cfl"])
        g.load_state(*s["state"])
        g.set_bc(s["bc"])
        g.set_harmonic(0, amp=[s["amp"]], omega=[s["omega"]], phase=[0.0],
                       uamp=[s["uamp"]], uomega=[s["omega"]], uphase=[0.0])
        sims.append(g)
    return sims


@pytest.mark.parametrize("graph", [True, False])
def test_batch_basins_bit_identical_to_plain_runs(graph):
    specs = workloads.ensemble(nb=9, lo=6, hi=180, seed=17)
    say(f"[gpu] batch vs plain, graph={graph}, shapes {[(s['nx'], s['ny']) for s in specs]}")
    B = build(specs, "batch")
    B.use_graph(graph)
    B.set_chunk_rows(16)
    plain = build(specs, "plain")
    assert B.advance(12) == 12
    for g in plain:
        assert g.advance(12) == 12
    for i, g in enumerate(plain):
        for a, b in zip(B.state(i), g.state()):
            assert np.array_equal(bits(a), bits(b)), i
        assert B.get_scalar(i, swe_amd.S_T) == g.t


def test_batch_autotune_keeps_state_and_results():
    specs = workloads.ensemble(nb=5, lo=40, hi=300, seed=23)
    say("[gpu] batch autotune")
    B = build(specs, "batch")
    before = [B.state(i) for i in range(len(specs))]
    rows, waves = B.autotune(steps=1)
    assert rows >= 8 and waves >= 1
    for i in range(len(specs)):
        for a, b in zip(B.state(i), before[i]):
            assert np.array_equal(bits(a), bits(b))
        assert B.get_scalar(i, swe_amd.S_STEPS) == 0
    plain = build(specs, "plain")
    B.advance(6)
    for i, g in enumerate(plain):
        g.advance(6)
        for a, b in zip(B.state(i), g.state()):
            assert np.array_equal(bits(a), bits(b)), i


def sphere_basin(limiter):
    """A 48x24 band of a small sphere (60S-60N), 2 km deep with a 5 m hump, walls
    at the band's edges, periodic in longitude."""
    nx, ny, R = 48, 24, 6.371e6 / 20
    rows = swe_amd.geometry_spherical(nx, ny, R=R, lat0=-60.0, lat1=60.0, omega=7.2921e-5)
    W, H = nx + 4, ny + 4
    X, Y = np.meshgrid(np.arange(W), np.arange(H))
    b = np.full(W * H, -2000.0)
    h = 2000.0 + 5.0 * np.exp(-(((X - W / 3) / 5.0) ** 2 + ((Y - H / 2) / 4.0) ** 2)).ravel()
    z = np.zeros(W * H)
    return dict(nx=nx, ny=ny, rows=rows, R=R, limiter=limiter, state=(b, h, z, z.copy()),
                bc=("periodic", "periodic", "reflect", "reflect"))


@pytest.mark.parametrize("limiters,with_sphere", [(("mc", "minmod", "central"), False),
                                                   (("minmod",), True), (("central",), False)])
def test_batch_limiters_and_row_kinds_equal_plain_runs(limiters, with_sphere):
    """Basins with different slope limiters in one batch (the kernel that switches
    on each task's limiter), and batches whose basins share one (the kernels
    specialised on it: with uniform rows on Cartesian grids, with per-row metrics
    once a spherical basin joins)."""
    specs = workloads.ensemble(nb=6, lo=6, hi=120, seed=29)
    for q, s in enumerate(specs):
        s["limiter"] = limiters[q % len(limiters)]
    say(f"[gpu] batch limiters {limiters}, sphere basin {with_sphere}")
    basins = [dict(nx=s["nx"], ny=s["ny"], manning=s["manning"], cfl=s["cfl"], limiter=s["limiter"],
                   rows=swe_amd.geometry_cartesian(s["nx"], s["ny"], s["dx"])) for s in specs]
    sph = sphere_basin(limiters[0]) if with_sphere else None
    if sph:
        basins.append(dict(nx=sph["nx"], ny=sph["ny"], rows=sph["rows"], sphere=True, R=sph["R"],
                           limiter=sph["limiter"], manning=0.0))
    B = swe_amd.Batch(basins)
    plain = []
    for i, s in enumerate(specs):
        B.load_state(i, *s["state"])
        B.set_bc(i, s["bc"])
        B.set_harmonic(i, 0, amp=[s["amp"]], omega=[s["omega"]], phase=[0.0],
                       uamp=[s["uamp"]], uomega=[s["omega"]], uphase=[0.0])
        g = swe_amd.Solver(s["nx"], s["ny"], swe_amd.geometry_cartesian(s["nx"], s["ny"], s["dx"]),
                           manning=s["manning"], cfl=s["cfl"], limiter=s["limiter"])
        g.load_state(*s["state"])
        g.set_bc(s["bc"])
        g.set_harmonic(0, amp=[s["amp"]], omega=[s["omega"]], phase=[0.0],
                       uamp=[s["uamp"]], uomega=[s["omega"]], uphase=[0.0])
        plain.append(g)
    if sph:
        q = len(specs)
        B.load_state(q, *sph["state"])
        B.set_bc(q, sph["bc"])
        g = swe_amd.Solver(sph["nx"], sph["ny"], sph["rows"], sphere=True, R=sph["R"], manning=0.0,
                           limiter=sph["limiter"])
        g.load_state(*sph["state"])
        g.set_bc(sph["bc"])
        plain.append(g)
    B.set_chunk_rows(16)
    assert B.advance(15) == 15
    for g in plain:
        assert g.advance(15) == 15
    for i, g in enumerate(plain):
        for a, b in zip(B.state(i), g.state()):
            assert np.array_equal(bits(a), bits(b)), i
        assert B.get_scalar(i, swe_amd.S_T) == g.t


def test_rows_made_uneven_after_a_captured_graph():
    """A Cartesian grid's rows are all alike, so its captured step graph runs the
    uniform-row stage kernels; once a row changes, the step must read every row
    again: the run equals a solver built with the uneven rows from the start."""
    s = workloads.ensemble(nb=1, lo=60, hi=90, seed=31)[0]
    nx, ny = s["nx"], s["ny"]
    rows = swe_amd.geometry_cartesian(nx, ny, s["dx"])
    uneven = np.array(rows["dxRow"], dtype=np.float64) * (1.0 + 0.01 * np.sin(np.arange(ny + 4)))
    say(f"[gpu] rows made uneven after capture, {nx}x{ny}")

    def solver(rw):
        g = swe_amd.Solver(nx, ny, rw, manning=s["manning"], cfl=s["cfl"])
        g.load_state(*s["state"])
        g.set_bc(s["bc"])
        g.set_harmonic(0, amp=[s["amp"]], omega=[s["omega"]], phase=[0.0],
                       uamp=[s["uamp"]], uomega=[s["omega"]], uphase=[0.0])
        g.use_graph(True)
        return g

    a = solver(rows)
    assert a.advance(5) == 5                 # graph captured on alike rows
    a.set_row("dxRow", uneven)
    ref_rows = dict(rows)
    ref_rows["dxRow"] = uneven
    b = solver(ref_rows)
    b.load_state(*a.state())
    b.set_time(a.t, 5)
    assert a.advance(7) == 7 and b.advance(7) == 7
    for x, y in zip(a.state(), b.state()):
        assert np.array_equal(bits(x), bits(y))
