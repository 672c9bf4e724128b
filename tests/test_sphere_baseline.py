"""The reference's recorded spherical results (tools/sphere-baseline.json,
copied to tests/golden/) against this build, with the reference's own
tolerances (tools/verify-sphere.mjs:1136-1198): cell areas within 1e-15 of the
recorded sums; resting-ocean dry cells exactly, max speed and max |eta| within
a factor of 4 of the recorded round-off residuals (exactly 0 where 0 was
recorded); Merian periods within 0.005 percentage points.

CPU: the area rows and the 6- and 4-degree resting oceans on the oracle.  GPU
(marked): every resting-ocean row and every Merian row on the HIP solver, the
Merian run device-resident with the projection recorder, and its 6-degree
series equal to the oracle's host loop bit for bit.
"""
import math

import numpy as np
import pytest

import oracle
import sphere_cases as sc
import swe_amd

BASE = sc.baseline()


def band(got, want):
    if got == 0 or want == 0:
        return got == want
    return 0.25 <= got / want <= 4


def oracle_make(nx, ny, rows, order):
    return oracle.OracleSim(nx, ny, rows, sphere=True, R=sc.R_REF, manning=0.0, cfl=0.45, order=order)


def gpu_make(nx, ny, rows, order):
    return swe_amd.Solver(nx, ny, rows, sphere=True, R=sc.R_REF, manning=0.0, cfl=0.45, order=order)


@pytest.mark.parametrize("row", BASE["areaIdentity"], ids=lambda r: f"{r['deg']}deg_{r['lat0']}_{r['lat1']}")
def test_area_identity_rows(row):
    nx, ny, rows, _ = sc.grid(row["deg"], row["lat0"], row["lat1"])
    A = sc.interior_area(rows, nx, ny)
    assert abs(A - row["area"]) <= 1e-15 * row["area"], (A, row["area"])
    want = 2 * math.pi * sc.R_REF ** 2 * (math.sin(row["lat1"] * sc.D) - math.sin(row["lat0"] * sc.D))
    assert abs((A - want) / want) < 1e-13


def check_rest(got, row):
    assert got["dry"] == row["dry"], (got, row)
    assert band(got["maxSpeed"], row["maxSpeed"]), (got, row)
    assert band(got["maxEta"], row["maxEta"]), (got, row)
    assert got["drift"] <= 1e-13


REST = [r for r in BASE["restingOcean"]]


@pytest.mark.parametrize("row", [r for r in REST if r["deg"] >= 4], ids=lambda r: f"{r['case']}_{r['deg']}deg")
def test_resting_ocean_rows_on_the_oracle(row):
    got = sc.resting_ocean(oracle_make, row["case"], row["deg"], row["order"], row["lat0"], row["lat1"], row["steps"])
    check_rest(got, row)


@pytest.mark.gpu
@pytest.mark.parametrize("row", REST, ids=lambda r: f"{r['case']}_{r['deg']}deg_o{r['order']}_{r['lat0']}")
def test_resting_ocean_rows_on_the_gpu(row):
    got = sc.resting_ocean(gpu_make, row["case"], row["deg"], row["order"], row["lat0"], row["lat1"], row["steps"])
    print(f"  {row['case']} {row['deg']}deg: {got} vs recorded {row['maxSpeed']:.3e} {row['maxEta']:.3e}", flush=True)
    check_rest(got, row)


def gpu_merian(n, deg):
    m = sc.merian_problem(n, deg)
    nx, ny = m["nx"], m["ny"]
    s = gpu_make(nx, ny, m["rows"], 2)
    try:
        z = np.zeros_like(m["h"])
        s.load_state(m["b"], m["h"], z, z)
        s.set_bc(("periodic", "periodic", "reflect", "reflect"))
        dt = s.max_dt()
        nsteps = sc.fixed_steps(dt, 1.3 * m["t_closed"])
        s.set_projection(m["w"], nsteps)
        s.autotune(reps=2)
        s.use_graph(True)
        assert s.advance(nsteps, swe_amd.DT_FIXED, dt) == nsteps
        times, vals = s.projection_series()
    finally:
        s.close()
    c0 = sc.projection(m["w"], m["b"], m["h"], nx, ny)
    return m, dt, times, vals, c0


@pytest.mark.gpu
@pytest.mark.parametrize("row", BASE["merian"], ids=lambda r: f"n{r['n']}_{r['deg']}deg")
def test_merian_rows_on_the_gpu(row):
    m, dt, times, vals, c0 = gpu_merian(row["n"], row["deg"])
    T, err, nz = sc.merian_period(m["t_closed"], times, vals, c0)
    print(f"  n={row['n']} {row['deg']}deg: errPct {err:.5f} vs recorded {row['errPct']:.5f} ({nz} zeros, "
          f"{len(vals)} steps)", flush=True)
    assert abs(err - row["errPct"]) <= 0.005
    assert abs(m["t_closed"] / 3600 - row["closed_h"]) <= 1e-12 * row["closed_h"]


@pytest.mark.gpu
def test_merian_series_equals_the_oracle_host_loop():
    n, deg = 2, 6
    m, dt, times, vals, c0 = gpu_merian(n, deg)
    o = oracle_make(m["nx"], m["ny"], m["rows"], 2)
    z = np.zeros_like(m["h"])
    o.load_state(m["b"], m["h"], z, z)
    o.set_bc(("periodic", "periodic", "reflect", "reflect"))
    assert o.max_dt() == dt
    ot, ov = [], []
    while o.t < 1.3 * m["t_closed"]:
        o.step(dt)
        ot.append(o.t)
        ov.append(sc.projection(m["w"], o.b, o.h, m["nx"], m["ny"]))
    assert np.array_equal(np.array(ot), times) and np.array_equal(np.array(ov), vals)
