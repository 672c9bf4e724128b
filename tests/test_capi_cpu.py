"""CPU-side checks of the C ABI library and the host mirror (no GPU calls).

- libswe_amd.so loads and exports every function include/swe_amd.h declares;
- swe_js_pow (host build of the device pow) equals the oracle's pow bit for bit
  and reproduces the engine's eps4 for the reference's default minDepth;
- the host restatement of geometry.mjs matches the reference's recorded rows:
  Cartesian bit for bit, spherical to 4 ulp (transcendentals from libm, not V8).
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import goldens
import oracle
import swe_amd

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "swe_amd.h")


def declared_functions():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(swe_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(swe_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_build_info():
    assert b"gfx950" in swe_amd.lib().swe_build_info()


def test_js_pow_host_matches_oracle_and_engine():
    L = swe_amd.lib()
    rng = np.random.default_rng(0)
    for x in 10.0 ** rng.uniform(-3.5, 4.5, 5000):
        assert L.swe_js_pow(x, 4.0 / 3.0) == oracle.js_pow(x, 4.0 / 3.0)
    fx = goldens.load("vel_probe")
    assert swe_amd.js_pow4(1e-3) == goldens.hexf(fx["eps4"])


def test_cartesian_rows_bit_exact():
    fx = goldens.load("lake_island")
    p = fx["params"]
    rows = swe_amd.geometry_cartesian(p["nx"], p["ny"], goldens.hexf(p["dx"]), goldens.hexf(p["dy"]))
    ref = goldens.rows(fx)
    for r in swe_amd.ROWS:
        assert np.array_equal(rows[r], ref[r]), r


@pytest.mark.parametrize("name", ["sphere_pole", "sphere_band_o1"])
def test_spherical_rows_close(name):
    fx = goldens.load(name)
    p = fx["params"]
    sp = p["sphere"]
    rows = swe_amd.geometry_spherical(p["nx"], p["ny"], R=sp["R"], lat0=sp["lat0"], lat1=sp["lat1"],
                                      omega=sp["omega"])
    ref = goldens.rows(fx)
    for r in swe_amd.ROWS:
        a, b = rows[r], ref[r]
        assert np.array_equal(np.isinf(a), np.isinf(b)), r
        fin = np.isfinite(b)
        err = np.abs(a[fin] - b[fin])
        tol = 4 * np.spacing(np.maximum(np.abs(b[fin]), 1e-300))
        assert np.all(err <= tol), (r, err.max())
    # the pole face is closed exactly (geometry.mjs:154-157)
    if sp["lat0"] == -90:
        H = p["ny"] + 4
        assert rows["dyRowS"][2] == math.inf and rows["dyRowN"][H - 3] == math.inf


def test_spherical_area_identity():
    rows = swe_amd.geometry_spherical(72, 36, R=1.0, omega=0.0)
    total = 72 * rows["area"][2:-2].sum()
    assert abs(total - 4 * math.pi) < 1e-13


def test_geometry_errors_match_reference():
    with pytest.raises(ValueError):
        swe_amd.geometry_spherical(9, 10)            # odd nx reaching a pole
    with pytest.raises(ValueError):
        swe_amd.geometry_spherical(8, 10, lat0=10, lat1=5)


def test_adjoint_memory_plan():
    # swe_adj_plan is host arithmetic: 3 W H doubles per checkpoint, two replay
    # stores of `inner` steps, the pass slots and scratch.  The 24-hour adjoint of
    # one 2160x1080 sphere member (about 1.48e6 steps at dt = 0.0585 s) needs two
    # levels to fit in one MI355X's 288 GB of HBM.
    ncells = (2160 + 4) * (1080 + 4)
    nsteps = int(round(86400 / 0.0585))
    k1 = int(round(nsteps ** 0.5))
    one = swe_amd.adjoint_plan_bytes(nsteps, ncells, k1, k1)
    k2 = int(round(nsteps ** (1 / 3)))
    two = swe_amd.adjoint_plan_bytes(nsteps, ncells, k2, k2 * k2)
    assert one > 288e9 > two
    # more checkpoints cost 3 W H doubles each
    a = swe_amd.adjoint_plan_bytes(1000, 100, 10, 10)
    b = swe_amd.adjoint_plan_bytes(1000, 100, 10, 20)
    assert a - b == pytest.approx(8.0 * 100 * 3 * (100 - 50) - 8.0 * 100 * 3 * 2)
    with pytest.raises(swe_amd.SweError):
        swe_amd.adjoint_plan_bytes(1000, 100, 10, 15)      # outer not a multiple of inner
