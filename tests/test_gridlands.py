"""The gridlands-terrain v1 import (water-simulator_amd/gridlands.py and
js/gridlands.mjs, restating src/gridlands.mjs) on the synthetic export that
tests/golden/make_gridlands.py writes.

CPU: the decoded layers, the exporter's invariants (two lakes perched on
purpose, both found and seated by settle), the refusals, the row-flip probe
and its mutant, the ghost clamp, the sampling at native resolution (bed, surface
and Manning n land on exactly the export's values), and the Python and JS
adapters sampling the same doubles bit for bit.
"""
import base64
import copy
import gzip
import json
import math
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import gridlands as gl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "gridlands_synth.terrain.json.gz")
NODE = shutil.which("node")


@pytest.fixture(scope="module")
def doc():
    with gzip.open(FIXTURE, "rt") as f:
        return json.load(f)


def test_fixture_is_the_generator_output(doc):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_gridlands
    assert make_gridlands.build() == doc


def test_invariants_and_settle(doc):
    t = gl.gridlands_terrain(doc)
    inv = gl.verify_invariants(t)
    assert inv["surface_below_bed"] == 0 and inv["lakes_not_flat"] == 0
    assert 0.10 * inv["cells"] < inv["wet"] < 0.95 * inv["cells"]
    assert inv["worst_ocean_perch"] <= 1e-6
    assert inv["perched_lakes"] == 2 and inv["worst_perch"] == 1.0
    s = gl.gridlands_terrain(doc, settle=True)
    assert s["settled"]["bodies_moved"] == 2 and s["settled"]["worst_drop"] == 1.0
    inv2 = gl.verify_invariants(s)
    assert inv2["perched_lakes"] == 0 and inv2["lakes_not_flat"] == 0
    # settling lowers water only
    assert np.all(s["surfaceM"] <= t["surfaceM"]) and np.any(s["surfaceM"] < t["surfaceM"])


def test_refusals(doc):
    bad = []
    for mutate in (lambda d: d.update(format="other"), lambda d: d.update(version=2),
                   lambda d: d["layers"]["bedM"].update(data=d["layers"]["bedM"]["data"][:1000]),
                   lambda d: d["layers"]["bedM"].update(data=d["layers"]["bedM"]["data"][:1001]),
                   lambda d: d["layers"]["surfaceM"].update(encoding="b64f64le"),
                   lambda d: d["layers"].pop("bedM"),
                   lambda d: d["grid"].update(width=0),
                   lambda d: d["grid"].update(rowOrder="south-first")):
        d = copy.deepcopy(doc)
        mutate(d)
        with pytest.raises(gl.GridlandsError):
            gl.gridlands_terrain(d)
        bad.append(d)
    assert len(bad) == 8
    d = copy.deepcopy(doc)
    d["layers"].pop("manningN")
    t = gl.gridlands_terrain(d)
    assert t["manningN"] is None
    with pytest.raises(gl.GridlandsError):
        gl.manning_field(t)


def test_row_flip_probe_and_its_mutant(doc):
    t = gl.gridlands_terrain(doc)
    W, H, C = t["W"], t["H"], t["cellMeters"]
    x = (np.arange(W) + 0.5) * C
    top = gl.bed_at(t, x, np.full(W, (H - 0.5) * C))
    assert np.array_equal(top, t["bedM"][:W].astype(np.float64))          # row 0 = north = top
    m = gl.gridlands_terrain(doc, flip_rows=False)
    assert not np.allclose(gl.bed_at(m, x, np.full(W, (H - 0.5) * C)), t["bedM"][:W], atol=1e-3)
    assert np.any(t["bedM"][:W] != t["bedM"][(H - 1) * W:])


def test_ghost_clamp(doc):
    t = gl.gridlands_terrain(doc)
    W, H, C = t["W"], t["H"], t["cellMeters"]
    xs = np.array([-1.5, -0.5, W - 0.5, W + 1.5]) * C
    ys = np.array([-1.5, -0.5, H - 0.5, H + 1.5]) * C
    X, Y = np.meshgrid(xs, ys)
    assert np.all(np.isfinite(gl.bed_at(t, X.ravel(), Y.ravel())))
    edge = gl.bed_at(t, 0.5 * C, 0.5 * C)
    assert gl.bed_at(t, -1.5 * C, 0.5 * C) == edge and gl.bed_at(t, -0.5 * C, 0.5 * C) == edge


def test_native_resolution_sampling_is_the_export(doc):
    """At 1:1 the bilinear bed and the nearest surface return the export's values."""
    for settle in (False, True):
        t = gl.gridlands_terrain(doc, settle=settle)
        W, H = t["W"], t["H"]
        b, h = gl.sample_state(t)
        n = gl.manning_field(t)
        inner = lambda a: a.reshape(H + 4, W + 4)[2:-2, 2:-2][::-1].ravel()   # north row first
        assert np.array_equal(inner(b), t["bedM"].astype(np.float64))
        eta = t["surfaceM"].astype(np.float64)
        assert np.array_equal(inner(h), np.maximum(0.0, eta - t["bedM"].astype(np.float64)))
        assert np.array_equal(inner(n), t["manningN"].astype(np.float64))
        assert len(np.unique(n)) >= 3


@pytest.mark.skipif(NODE is None, reason="node is not installed")
@pytest.mark.parametrize("settle", [False, True])
def test_python_and_js_adapters_sample_the_same_doubles(doc, settle):
    t = gl.gridlands_terrain(doc, settle=settle)
    b, h = gl.sample_state(t)
    n = gl.manning_field(t)
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "s.f64")
        subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "gridlands_sample.mjs"), FIXTURE, out,
                        "settle" if settle else "as-is"], check=True, timeout=120)
        js = np.fromfile(out, dtype="<f8")
    N = b.size
    assert js.size == 3 * N
    for a, want in ((js[:N], b), (js[N:2 * N], h), (js[2 * N:], n)):
        assert np.array_equal(a.view(np.uint64), want.view(np.uint64))
