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


def _flood_fill_bodies(bed, surf, W, H):
    """Brute force: breadth-first flood fill from each unvisited wet cell in index
    order; per body its cells (sorted), level (first cell's surface), spread and
    the lowest bed of its dry 4-neighbours."""
    wet = (surf.astype(np.float64) - bed.astype(np.float64)) > 0
    seen = np.zeros(W * H, bool)
    out = []
    for k0 in range(W * H):
        if seen[k0] or not wet[k0]:
            continue
        seen[k0] = True
        queue, cells, ring = [k0], [], []
        while queue:
            k = queue.pop(0)
            cells.append(k)
            r, c = divmod(k, W)
            for rr, cc in ((r, c - 1), (r, c + 1), (r - 1, c), (r + 1, c)):
                if 0 <= rr < H and 0 <= cc < W:
                    m = rr * W + cc
                    if not wet[m]:
                        ring.append(float(bed[m]))
                    elif not seen[m]:
                        seen[m] = True
                        queue.append(m)
        level = float(surf[k0])
        out.append(dict(level=level, cells=sorted(cells), ring_min=min(ring, default=math.inf),
                        spread=max(abs(float(surf[k]) - level) for k in cells)))
    return out


def _random_grids(seed, n=12):
    rng = np.random.default_rng(seed)
    grids = []
    for q in range(n):
        W, H = int(rng.integers(1, 30)), int(rng.integers(1, 25))
        bed = rng.normal(0, 1, W * H).astype(np.float32)
        # lakes at a few shared levels (some perched, some uneven) over random beds
        levels = rng.choice([-0.5, 0.0, 0.3, 0.8], W * H)
        surf = np.where(rng.random(W * H) < 0.6, np.maximum(levels, bed), bed).astype(np.float32)
        if q % 3 == 0:
            surf = np.where(rng.random(W * H) < 0.1, surf + np.float32(0.01), surf).astype(np.float32)
        grids.append((W, H, bed, surf))
    return grids


@pytest.mark.skipif(NODE is None, reason="node is not installed")
def test_waterbodies_against_brute_force_flood_fill():
    """Component labelling (both hosts) against a plain breadth-first flood fill on
    random grids, single rows and columns included; then the invariant reports and
    the settle step of the two hosts against each other."""
    grids = _random_grids(7)
    with tempfile.TemporaryDirectory() as d:
        inp, out = os.path.join(d, "in.json"), os.path.join(d, "out.json")
        with open(inp, "w") as f:
            json.dump({"grids": [dict(W=W, H=H, bed=b.tolist(), surface=s.tolist()) for W, H, b, s in grids]}, f)
        subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "gridlands_bodies.mjs"), inp, out], check=True,
                       timeout=120)
        with open(out) as f:
            js = json.load(f)
    for (W, H, bed, surf), j in zip(grids, js):
        want = _flood_fill_bodies(bed, surf, W, H)
        py = gl.waterbodies(bed, surf, W, H)
        assert len(py) == len(want) == len(j["bodies"])
        for w, p, q in zip(want, py, j["bodies"]):
            assert list(p["cells"]) == w["cells"] == q["cells"]
            assert p["level"] == w["level"] == q["level"]
            assert p["spread"] == w["spread"] == q["spread"]
            assert p["ring_min"] == w["ring_min"] == (math.inf if q["ringMin"] is None else q["ringMin"])
        inv = gl.verify_invariants(dict(bedM=bed, surfaceM=surf, W=W, H=H))
        for k_py, k_js in (("wet", "wet"), ("surface_below_bed", "surfaceBelowBed"), ("bodies", "bodies"),
                           ("lake_bodies", "lakeBodies"), ("lakes_not_flat", "lakesNotFlat"),
                           ("perched_lakes", "perchedLakes"), ("perched_cells", "perchedCells"),
                           ("largest_body", "largestBody"), ("worst_perch", "worstPerch")):
            assert inv[k_py] == j["inv"][k_js], (k_py, inv[k_py], j["inv"][k_js])
        s2 = surf.copy()
        moved = gl.settle_lakes(bed, s2, W, H)
        assert moved["bodies_moved"] == j["moved"]
        assert np.array_equal(s2, np.asarray(j["settledSurface"], dtype=np.float32))
        # after settling nothing is perched any more
        assert gl.verify_invariants(dict(bedM=bed, surfaceM=s2, W=W, H=H))["perched_lakes"] == 0
