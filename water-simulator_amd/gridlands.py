"""Terrain import from gridlands' `gridlands-terrain` v1 export (src/gridlands.mjs
of the reference) for the Python host: decode the float32 layers, check the
exporter's invariants, optionally seat perched lakes, and sample the bed,
initial depth and Manning field at the solver's cell centres (two ghost rings
included) exactly as the reference's constructor samples its closures
(swe.mjs:222-231: x = (i - ng + 0.5) dx, y = (j - ng + 0.5) dy, row by row).

    doc = json.load(open("map.terrain.json"))
    t = gridlands.gridlands_terrain(doc, settle=True)
    b, h = gridlands.sample_state(t)            # W*H arrays for Solver.load_state
    n = gridlands.manning_field(t)              # W*H, for Solver.set_manning_field
"""
import base64
import binascii
import json
import math
import sys

import numpy as np


class GridlandsError(ValueError):
    pass


def _layer(layer, n, name):
    """One b64f32le layer as n float32 values (gridlands.mjs:86-110)."""
    if not layer:
        raise GridlandsError(f"gridlands: layer {name} is missing")
    if layer.get("encoding") != "b64f32le":
        raise GridlandsError(f"gridlands: layer {name} is encoded as {layer.get('encoding')}; only b64f32le is read")
    try:
        raw = base64.b64decode(layer["data"], validate=True)
    except (binascii.Error, ValueError, TypeError) as e:
        raise GridlandsError(f"gridlands: layer {name} is not valid base64 ({e})") from None
    if len(raw) != 4 * n:
        raise GridlandsError(f"gridlands: layer {name} holds {len(raw)} bytes; the grid needs {4 * n} ({n} cells x 4)")
    return np.frombuffer(raw, dtype="<f4").copy()


def waterbodies(bed, surface, W, H):
    """Connected wet bodies in order of their first cell: level, cells, spread,
    ring minimum (spill level), perch, ocean flag (gridlands.mjs:119-155)."""
    n = W * H
    wet = (surface.astype(np.float64) - bed.astype(np.float64)) > 0
    seen = np.zeros(n, bool)
    bodies = []
    for start in np.flatnonzero(wet):
        if seen[start]:
            continue
        level = float(surface[start])
        cells, ring, spread = [], set(), 0.0
        seen[start] = True
        todo = [int(start)]
        while todo:
            k = todo.pop()
            cells.append(k)
            s = float(surface[k])
            if s != level:
                spread = max(spread, abs(s - level))
            col, row = k % W, k // W
            for m, ok in ((k - 1, col > 0), (k + 1, col < W - 1), (k - W, row > 0), (k + W, row < H - 1)):
                if not ok:
                    continue
                if not wet[m]:
                    ring.add(m)
                elif not seen[m]:
                    seen[m] = True
                    todo.append(m)
        ring_min = min((float(bed[m]) for m in ring), default=math.inf)
        bodies.append(dict(level=level, cells=cells, spread=spread, ring_min=ring_min, perch=level - ring_min,
                           is_ocean=level == 0))
    return bodies


def verify_invariants(t):
    """The exporter's guarantees re-derived (gridlands.mjs:157-206)."""
    bed, surf, W, H = t["bedM"], t["surfaceM"], t["W"], t["H"]
    d = surf.astype(np.float64) - bed.astype(np.float64)
    bodies = waterbodies(bed, surf, W, H)
    lakes = [b for b in bodies if not b["is_ocean"]]
    uneven = [b for b in bodies if b["spread"] > 0]
    perched = sorted((b for b in lakes if b["perch"] > 1e-6), key=lambda b: -b["perch"])
    oceans = [b["perch"] for b in bodies if b["is_ocean"]]
    return dict(
        cells=W * H, wet=int((d > 0).sum()), ocean_cells=int(((surf == 0) & (bed < 0)).sum()),
        surface_below_bed=int((d < 0).sum()), worst_below_bed=float(min(0.0, d.min())),
        bodies=len(bodies), lake_bodies=len(lakes), lakes_not_flat=len(uneven),
        worst_lake_spread=max((b["spread"] for b in uneven), default=0.0),
        largest_body=max((len(b["cells"]) for b in bodies), default=0),
        perched_lakes=len(perched), worst_perch=perched[0]["perch"] if perched else 0.0,
        perched_cells=sum(len(b["cells"]) for b in perched),
        perched_volume_cells=sum(len(b["cells"]) * b["perch"] for b in perched),
        worst_ocean_perch=max(oceans, default=-math.inf))


def settle_lakes(bed, surface, W, H):
    """Lower every body more than 1e-6 m above its spill level to it, in place
    (gridlands.mjs:208-219)."""
    moved = []
    for b in waterbodies(bed, surface, W, H):
        if not b["perch"] > 1e-6:
            continue
        idx = np.asarray(b["cells"])
        surface[idx] = np.maximum(bed[idx], np.float32(b["ring_min"]))
        moved.append(dict(frm=b["level"], to=b["ring_min"], drop=b["perch"], cells=len(idx)))
    return dict(bodies_moved=len(moved), worst_drop=max((m["drop"] for m in moved), default=0.0), moved=moved)


def gridlands_terrain(doc, flip_rows=True, settle=False):
    """A parsed gridlands-terrain v1 document as solver inputs (gridlands.mjs:221-302)."""
    if not isinstance(doc, dict):
        raise GridlandsError("gridlands: not a document")
    if doc.get("format") != "gridlands-terrain" or doc.get("version") != 1:
        raise GridlandsError(f"gridlands: {doc.get('format')} v{doc.get('version')} is not read here; "
                             "this adapter reads gridlands-terrain v1 only")
    g = doc.get("grid") or {}
    W, H, C = g.get("width"), g.get("height"), g.get("cellMeters")
    if not (isinstance(W, int) and isinstance(H, int) and W > 0 and H > 0 and C and C > 0):
        raise GridlandsError(f"gridlands: bad grid {json.dumps(g)}")
    if g.get("rowOrder") and g["rowOrder"] != "north-first":
        raise GridlandsError(f"gridlands: rows are {g['rowOrder']}; only north-first is read")
    n = W * H
    layers = doc.get("layers") or {}
    bed = _layer(layers.get("bedM"), n, "bedM")
    surf = _layer(layers.get("surfaceM"), n, "surfaceM")
    mann = _layer(layers["manningN"], n, "manningN") if layers.get("manningN") else None
    settled = settle_lakes(bed, surf, W, H) if settle else None
    stats, meta = doc.get("stats") or {}, doc.get("meta") or {}
    sug = stats.get("suggestedManningScalar")
    return dict(W=W, H=H, cellMeters=float(C), bedM=bed, surfaceM=surf, manningN=mann, flip_rows=flip_rows,
                domain=dict(nx=W, ny=H, dx=float(C), dy=float(C)), manning=0.03 if sug is None else sug,
                settled=settled, stats=stats, meta=meta,
                label=f"gridlands {meta.get('preset')}/{meta.get('seed')} #{meta.get('mapId')}",
                shortest_honest_wavelength=40 * C)


def _coords(t, x, y):
    """Fractional export column and row of simulation points, clamped."""
    W, H, C = t["W"], t["H"], t["cellMeters"]
    fc = np.minimum(W - 1, np.maximum(0.0, x / C - 0.5))
    fr = np.minimum(H - 1, np.maximum(0.0, y / C - 0.5))
    if t["flip_rows"]:
        fr = (H - 1) - fr
    return fc, fr


def _js_round(v):
    """Math.round: the nearest integer, halves toward +infinity."""
    f = np.floor(v)
    return (f + (v - f >= 0.5)).astype(np.int64)


def bed_at(t, x, y):
    """The bilinear bed sampler of gridlands.mjs:259-268, vectorised."""
    W, H, bm = t["W"], t["H"], t["bedM"].astype(np.float64)
    fc, fr = _coords(t, np.asarray(x, np.float64), np.asarray(y, np.float64))
    c0, r0 = np.floor(fc).astype(np.int64), np.floor(fr).astype(np.int64)
    c1, r1 = np.minimum(W - 1, c0 + 1), np.minimum(H - 1, r0 + 1)
    tc, tr = fc - c0, fr - r0
    v00, v10 = bm[r0 * W + c0], bm[r0 * W + c1]
    v01, v11 = bm[r1 * W + c0], bm[r1 * W + c1]
    return (v00 * (1 - tc) + v10 * tc) * (1 - tr) + (v01 * (1 - tc) + v11 * tc) * tr


def _nearest(t, layer, x, y):
    W = t["W"]
    fc, fr = _coords(t, np.asarray(x, np.float64), np.asarray(y, np.float64))
    return layer.astype(np.float64)[_js_round(fr) * W + _js_round(fc)]


def eta_at(t, x, y):
    """The nearest-cell surface sampler of gridlands.mjs:278."""
    return _nearest(t, t["surfaceM"], x, y)


def cell_centres(t, ng=2):
    """x, y of every cell of the solver's W*H grid, ghosts included (swe.mjs:222-227, :251)."""
    nx, ny, dx, dy = t["domain"]["nx"], t["domain"]["ny"], t["domain"]["dx"], t["domain"]["dy"]
    i = np.arange(nx + 2 * ng) - ng
    j = np.arange(ny + 2 * ng) - ng
    X = np.broadcast_to((i + 0.5) * dx, (ny + 2 * ng, nx + 2 * ng))
    Y = np.broadcast_to(((j + 0.5) * dy)[:, None], (ny + 2 * ng, nx + 2 * ng))
    return X.ravel(), Y.ravel()


def sample_state(t, ng=2):
    """(b, h) over the solver's W*H cells as the reference constructor fills them:
    b = bed(x, y), h = max(0, eta0(x, y) - b)."""
    x, y = cell_centres(t, ng)
    b = bed_at(t, x, y)
    return b, np.maximum(0.0, eta_at(t, x, y) - b)


def manning_field(t, ng=2):
    """The per-cell Manning layer at the solver's W*H cells (nearest export cell,
    the mapping tools/verify-gridlands.mjs:332-333 uses)."""
    if t["manningN"] is None:
        raise GridlandsError("gridlands: the export carries no manningN layer")
    x, y = cell_centres(t, ng)
    return _nearest(t, t["manningN"], x, y)


def load(path, **kw):
    with open(path) as f:
        return gridlands_terrain(json.load(f), **kw)


if __name__ == "__main__":
    t = load(sys.argv[1])
    print(t["label"], json.dumps(verify_invariants(t)))
