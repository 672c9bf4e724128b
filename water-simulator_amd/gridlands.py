"""Terrain import from gridlands' `gridlands-terrain` v1 export for the Python
host: the data format on the input side of the solver (the reference's
src/gridlands.mjs has the JavaScript side; js/gridlands.mjs is this design for
the Node drop-in).

A document is a W x H grid at `cellMeters`, row 0 = north, whose layers `bedM`,
`surfaceM` (standing water) and optionally `manningN` are float32
little-endian arrays in base64.  The module works in whole-array passes:

1. label_water: the wet cells (surface above bed) split into 4-connected
   waterbodies, numbered in order of each body's first cell in raster order
   (scipy.ndimage.label's two-pass labelling numbers them that way);
2. body_stats: per body its first cell and level (that cell's surface), size,
   largest departure of a surface from the level, and the lowest bed among the
   dry cells bordering it (the spill level), as per-label reductions;
3. the invariant report and the settle step read those per-body arrays.

Sampling then fills the solver's W*H bed, depth and Manning arrays at the
constructor's cell centres, two ghost rings included (swe.mjs:222-231: x = (i -
ng + 0.5) dx, y = (j - ng + 0.5) dy), with the same doubles as the Node adapter.

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
from scipy import ndimage

# 40 cells per wavelength: the shortest wave the grid carries without visible
# numerical dispersion, as the reference states it for its maps
CELLS_PER_WAVELENGTH = 40
# a body whose level stands more than this above its spill level is perched
PERCH_TOLERANCE = 1e-6
_CROSS = np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]])


class GridlandsError(ValueError):
    pass


def _decode_layer(layers, name, cells, required):
    """The float32 values of one layer: `b64f32le`, exactly `cells` of them."""
    layer = layers.get(name)
    if not layer:
        if required:
            raise GridlandsError(f"gridlands: layer {name} is missing")
        return None
    if layer.get("encoding") != "b64f32le":
        raise GridlandsError(f"gridlands: layer {name} is encoded as {layer.get('encoding')}; only b64f32le is read")
    try:
        raw = base64.b64decode(layer["data"], validate=True)
    except (binascii.Error, ValueError, TypeError, KeyError) as e:
        raise GridlandsError(f"gridlands: layer {name} is not valid base64 ({e})") from None
    if len(raw) != 4 * cells:
        raise GridlandsError(f"gridlands: layer {name} holds {len(raw)} bytes where {cells} cells need {4 * cells}")
    return np.frombuffer(raw, dtype="<f4").copy()


def label_water(bed, surface, W, H):
    """Waterbody labels over the W*H cells: 0, 1, ... in order of each body's
    first cell (raster order), -1 on dry cells; and the number of bodies."""
    wet = (surface.astype(np.float64) - bed.astype(np.float64)) > 0
    lab, count = ndimage.label(wet.reshape(H, W), structure=_CROSS)
    return lab.ravel().astype(np.int64) - 1, int(count)


def body_stats(bed, surface, W, H):
    """Per-body first cell, size, level, spread and spill level (dict of arrays)."""
    label, count = label_water(bed, surface, W, H)
    s64 = surface.astype(np.float64)
    b64 = bed.astype(np.float64)
    wet = label >= 0
    cells = np.flatnonzero(wet)
    lw = label[cells]
    first = np.full(count, -1, np.int64)
    # the first cell of a body: the smallest index carrying its label
    first[lw[::-1]] = cells[::-1]
    level = s64[first] if count else np.zeros(0)
    size = np.bincount(lw, minlength=count)
    spread = np.zeros(count)
    np.maximum.at(spread, lw, np.abs(s64[cells] - level[lw]))
    ring = np.full(count, math.inf)
    grid = label.reshape(H, W)
    for dr, dc in ((0, -1), (0, 1), (-1, 0), (1, 0)):
        # wet cell (r, c) next to a dry neighbour (r + dr, c + dc) inside the grid
        r0, r1 = max(0, -dr), H - max(0, dr)
        c0, c1 = max(0, -dc), W - max(0, dc)
        here = grid[r0:r1, c0:c1]
        there = grid[r0 + dr:r1 + dr, c0 + dc:c1 + dc]
        m = (here >= 0) & (there < 0)
        nb = b64.reshape(H, W)[r0 + dr:r1 + dr, c0 + dc:c1 + dc]
        np.minimum.at(ring, here[m], nb[m])
    return dict(label=label, count=count, first=first, size=size, level=level, spread=spread, ring_min=ring)


def waterbodies(bed, surface, W, H):
    """The connected waterbodies in order of their first cell: level (that
    cell's surface), cells (raster order), spread (largest departure of a
    surface from the level), ring_min (the spill level: lowest bed on the ring
    of dry neighbours, inf when none), perch (level - ring_min), is_ocean
    (level exactly 0)."""
    s = body_stats(bed, surface, W, H)
    order = np.argsort(s["label"], kind="stable")
    cut = np.cumsum(np.concatenate(([np.count_nonzero(s["label"] < 0)], s["size"])))
    out = []
    for b in range(s["count"]):
        lv, rm = float(s["level"][b]), float(s["ring_min"][b])
        out.append(dict(level=lv, cells=order[cut[b]:cut[b + 1]], spread=float(s["spread"][b]), ring_min=rm,
                        perch=lv - rm, is_ocean=lv == 0))
    return out


def verify_invariants(t):
    """The exporter's guarantees re-derived from the decoded layers, and the one
    the solver needs that the format does not promise (no lake above the lowest
    bed of its own shoreline).  A report; raises nothing."""
    bed, surf, W, H = t["bedM"], t["surfaceM"], t["W"], t["H"]
    depth = surf.astype(np.float64) - bed.astype(np.float64)
    s = body_stats(bed, surf, W, H)
    perch = s["level"] - s["ring_min"]
    ocean = s["level"] == 0
    lake = ~ocean
    perched = lake & (perch > PERCH_TOLERANCE)
    uneven = s["spread"] > 0
    below = depth < 0
    return dict(
        cells=W * H, wet=int((depth > 0).sum()), ocean_cells=int(((surf == 0) & (bed < 0)).sum()),
        surface_below_bed=int(below.sum()), worst_below_bed=float(min(0.0, depth.min())),
        bodies=s["count"], lake_bodies=int(lake.sum()), lakes_not_flat=int(uneven.sum()),
        worst_lake_spread=float(s["spread"][uneven].max()) if uneven.any() else 0.0,
        largest_body=int(s["size"].max()) if s["count"] else 0,
        perched_lakes=int(perched.sum()), worst_perch=float(perch[perched].max()) if perched.any() else 0.0,
        perched_cells=int(s["size"][perched].sum()),
        perched_volume_cells=float(np.sum(s["size"][perched] * perch[perched])),
        worst_ocean_perch=float(perch[ocean].max()) if ocean.any() else -math.inf)


def settle_lakes(bed, surface, W, H):
    """Seat every body standing more than 1e-6 m above its spill level at that
    level, in place: its cells get surface max(bed, spill), so cells whose bed
    is above the spill level fall dry.  Returns what moved."""
    s = body_stats(bed, surface, W, H)
    drop = s["level"] - s["ring_min"]
    perched = drop > PERCH_TOLERANCE
    lab = s["label"]
    cells = np.flatnonzero((lab >= 0) & perched[np.maximum(lab, 0)])
    surface[cells] = np.maximum(bed[cells], s["ring_min"][lab[cells]].astype(np.float32))
    moved = [dict(frm=float(s["level"][b]), to=float(s["ring_min"][b]), drop=float(drop[b]), cells=int(s["size"][b]))
             for b in np.flatnonzero(perched)]
    return dict(bodies_moved=len(moved), worst_drop=max((m["drop"] for m in moved), default=0.0), moved=moved)


_DOCUMENT_CHECKS = (
    lambda d: None if isinstance(d, dict) else "not a document",
    lambda d: None if d.get("format") == "gridlands-terrain" and d.get("version") == 1
    else f"{d.get('format')} v{d.get('version')} is not read here; this adapter reads gridlands-terrain v1 only",
    lambda d: None if all(isinstance((d.get("grid") or {}).get(k), int) and d["grid"][k] > 0
                          for k in ("width", "height")) and ((d.get("grid") or {}).get("cellMeters") or 0) > 0
    else f"bad grid {json.dumps(d.get('grid'))}",
    lambda d: None if d["grid"].get("rowOrder") in (None, "", "north-first")
    else f"rows are {d['grid']['rowOrder']}; only north-first is read",
)


def gridlands_terrain(doc, flip_rows=True, settle=False):
    """A parsed gridlands-terrain v1 document as solver inputs: the decoded
    layers, the domain, the Manning scalar the export suggests (0.03 when none),
    the provenance it carries.  flip_rows (default) puts the export's north row
    at the top of the simulation; settle seats perched lakes."""
    for check in _DOCUMENT_CHECKS:
        msg = check(doc)
        if msg:
            raise GridlandsError(f"gridlands: {msg}")
    g = doc["grid"]
    W, H, C = g["width"], g["height"], float(g["cellMeters"])
    layers = doc.get("layers") or {}
    bed = _decode_layer(layers, "bedM", W * H, True)
    surf = _decode_layer(layers, "surfaceM", W * H, True)
    mann = _decode_layer(layers, "manningN", W * H, False)
    settled = settle_lakes(bed, surf, W, H) if settle else None
    stats, meta = doc.get("stats") or {}, doc.get("meta") or {}
    suggested = stats.get("suggestedManningScalar")
    return dict(W=W, H=H, cellMeters=C, bedM=bed, surfaceM=surf, manningN=mann, flip_rows=flip_rows,
                domain=dict(nx=W, ny=H, dx=C, dy=C), manning=0.03 if suggested is None else suggested,
                settled=settled, stats=stats, meta=meta,
                label=f"gridlands {meta.get('preset')}/{meta.get('seed')} #{meta.get('mapId')}",
                shortest_honest_wavelength=CELLS_PER_WAVELENGTH * C)


def _grid_coords(t, x, y):
    """Fractional export column and row of simulation points, clamped to the
    grid; with flip_rows the export's row 0 (north) is the top of the domain."""
    W, H, C = t["W"], t["H"], t["cellMeters"]
    fc = np.minimum(W - 1, np.maximum(0.0, np.asarray(x, np.float64) / C - 0.5))
    fr = np.minimum(H - 1, np.maximum(0.0, np.asarray(y, np.float64) / C - 0.5))
    return fc, ((H - 1) - fr if t["flip_rows"] else fr)


def _js_round(v):
    """Math.round: the nearest integer, halves toward +infinity."""
    f = np.floor(v)
    return (f + (v - f >= 0.5)).astype(np.int64)


def bed_at(t, x, y):
    """The bed between the four surrounding export cells, bilinearly."""
    W, H, layer = t["W"], t["H"], t["bedM"].astype(np.float64)
    fc, fr = _grid_coords(t, x, y)
    c0, r0 = np.floor(fc).astype(np.int64), np.floor(fr).astype(np.int64)
    c1, r1 = np.minimum(W - 1, c0 + 1), np.minimum(H - 1, r0 + 1)
    tc, tr = fc - c0, fr - r0
    north = layer[r0 * W + c0] * (1 - tc) + layer[r0 * W + c1] * tc
    south = layer[r1 * W + c0] * (1 - tc) + layer[r1 * W + c1] * tc
    return north * (1 - tr) + south * tr


def _nearest(t, layer, x, y):
    """The nearest export cell's value (Math.round on both coordinates)."""
    fc, fr = _grid_coords(t, x, y)
    return layer.astype(np.float64)[_js_round(fr) * t["W"] + _js_round(fc)]


def eta_at(t, x, y):
    """The surface of the nearest export cell, so a flat waterbody stays flat."""
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
