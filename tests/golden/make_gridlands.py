"""Writes a synthetic gridlands-terrain v1 export for the terrain-import tests:

    python tests/golden/make_gridlands.py   ->  tests/golden/gridlands_synth.terrain.json.gz

No gridlands export ships with the reference (tools/verify-gridlands.mjs reads
one from its author's disk and skips without it), so this builds one in the
format src/gridlands.mjs reads: 120 x 90 cells at 50 m, row 0 = north, float32
little-endian base64 layers bedM, surfaceM and manningN, meta and stats.  The
map: an ocean seated at 0 m on the west, land rising east with hills, and five
depressions carved into the land, filled to their spill level by a priority
flood (Barnes et al. 2014), the way gridlands seats its lakes (four hold
water); two of the lakes are
then raised above their spill level (1 m and 0.5 m), the exporter defect that
src/gridlands.mjs's perched-lake check and settleLakes() exist for.  Beds are
multiples of 0.25 m, so every wet cell is at least 0.25 m deep.  Roughness:
water 0.025, land 0.035, a forest 0.10, a playa 0.020.  Deterministic
(SplitMix64, seed 7).
"""
import base64
import gzip
import heapq
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "gridlands_synth.terrain.json.gz")
W, H, C = 120, 90, 50.0


def splitmix64(seed, n):
    out = np.empty(n)
    x = seed & 0xFFFFFFFFFFFFFFFF
    for i in range(n):
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        out[i] = ((z ^ (z >> 31)) >> 11) / float(1 << 53)
    return out


def priority_flood(bed, outlet):
    """Water surface of every cell once depressions fill to their spill level:
    a min-heap flood inward from the outlets (Barnes, Lehman & Mulla 2014)."""
    filled = bed.copy()
    done = np.zeros(bed.shape, bool)
    heap = []
    for r, c in zip(*np.nonzero(outlet)):
        heapq.heappush(heap, (filled[r, c], r, c))
        done[r, c] = True
    while heap:
        z, r, c = heapq.heappop(heap)
        for rr, cc in ((r - 1, c), (r + 1, c), (r, c - 1), (r, c + 1)):
            if 0 <= rr < bed.shape[0] and 0 <= cc < bed.shape[1] and not done[rr, cc]:
                done[rr, cc] = True
                filled[rr, cc] = max(bed[rr, cc], z)
                heapq.heappush(heap, (filled[rr, cc], rr, cc))
    return filled


def build():
    u = splitmix64(7, 64)
    row, col = np.mgrid[0:H, 0:W]              # row 0 = north
    x = (col + 0.5) * C
    y = (H - row - 0.5) * C                    # metres north of the south edge
    coast = 30 * C + 4 * C * np.sin(y / 700.0)
    bed = np.where(x < coast, -40.0 + 39.0 * x / coast, 0.5 + 0.008 * (x - coast))
    for q in range(6):                         # hills
        cx, cy = (0.45 + 0.5 * u[q]) * W * C, u[10 + q] * H * C
        bed += (x > coast) * (25 + 40 * u[20 + q]) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (900.0 + 400 * u[30 + q]) ** 2)
    bowls = [(62, 20, 12.0, 5.0), (85, 60, 16.0, 6.5), (100, 25, 14.0, 5.5), (70, 72, 11.0, 4.5), (108, 78, 15.0, 5.0)]
    for (bc, br, depth, rad) in bowls:         # closed depressions on the land
        bed -= depth * np.exp(-((col - bc) ** 2 + (row - br) ** 2) / rad ** 2)
    bed = np.round(bed * 4.0) / 4.0            # multiples of 0.25 m
    ocean_bed = bed < 0
    # the ocean: sub-zero cells connected to the west edge
    ocean = np.zeros_like(ocean_bed)
    stack = [(r, 0) for r in range(H) if ocean_bed[r, 0]]
    for r, c in stack:
        ocean[r, c] = True
    while stack:
        r, c = stack.pop()
        for rr, cc in ((r - 1, c), (r + 1, c), (r, c - 1), (r, c + 1)):
            if 0 <= rr < H and 0 <= cc < W and ocean_bed[rr, cc] and not ocean[rr, cc]:
                ocean[rr, cc] = True
                stack.append((rr, cc))
    edge = np.zeros_like(ocean)
    edge[0, :] = edge[-1, :] = edge[:, 0] = edge[:, -1] = True
    land_bed = np.where(ocean, 0.0, bed)       # the sea surface is the outlet level
    filled = priority_flood(land_bed, ocean | edge)
    surface = np.where(ocean, 0.0, filled)
    lakes = (~ocean) & (surface > bed)
    # two lakes perched above their spill level
    lab = np.zeros((H, W), int)
    n_lab = 0
    for r, c in zip(*np.nonzero(lakes)):
        if lab[r, c]:
            continue
        n_lab += 1
        st = [(r, c)]
        lab[r, c] = n_lab
        while st:
            a, b = st.pop()
            for aa, bb in ((a - 1, b), (a + 1, b), (a, b - 1), (a, b + 1)):
                if 0 <= aa < H and 0 <= bb < W and lakes[aa, bb] and not lab[aa, bb]:
                    lab[aa, bb] = n_lab
                    st.append((aa, bb))
    assert n_lab >= 4, n_lab
    surface = surface + 1.0 * (lab == 1) + 0.5 * (lab == 4)
    mann = np.where(ocean | lakes, 0.025, 0.035)
    mann = np.where((~ocean) & (~lakes) & ((col - 95) ** 2 + (row - 40) ** 2 < 100), 0.10, mann)
    mann = np.where((~ocean) & (~lakes) & ((col - 50) ** 2 + (row - 50) ** 2 < 30), 0.020, mann)

    def layer(a):
        return {"encoding": "b64f32le", "data": base64.b64encode(np.ascontiguousarray(a, "<f4").tobytes()).decode()}

    return {
        "format": "gridlands-terrain", "version": 1,
        "grid": {"width": W, "height": H, "cellMeters": C, "rowOrder": "north-first"},
        "layers": {"bedM": layer(bed), "surfaceM": layer(surface), "manningN": layer(mann)},
        "stats": {"suggestedManningScalar": 0.035, "lakes": int(n_lab)},
        "meta": {"tool": "gridlands (synthetic test export)", "mapId": "synth-7", "preset": "continent", "seed": 7,
                 "spec": "gridlands TERRAIN-FORMAT v1, synthetic export for the import tests"},
    }


if __name__ == "__main__":
    doc = build()
    with open(OUT, "wb") as raw, gzip.GzipFile(fileobj=raw, mode="wb", mtime=0) as f:
        f.write(json.dumps(doc).encode())
    print(OUT, os.path.getsize(OUT), "bytes,", doc["stats"]["lakes"], "lakes")
