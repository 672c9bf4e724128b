// Terrain import from gridlands' `gridlands-terrain` v1 export: the data format
// on the input side of the solver (src/gridlands.mjs of the reference, same
// exports and semantics).  A document carries a W x H grid at `cellMeters`
// spacing, row 0 = north, with float32 little-endian base64 layers `bedM`,
// `surfaceM` (standing water) and optionally `manningN`.  The adapter hands the
// solver's constructor `bed(x, y)` / `eta0(x, y)` samplers (row flip, edge
// clamp, bilinear bed, nearest surface), checks the exporter's invariants, and
// can seat perched lakes at their spill level (opt-in, reported).
//
// Pure JS: it runs wherever the drop-in does, on Node 12 as well.

const LITTLE_ENDIAN = new Uint8Array(new Uint32Array([1]).buffer)[0] === 1;

function bytesOfBase64(text) {
  if (typeof Buffer !== 'undefined' && typeof Buffer.from === 'function') {
    const b = Buffer.from(text, 'base64');
    // a small Buffer can be a view into a shared pool: keep its own window only
    return new Uint8Array(b.buffer, b.byteOffset, b.byteLength);
  }
  const s = atob(text);
  const out = new Uint8Array(s.length);
  for (let i = 0; i < s.length; i++) out[i] = s.charCodeAt(i);
  return out;
}

// One `b64f32le` layer as exactly n float32 values (gridlands.mjs:86-110): a
// missing layer, another encoding, a big-endian host or a byte count other than
// 4n is an error, never a strange-looking map.
function layerF32(layer, n, name) {
  if (!layer) throw new Error(`gridlands: layer ${name} is missing`);
  if (layer.encoding !== 'b64f32le') {
    throw new Error(`gridlands: layer ${name} is encoded as ${layer.encoding}; only b64f32le is read`);
  }
  if (!LITTLE_ENDIAN) throw new Error('gridlands: b64f32le layers are read on little-endian hosts only');
  const bytes = bytesOfBase64(layer.data);
  if (bytes.byteLength !== 4 * n) {
    throw new Error(`gridlands: layer ${name} holds ${bytes.byteLength} bytes; the grid needs ${4 * n} `
      + `(${n} cells x 4)`);
  }
  return bytes.byteOffset % 4 === 0 ? new Float32Array(bytes.buffer, bytes.byteOffset, n)
    : new Float32Array(bytes.slice().buffer);
}

/**
 * The connected waterbodies (cells with surface above bed, 4-neighbour
 * connected), in order of their first cell, each with its level (the surface
 * of that first cell), its cells, the largest departure of a cell's surface
 * from the level, and the lowest bed on its ring of dry neighbours -- the
 * spill level (gridlands.mjs:119-155).
 */
export function waterbodies(bedM, surfaceM, W, H) {
  const n = W * H;
  const isWet = (k) => surfaceM[k] - bedM[k] > 0;
  const visited = new Uint8Array(n);
  const bodies = [];
  const todo = [];
  for (let start = 0; start < n; start++) {
    if (visited[start] || !isWet(start)) continue;
    const level = surfaceM[start];
    const cells = [];
    const ring = new Set();
    let spread = 0;
    visited[start] = 1;
    todo.length = 0;
    todo.push(start);
    while (todo.length) {
      const k = todo.pop();
      cells.push(k);
      if (surfaceM[k] !== level) spread = Math.max(spread, Math.abs(surfaceM[k] - level));
      const col = k % W, row = (k - col) / W;
      const around = [];
      if (col > 0) around.push(k - 1);
      if (col < W - 1) around.push(k + 1);
      if (row > 0) around.push(k - W);
      if (row < H - 1) around.push(k + W);
      for (const m of around) {
        if (!isWet(m)) ring.add(m);
        else if (!visited[m]) { visited[m] = 1; todo.push(m); }
      }
    }
    let ringMin = Infinity;
    for (const m of ring) ringMin = Math.min(ringMin, bedM[m]);
    bodies.push({ level, cells, spread, ringMin, perch: level - ringMin, isOcean: level === 0 });
  }
  return bodies;
}

/**
 * The exporter's guarantees re-derived from the decoded layers
 * (gridlands.mjs:157-206), plus the one the solver needs and the format does
 * not promise: no lake above the lowest bed on its own shoreline (perched).
 * Reports, throws nothing.
 */
export function verifyInvariants(t) {
  const { bedM, surfaceM, W, H } = t;
  const n = W * H;
  let surfaceBelowBed = 0, worstBelowBed = 0, wet = 0, oceanCells = 0;
  for (let k = 0; k < n; k++) {
    const d = surfaceM[k] - bedM[k];
    if (d < 0) { surfaceBelowBed++; worstBelowBed = Math.min(worstBelowBed, d); }
    if (d > 0) wet++;
    if (surfaceM[k] === 0 && bedM[k] < 0) oceanCells++;
  }
  const bodies = waterbodies(bedM, surfaceM, W, H);
  const lakes = bodies.filter((b) => !b.isOcean);
  const uneven = bodies.filter((b) => b.spread > 0);
  const perched = lakes.filter((b) => b.perch > 1e-6).sort((a, b) => b.perch - a.perch);
  let perchedVolumeCells = 0;
  for (const b of perched) perchedVolumeCells += b.cells.length * b.perch;
  return {
    cells: n, wet, oceanCells, surfaceBelowBed, worstBelowBed,
    dryMismatch: 0,
    bodies: bodies.length,
    lakeBodies: lakes.length,
    lakesNotFlat: uneven.length,
    worstLakeSpread: uneven.reduce((m, b) => Math.max(m, b.spread), 0),
    largestBody: bodies.reduce((m, b) => Math.max(m, b.cells.length), 0),
    perchedLakes: perched.length,
    worstPerch: perched.length ? perched[0].perch : 0,
    perchedCells: perched.reduce((m, b) => m + b.cells.length, 0),
    perchedVolumeCells,
    worstOceanPerch: bodies.filter((b) => b.isOcean).reduce((m, b) => Math.max(m, b.perch), -Infinity),
  };
}

/**
 * Lower every body standing more than 1e-6 m above its spill level to that
 * level, in place (cells left above it become dry), and report what moved
 * (gridlands.mjs:208-219).  Opt-in: a repair on the importing side.
 */
export function settleLakes(bedM, surfaceM, W, H) {
  const moved = [];
  for (const b of waterbodies(bedM, surfaceM, W, H)) {
    if (!(b.perch > 1e-6)) continue;
    for (const k of b.cells) surfaceM[k] = Math.max(bedM[k], b.ringMin);
    moved.push({ from: b.level, to: b.ringMin, drop: b.perch, cells: b.cells.length });
  }
  return { bodiesMoved: moved.length, worstDrop: moved.reduce((m, x) => Math.max(m, x.drop), 0), moved };
}

/**
 * A parsed gridlands-terrain v1 document as ShallowWater arguments
 * (gridlands.mjs:221-302).  flipRows (default true) maps the export's row 0
 * (north) to the top of the simulation (+y up); settle seats perched lakes.
 */
export function gridlandsTerrain(doc, { flipRows = true, settle = false } = {}) {
  if (!doc || typeof doc !== 'object') throw new Error('gridlands: not a document');
  if (doc.format !== 'gridlands-terrain' || doc.version !== 1) {
    throw new Error(`gridlands: ${doc.format} v${doc.version} is not read here; `
      + 'this adapter reads gridlands-terrain v1 only');
  }
  const grid = doc.grid || {};
  const W = grid.width, H = grid.height, C = grid.cellMeters;
  if (!(W > 0 && H > 0 && C > 0)) throw new Error(`gridlands: bad grid ${JSON.stringify(grid)}`);
  if (grid.rowOrder && grid.rowOrder !== 'north-first') {
    throw new Error(`gridlands: rows are ${grid.rowOrder}; only north-first is read`);
  }
  const n = W * H;
  const layers = doc.layers || {};
  const bedM = layerF32(layers.bedM, n, 'bedM');
  const surfaceM = layerF32(layers.surfaceM, n, 'surfaceM');
  const manningN = layers.manningN ? layerF32(layers.manningN, n, 'manningN') : null;
  const settled = settle ? settleLakes(bedM, surfaceM, W, H) : null;

  // fractional export row and column of a simulation point, clamped to the
  // grid (the constructor samples two ghost rings beyond it as well)
  const colOf = (x) => Math.min(W - 1, Math.max(0, x / C - 0.5));
  const rowOf = flipRows
    ? (y) => (H - 1) - Math.min(H - 1, Math.max(0, y / C - 0.5))
    : (y) => Math.min(H - 1, Math.max(0, y / C - 0.5));
  // bed: bilinear between the four surrounding export cells
  const bed = (x, y) => {
    const fc = colOf(x), fr = rowOf(y);
    const c0 = Math.floor(fc), r0 = Math.floor(fr);
    const c1 = Math.min(W - 1, c0 + 1), r1 = Math.min(H - 1, r0 + 1);
    const tc = fc - c0, tr = fr - r0;
    const v00 = bedM[r0 * W + c0], v10 = bedM[r0 * W + c1];
    const v01 = bedM[r1 * W + c0], v11 = bedM[r1 * W + c1];
    return (v00 * (1 - tc) + v10 * tc) * (1 - tr) + (v01 * (1 - tc) + v11 * tc) * tr;
  };
  // surface: the nearest export cell, so each waterbody stays one flat level
  const eta0 = (x, y) => surfaceM[Math.round(rowOf(y)) * W + Math.round(colOf(x))];

  const stats = doc.stats || {};
  const meta = doc.meta || {};
  const suggested = stats.suggestedManningScalar;
  return {
    bed, eta0,
    domain: { nx: W, ny: H, dx: C, dy: C },
    manning: suggested === undefined || suggested === null ? 0.03 : suggested,
    manningN,
    bedM, surfaceM, W, H, cellMeters: C,
    settled,
    stats,
    meta,
    label: `gridlands ${meta.preset}/${meta.seed} #${meta.mapId}`,
    provenance: {
      tool: meta.tool, mapId: meta.mapId, preset: meta.preset, seed: meta.seed, spec: meta.spec,
      format: doc.format, version: doc.version,
      grid: `${W}x${H} @ ${C} m = ${(W * C / 1000).toFixed(1)} x ${(H * C / 1000).toFixed(1)} km`,
    },
    // 40 cells per wavelength: the shortest wave the grid resolves honestly
    shortestHonestWavelength: 40 * C,
  };
}

/** Parse (if given text) and adapt in one call. */
export function loadGridlandsTerrain(text, opts) {
  return gridlandsTerrain(typeof text === 'string' ? JSON.parse(text) : text, opts);
}

/**
 * The per-cell Manning layer as a setManningField sampler: the nearest export
 * cell under the same row flip and clamp as the surface (the mapping
 * tools/verify-gridlands.mjs:332-333 uses).  Throws when the export has none.
 */
export function manningSampler(t, { flipRows = true } = {}) {
  if (!t.manningN) throw new Error('gridlands: the export carries no manningN layer');
  const { W, H, cellMeters: C, manningN } = t;
  const colOf = (x) => Math.min(W - 1, Math.max(0, x / C - 0.5));
  const rowOf = flipRows
    ? (y) => (H - 1) - Math.min(H - 1, Math.max(0, y / C - 0.5))
    : (y) => Math.min(H - 1, Math.max(0, y / C - 0.5));
  return (x, y) => manningN[Math.round(rowOf(y)) * W + Math.round(colOf(x))];
}
