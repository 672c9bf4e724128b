// Terrain import from gridlands' `gridlands-terrain` v1 export for the Node
// drop-in: the data format on the input side of the solver, with the exports of
// the reference's src/gridlands.mjs (gridlandsTerrain, loadGridlandsTerrain,
// verifyInvariants, waterbodies, settleLakes) and manningSampler.
//
// A document is a W x H grid at `cellMeters`, row 0 = north, whose layers
// `bedM`, `surfaceM` (standing water) and optionally `manningN` are float32
// little-endian arrays in base64.  The module works in three passes over the
// decoded grid, each a plain loop over cells in raster order:
//   1. labelWater: the wet cells (surface above bed) are split into 4-connected
//      waterbodies by two-pass connected-component labelling (Rosenfeld and
//      Pfaltz): a raster scan that unions each wet cell with its wet west and
//      north neighbours in a disjoint-set forest, then a pass that numbers the
//      roots in order of their first cell;
//   2. bodyStats: per body, in one pass over the labels, its first cell and
//      level (that cell's surface), size, the largest departure of a surface
//      from the level, and the lowest bed among the dry cells bordering it (the
//      spill level);
//   3. the invariant report and the settle step read those per-body arrays.
// water-simulator_amd/gridlands.py is the same design for the Python host.
// Pure JS: it runs on Node 12, where the drop-in runs.

const LITTLE_ENDIAN = new Uint8Array(new Uint32Array([1]).buffer)[0] === 1;
// 40 cells per wavelength: the shortest wave the grid carries without visible
// numerical dispersion, as the reference states it for its maps
const CELLS_PER_WAVELENGTH = 40;
// a body whose level stands more than this above its spill level is perched
const PERCH_TOLERANCE = 1e-6;

function fail(msg) { throw new Error(`gridlands: ${msg}`); }

function base64Bytes(text) {
  if (typeof Buffer !== 'undefined' && typeof Buffer.from === 'function') {
    const buf = Buffer.from(text, 'base64');
    return new Uint8Array(buf.buffer, buf.byteOffset, buf.byteLength);
  }
  const bin = atob(text);
  const out = new Uint8Array(bin.length);
  for (let i = 0; i < bin.length; i++) out[i] = bin.charCodeAt(i);
  return out;
}

// The float32 values of one layer: `b64f32le`, exactly `cells` of them.
function decodeLayer(layers, name, cells, required) {
  const layer = layers[name];
  if (!layer) {
    if (required) fail(`layer ${name} is missing`);
    return null;
  }
  if (layer.encoding !== 'b64f32le') fail(`layer ${name} is encoded as ${layer.encoding}; only b64f32le is read`);
  if (!LITTLE_ENDIAN) fail('b64f32le layers are read on little-endian hosts only');
  const bytes = base64Bytes(layer.data);
  if (bytes.byteLength !== 4 * cells) {
    fail(`layer ${name} holds ${bytes.byteLength} bytes where ${cells} cells need ${4 * cells}`);
  }
  // copy into an aligned buffer of its own (a decoded Buffer may share a pool)
  const out = new Float32Array(cells);
  new Uint8Array(out.buffer).set(bytes);
  return out;
}

// ---- pass 1: waterbody labels ----------------------------------------------------

function isWet(bedM, surfaceM, k) { return surfaceM[k] - bedM[k] > 0; }

/**
 * Labels of the 4-connected waterbodies: an Int32Array over the W*H cells
 * holding the body number (0, 1, ... in order of each body's first cell in
 * raster order) of a wet cell and -1 of a dry one, and the number of bodies.
 */
export function labelWater(bedM, surfaceM, W, H) {
  const n = W * H;
  const parent = new Int32Array(n).fill(-1);
  const find = (k) => {
    while (parent[k] !== k) { parent[k] = parent[parent[k]]; k = parent[k]; }
    return k;
  };
  const union = (a, b) => {
    const ra = find(a), rb = find(b);
    if (ra === rb) return;
    // the root with the smaller index survives, so a root is its body's first cell
    if (ra < rb) parent[rb] = ra; else parent[ra] = rb;
  };
  for (let k = 0; k < n; k++) {
    if (!isWet(bedM, surfaceM, k)) continue;
    parent[k] = k;
    if (k % W > 0 && parent[k - 1] >= 0) union(k, k - 1);
    if (k >= W && parent[k - W] >= 0) union(k, k - W);
  }
  const label = new Int32Array(n).fill(-1);
  let bodies = 0;
  for (let k = 0; k < n; k++) {
    if (parent[k] < 0) continue;
    const r = find(k);
    label[k] = r === k ? bodies++ : label[r];
  }
  return { label, count: bodies };
}

// ---- pass 2: per-body statistics -------------------------------------------------

function bodyStats(bedM, surfaceM, W, H) {
  const { label, count } = labelWater(bedM, surfaceM, W, H);
  const first = new Int32Array(count).fill(-1);
  const size = new Int32Array(count);
  const level = new Float64Array(count);
  const spread = new Float64Array(count);
  const ringMin = new Float64Array(count).fill(Infinity);
  const n = W * H;
  for (let k = 0; k < n; k++) {
    const b = label[k];
    if (b < 0) continue;
    if (first[b] < 0) { first[b] = k; level[b] = surfaceM[k]; }
    size[b]++;
    const dev = Math.abs(surfaceM[k] - level[b]);
    if (dev > spread[b]) spread[b] = dev;
    // the dry cells bordering this one hold the body in
    const col = k % W;
    if (col > 0 && label[k - 1] < 0) ringMin[b] = Math.min(ringMin[b], bedM[k - 1]);
    if (col < W - 1 && label[k + 1] < 0) ringMin[b] = Math.min(ringMin[b], bedM[k + 1]);
    if (k >= W && label[k - W] < 0) ringMin[b] = Math.min(ringMin[b], bedM[k - W]);
    if (k + W < n && label[k + W] < 0) ringMin[b] = Math.min(ringMin[b], bedM[k + W]);
  }
  return { label, count, first, size, level, spread, ringMin };
}

/**
 * The connected waterbodies in order of their first cell: each with its level
 * (the surface of that first cell), its cells (raster order), the largest
 * departure of a cell's surface from the level, the lowest bed on its ring of
 * dry neighbours (the spill level, Infinity when it has none), how far the
 * level stands above it, and whether it is the ocean (level exactly 0).
 */
export function waterbodies(bedM, surfaceM, W, H) {
  const s = bodyStats(bedM, surfaceM, W, H);
  // bucket the cells by body, keeping raster order within a body
  const start = new Int32Array(s.count + 1);
  for (let b = 0; b < s.count; b++) start[b + 1] = start[b] + s.size[b];
  const fill = start.slice(0, s.count);
  const order = new Int32Array(start[s.count]);
  for (let k = 0; k < s.label.length; k++) if (s.label[k] >= 0) order[fill[s.label[k]]++] = k;
  const out = [];
  for (let b = 0; b < s.count; b++) {
    out.push({
      level: s.level[b], cells: order.subarray(start[b], start[b + 1]), spread: s.spread[b],
      ringMin: s.ringMin[b], perch: s.level[b] - s.ringMin[b], isOcean: s.level[b] === 0,
    });
  }
  return out;
}

// ---- pass 3: the invariant report and the settle step ------------------------------

/**
 * The exporter's guarantees re-derived from the decoded layers, and the one the
 * solver needs that the format does not promise (no lake standing above the
 * lowest bed of its own shoreline).  A report; it throws nothing.
 */
export function verifyInvariants(t) {
  const { bedM, surfaceM, W, H } = t;
  const n = W * H;
  const r = {
    cells: n, wet: 0, oceanCells: 0, surfaceBelowBed: 0, worstBelowBed: 0,
    bodies: 0, lakeBodies: 0, lakesNotFlat: 0, worstLakeSpread: 0, largestBody: 0,
    perchedLakes: 0, worstPerch: 0, perchedCells: 0, perchedVolumeCells: 0, worstOceanPerch: -Infinity,
  };
  for (let k = 0; k < n; k++) {
    const depth = surfaceM[k] - bedM[k];
    if (depth > 0) r.wet++;
    else if (depth < 0) { r.surfaceBelowBed++; if (depth < r.worstBelowBed) r.worstBelowBed = depth; }
    if (surfaceM[k] === 0 && bedM[k] < 0) r.oceanCells++;
  }
  const s = bodyStats(bedM, surfaceM, W, H);
  r.bodies = s.count;
  for (let b = 0; b < s.count; b++) {
    const perch = s.level[b] - s.ringMin[b];
    if (s.size[b] > r.largestBody) r.largestBody = s.size[b];
    if (s.spread[b] > 0) {
      r.lakesNotFlat++;
      if (s.spread[b] > r.worstLakeSpread) r.worstLakeSpread = s.spread[b];
    }
    if (s.level[b] === 0) {
      if (perch > r.worstOceanPerch) r.worstOceanPerch = perch;
      continue;
    }
    r.lakeBodies++;
    if (perch > PERCH_TOLERANCE) {
      r.perchedLakes++;
      r.perchedCells += s.size[b];
      r.perchedVolumeCells += s.size[b] * perch;
      if (perch > r.worstPerch) r.worstPerch = perch;
    }
  }
  return r;
}

/**
 * Seat every body standing more than 1e-6 m above its spill level at that
 * level, in place: each of its cells gets surface max(bed, spill), so cells
 * whose bed is above the spill level fall dry.  Returns what moved.  An opt-in
 * repair on the importing side.
 */
export function settleLakes(bedM, surfaceM, W, H) {
  const s = bodyStats(bedM, surfaceM, W, H);
  const perched = new Uint8Array(s.count);
  const moved = [];
  for (let b = 0; b < s.count; b++) {
    const drop = s.level[b] - s.ringMin[b];
    if (!(drop > PERCH_TOLERANCE)) continue;
    perched[b] = 1;
    moved.push({ from: s.level[b], to: s.ringMin[b], drop, cells: s.size[b] });
  }
  for (let k = 0; k < s.label.length; k++) {
    const b = s.label[k];
    if (b >= 0 && perched[b]) surfaceM[k] = Math.max(bedM[k], s.ringMin[b]);
  }
  let worstDrop = 0;
  for (const m of moved) if (m.drop > worstDrop) worstDrop = m.drop;
  return { bodiesMoved: moved.length, worstDrop, moved };
}

// ---- sampling at the solver's cell centres ------------------------------------------

// Export grid coordinates of a simulation point: fractional column and row,
// clamped to the grid (the constructor also samples two ghost rings beyond
// it); with flipRows the export's row 0 (north) is the top of the domain.
function gridCoords(W, H, C, flipRows) {
  const col = (x) => Math.min(W - 1, Math.max(0, x / C - 0.5));
  const row = flipRows
    ? (y) => (H - 1) - Math.min(H - 1, Math.max(0, y / C - 0.5))
    : (y) => Math.min(H - 1, Math.max(0, y / C - 0.5));
  return { col, row };
}

// Bilinear interpolation of a layer between the four surrounding export cells.
function bilinear(layer, W, H, coords) {
  return (x, y) => {
    const fc = coords.col(x), fr = coords.row(y);
    const c0 = Math.floor(fc), r0 = Math.floor(fr);
    const c1 = Math.min(W - 1, c0 + 1), r1 = Math.min(H - 1, r0 + 1);
    const tc = fc - c0, tr = fr - r0;
    const north = layer[r0 * W + c0] * (1 - tc) + layer[r0 * W + c1] * tc;
    const south = layer[r1 * W + c0] * (1 - tc) + layer[r1 * W + c1] * tc;
    return north * (1 - tr) + south * tr;
  };
}

// The nearest export cell's value (Math.round on both coordinates), so that a
// flat waterbody samples to one flat level.
function nearest(layer, W, coords) {
  return (x, y) => layer[Math.round(coords.row(y)) * W + Math.round(coords.col(x))];
}

// The document checks, in the order they are made; each returns an error
// message or nothing.
const DOCUMENT_CHECKS = [
  (d) => (d && typeof d === 'object' ? null : 'not a document'),
  (d) => (d.format === 'gridlands-terrain' && d.version === 1 ? null
    : `${d.format} v${d.version} is not read here; this adapter reads gridlands-terrain v1 only`),
  (d) => {
    const g = d.grid || {};
    return g.width > 0 && g.height > 0 && g.cellMeters > 0 ? null : `bad grid ${JSON.stringify(g)}`;
  },
  (d) => (!d.grid.rowOrder || d.grid.rowOrder === 'north-first' ? null
    : `rows are ${d.grid.rowOrder}; only north-first is read`),
];

/**
 * A parsed gridlands-terrain v1 document as ShallowWater arguments: bed(x, y)
 * and eta0(x, y) samplers, the domain, the Manning scalar the export suggests
 * (0.03 when none) and its per-cell layer, the decoded layers, and the
 * provenance the map carries.  flipRows (default true) puts the export's
 * north row at the top of the simulation (+y up); settle seats perched lakes.
 */
export function gridlandsTerrain(doc, { flipRows = true, settle = false } = {}) {
  for (const check of DOCUMENT_CHECKS) {
    const msg = check(doc);
    if (msg) fail(msg);
  }
  const { width: W, height: H, cellMeters: C } = doc.grid;
  const layers = doc.layers || {};
  const bedM = decodeLayer(layers, 'bedM', W * H, true);
  const surfaceM = decodeLayer(layers, 'surfaceM', W * H, true);
  const manningN = decodeLayer(layers, 'manningN', W * H, false);
  const settled = settle ? settleLakes(bedM, surfaceM, W, H) : null;
  const coords = gridCoords(W, H, C, flipRows);
  const meta = doc.meta || {};
  const stats = doc.stats || {};
  const km = (m) => (m / 1000).toFixed(1);
  const suggested = stats.suggestedManningScalar;
  return {
    W, H, cellMeters: C, bedM, surfaceM, manningN, settled, stats, meta,
    domain: { nx: W, ny: H, dx: C, dy: C },
    bed: bilinear(bedM, W, H, coords),
    eta0: nearest(surfaceM, W, coords),
    manning: suggested == null ? 0.03 : suggested,
    label: `gridlands ${meta.preset}/${meta.seed} #${meta.mapId}`,
    provenance: {
      tool: meta.tool, mapId: meta.mapId, preset: meta.preset, seed: meta.seed, spec: meta.spec,
      format: doc.format, version: doc.version,
      grid: `${W} x ${H} cells of ${C} m (${km(W * C)} km by ${km(H * C)} km)`,
    },
    shortestHonestWavelength: CELLS_PER_WAVELENGTH * C,
    flipRows,
  };
}

/** Parse (when given text) and adapt in one call. */
export function loadGridlandsTerrain(text, opts) {
  return gridlandsTerrain(typeof text === 'string' ? JSON.parse(text) : text, opts);
}

/**
 * The per-cell Manning layer as a setManningField sampler: the nearest export
 * cell under the terrain's row flip and clamp, the mapping
 * tools/verify-gridlands.mjs:332-333 builds by hand.  Throws when the export
 * has no manningN layer.
 */
export function manningSampler(t) {
  if (!t.manningN) fail('the export carries no manningN layer');
  return nearest(t.manningN, t.W, gridCoords(t.W, t.H, t.cellMeters, t.flipRows !== false));
}
