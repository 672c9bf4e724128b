// The terrain import (water-simulator_amd/js/gridlands.mjs) driving the drop-in
// solver on the GPU: the acceptance gates of the reference's
// tools/verify-gridlands.mjs, restated on this build's synthetic export
// (tests/golden/gridlands_synth.terrain.json.gz, written by make_gridlands.py;
// two of its lakes deliberately perched above their spill level).
//
//   node tests/js/gridlands_gpu.mjs
//
// PASS/FAIL lines are this side's checks; EXPORT lines are findings about the
// export (the perched lakes), which this synthetic map must produce.
import fs from 'fs';
import zlib from 'zlib';
import path from 'path';
import url from 'url';
import { ShallowWater, reflect } from '../../water-simulator_amd/js/swe.mjs';
import { gridlandsTerrain, verifyInvariants, manningSampler } from '../../water-simulator_amd/js/gridlands.mjs';

const HERE = path.dirname(url.fileURLToPath(import.meta.url));
const FIXTURE = path.join(HERE, '..', 'golden', 'gridlands_synth.terrain.json.gz');
const G = 9.80665, EPS = Number.EPSILON;
let pass = 0, fail = 0, theirs = 0;
function check(name, ok, note = '') {
  if (ok) pass++; else fail++;
  console.log(`  ${ok ? 'PASS' : 'FAIL'}  ${name}${note ? '   ' + note : ''}`);
}
function finding(name, ok, note = '') {
  if (ok) pass++; else theirs++;
  console.log(`  ${ok ? 'PASS' : 'EXPORT'}  ${name}${note ? '   ' + note : ''}`);
}
const near = (a, b, tol) => Math.abs(a - b) <= tol;

const raw = zlib.gunzipSync(fs.readFileSync(FIXTURE)).toString('utf8');
const doc = JSON.parse(raw);
const t = gridlandsTerrain(doc);
console.log(`${t.label}   ${t.provenance.grid}`);

// ---- their invariants, re-derived ------------------------------------------------
const inv = verifyInvariants(t);
check('surfaceM is never below bedM', inv.surfaceBelowBed === 0);
check('every non-ocean waterbody has one surface level', inv.lakesNotFlat === 0, `${inv.lakeBodies} lakes`);
check('the map has water to test', inv.wet > 0.10 * inv.cells, `${inv.wet} of ${inv.cells} wet`);
check('the map has dry land too', inv.wet < 0.95 * inv.cells);
check('the ocean is seated at or below its shoreline', inv.worstOceanPerch <= 1e-6, `${inv.worstOceanPerch}`);
finding('no lake is perched above its spill level', inv.perchedLakes === 0,
  `${inv.perchedLakes} perched, worst ${inv.worstPerch} m`);

// ---- provenance and refusals -----------------------------------------------------
check('label carries preset, seed and mapId', /gridlands .+\/.+ #.+/.test(t.label) && !!t.provenance.mapId, t.label);
check('the spec is carried through', /gridlands/.test(t.provenance.spec || ''));
for (const [what, mutate] of [['format', (d) => Object.assign({}, d, { format: 'other' })],
  ['version', (d) => Object.assign({}, d, { version: 2 })]]) {
  let threw = false;
  try { gridlandsTerrain(mutate(doc)); } catch (e) { threw = true; }
  check(`a wrong ${what} is refused`, threw);
}
{
  let threw = false;
  try {
    const d = JSON.parse(raw);
    d.layers.bedM.data = d.layers.bedM.data.slice(0, 1000);
    gridlandsTerrain(d);
  } catch (e) { threw = true; }
  check('a truncated layer is refused', threw);
}

// ---- the row flip, probed by geometry --------------------------------------------
const C = t.cellMeters, W = t.W, H = t.H;
const yTop = (H - 0.5) * C;
let north = 0, flippedNorth = 0;
const tFlip = gridlandsTerrain(doc, { flipRows: false });
for (let i = 0; i < W; i++) {
  const x = (i + 0.5) * C;
  if (near(t.bed(x, yTop), t.bedM[i], 1e-3)) north++;
  if (near(tFlip.bed(x, yTop), t.bedM[i], 1e-3)) flippedNorth++;
}
check('sim +y is north: the top row samples the export\'s row 0', north === W, `${north}/${W}`);
check('the mutant flipRows:false is caught by that probe', flippedNorth < W, `${flippedNorth}/${W}`);
let differ = false;
for (let i = 0; i < W; i++) if (t.bedM[i] !== t.bedM[(H - 1) * W + i]) differ = true;
check('north and south rows differ (the probe is not vacuous)', differ);
{
  const xs = [-1.5 * C, -0.5 * C, (W - 0.5) * C, (W + 1.5) * C], ys = [-1.5 * C, -0.5 * C, (H - 0.5) * C, (H + 1.5) * C];
  let finite = 0;
  for (const x of xs) for (const y of ys) if (Number.isFinite(t.bed(x, y))) finite++;
  check('bed() is finite on both ghost rings', finite === 16);
  const edge = t.bed(0.5 * C, 0.5 * C);
  check('the clamp extrapolates the edge value', near(t.bed(-1.5 * C, 0.5 * C), edge, 1e-9) && near(t.bed(-0.5 * C, 0.5 * C), edge, 1e-9));
}

// ---- lake at rest, imported ------------------------------------------------------
function restRun(terrain, steps) {
  const sim = new ShallowWater({ nx: terrain.domain.nx, ny: terrain.domain.ny, dx: terrain.domain.dx,
    dy: terrain.domain.dy, bed: terrain.bed, eta0: terrain.eta0, manning: 0, cfl: 0.45 });
  sim.boundaries = { west: reflect, east: reflect, south: reflect, north: reflect };
  const eta0 = new Float64Array(sim.h.length);
  for (let k = 0; k < sim.h.length; k++) eta0[k] = sim.b[k] + sim.h[k];
  let hMax = 0;
  for (let j = 0; j < sim.ny; j++) for (let i = 0; i < sim.nx; i++) hMax = Math.max(hMax, sim.depth(i, j));
  const v0 = sim.volume(), dt = sim.maxDt();
  const t0 = Date.now();
  for (let n = 0; n < steps; n++) sim.step(dt);
  let dEta = 0;
  for (let j = 0; j < sim.ny; j++) for (let i = 0; i < sim.nx; i++) {
    const k = sim.idx(i, j);
    if (sim.h[k] > sim.minDepth) dEta = Math.max(dEta, Math.abs(sim.b[k] + sim.h[k] - eta0[k]));
  }
  return { sim, dt, hMax, dEta, spd: sim.maxSpeed(), drift: Math.abs((sim.volume() - v0) / v0),
    secs: (Date.now() - t0) / 1000 };
}
const STEPS = 1000;
const asImported = restRun(t, STEPS);
const uTol = 4 * EPS * G * asImported.hMax * asImported.dt * STEPS / C;
console.log(`  ${STEPS} steps of dt ${asImported.dt.toFixed(4)} s in ${asImported.secs.toFixed(2)} s; tolerance ${uTol.toExponential(3)} m/s`);
finding('as imported, the still world stays still', asImported.spd < uTol, `max speed ${asImported.spd.toExponential(3)} m/s`);
finding('as imported, the surface does not move', asImported.dEta < 1e-6, `max |eta - eta0| ${asImported.dEta.toExponential(3)} m`);
check('as imported, the state stays finite', asImported.sim.finite());
const tSettled = gridlandsTerrain(doc, { settle: true });
const seated = restRun(tSettled, STEPS);
check('settled, the still world stays still', seated.spd < uTol,
  `max speed ${seated.spd.toExponential(3)} m/s, margin ${(uTol / Math.max(seated.spd, 1e-300)).toFixed(1)}x`);
check('settled, the surface does not move', seated.dEta < 1e-9, `max |eta - eta0| ${seated.dEta.toExponential(3)} m`);
check('settled, the state stays finite', seated.sim.finite());
check('settling changed something, so the pair is a comparison',
  tSettled.settled.bodiesMoved > 0 && asImported.spd > seated.spd * 1e3, `${tSettled.settled.bodiesMoved} bodies moved`);

// ---- volume conservation ---------------------------------------------------------
check('a closed box conserves water (settled)', seated.drift < 1e-13, `drift ${seated.drift.toExponential(3)}`);
check('and the as-imported run conserves too', asImported.drift < 1e-13, `drift ${asImported.drift.toExponential(3)}`);
const bumpAt = (x0, y0, a, r) => (x, y) => a * Math.exp(-(((x - x0) / r) ** 2 + ((y - y0) / r) ** 2));
{
  const bump = bumpAt(900, 2200, 2.0, 500);
  const sim2 = new ShallowWater({ nx: W, ny: H, dx: C, dy: C, bed: tSettled.bed,
    eta0: (x, y) => { const s = tSettled.eta0(x, y), b = tSettled.bed(x, y); return s > b ? s + bump(x, y) : s; },
    manning: 0, cfl: 0.45 });
  sim2.boundaries = { west: reflect, east: reflect, south: reflect, north: reflect };
  const w0 = sim2.volume();
  for (let n = 0; n < 400; n++) sim2.step();
  const drift2 = Math.abs((sim2.volume() - w0) / w0);
  check('a 2 m bump propagating over the terrain conserves', drift2 < 1e-12, `drift ${drift2.toExponential(3)}`);
  check('and it moved', sim2.maxSpeed() > 0.05, `max speed ${sim2.maxSpeed().toFixed(3)} m/s`);
}

// ---- per-cell Manning ------------------------------------------------------------
{
  const nf = t.manningN;
  check('the export carries a manningN layer', !!nf && nf.length === W * H);
  const nAt = manningSampler(t);
  const simM = new ShallowWater({ nx: W, ny: H, dx: C, dy: C, bed: tSettled.bed, eta0: tSettled.eta0,
    manning: t.manning, cfl: 0.45 });
  simM.setManningField(nAt);
  let mism = 0;
  const distinct = new Set();
  for (let j = 0; j < simM.ny; j++) for (let i = 0; i < simM.nx; i++) {
    const kk = ((H - 1) - j) * W + i;
    if (Math.abs(simM.manningField[simM.idx(i, j)] - nf[kk]) > 1e-12) mism++;
    distinct.add(nf[kk]);
  }
  check('the field lands on the cells the bed does', mism === 0, `${mism} mismatches`);
  check('the field varies', distinct.size >= 3, `${distinct.size} distinct values`);
  const flood = (field) => {
    const bump = bumpAt(900, 2200, 3.0, 400);
    const s = new ShallowWater({ nx: W, ny: H, dx: C, dy: C, bed: tSettled.bed,
      eta0: (x, y) => { const su = tSettled.eta0(x, y), b = tSettled.bed(x, y); return su > b ? su + bump(x, y) : su; },
      manning: t.manning, cfl: 0.45 });
    if (field) s.setManningField(nAt);
    s.boundaries = { west: reflect, east: reflect, south: reflect, north: reflect };
    for (let n = 0; n < 300; n++) s.step();
    return s;
  };
  const sS = flood(false), sF = flood(true);
  const vS = sS.maxSpeed(), vF = sF.maxSpeed();
  check('a per-cell field gives a different answer from the scalar', Math.abs(vS - vF) > 1e-9,
    `scalar ${vS.toFixed(4)} m/s, per-cell ${vF.toFixed(4)} m/s`);
  check('both stay finite', sS.finite() && sF.finite());
  let refused = 0;
  for (const bad of [() => -1, () => NaN, () => Infinity]) {
    try { sS.setManningField(bad); } catch (e) { refused++; }
  }
  check('a negative, NaN or infinite n is refused', refused === 3);
  sF.setManningField(null);
  check('setManningField(null) returns to the scalar', sF.manningField === null);
}

console.log(`${fail ? fail + ' FAILURES' : 'OUR SIDE GREEN'} -- ${pass}/${pass + fail + theirs} checks, ${theirs} exporter-side`);
// the synthetic export perches two lakes on purpose: the three exporter-side findings must show
if (theirs !== 3) { console.log(`expected 3 exporter-side findings, saw ${theirs}`); fail++; }
process.exitCode = fail ? 1 : 0;
