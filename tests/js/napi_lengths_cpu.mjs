// The addon's length checks, run against the stand-in library of
// tests/napi_mock (no GPU): every binding that hands a typed array to the C
// library throws a RangeError for an array of the wrong length, and the
// library is never called for it; arrays of the right length go through.
import { createRequire } from 'module';
import { ShallowWater, Adjoint } from '../../water-simulator_amd/js/swe.mjs';

const require = createRequire(import.meta.url);
const addon = require('../../water-simulator_amd/napi/swe_napi.node');
const calls = () => Number(addon.buildInfo().split(' ')[1]);
let fails = 0, checks = 0;
const check = (ok, what) => { checks++; if (!ok) { fails++; console.log('  FAIL ' + what); } };

if (!addon.buildInfo().startsWith('mock ')) {
  console.log('this check runs against the stand-in library (LD_LIBRARY_PATH=<dir of the mock libswe_amd.so>)');
  process.exit(2);
}

// refused: RangeError, and no library call with the array (a series read first
// asks the library for its sample count, with no array: `queries` such calls);
// accepted: one or more library calls
function refused(what, fn, queries = 0) {
  const c0 = calls();
  let err = null;
  try { fn(); } catch (e) { err = e; }
  check(err instanceof RangeError, `${what}: expected a RangeError, got ${err}`);
  check(calls() === c0 + queries, `${what}: the library was called with the array`);
}
function accepted(what, fn) {
  const c0 = calls();
  try { fn(); } catch (e) { check(false, `${what}: threw ${e}`); return; }
  check(calls() > c0, `${what}: the library was not called`);
}

const nx = 8, ny = 6, W = nx + 4, H = ny + 4, N = W * H;
const sim = new ShallowWater({ nx, ny, dx: 10, bed: () => -1 });
const d = sim._dev;
const f = (n) => new Float64Array(n);

for (const [name, fn] of [['setField', addon.setField], ['getField', addon.getField], ['getGhosts', addon.getGhosts]]) {
  refused(`${name} short`, () => fn(d, 1, f(N - 1)));
  refused(`${name} long`, () => fn(d, 1, f(N + 1)));
  accepted(`${name} W*H`, () => fn(d, 1, f(N)));
}
accepted('setField null (Manning field off)', () => addon.setField(d, 4, null));
refused('setRow short', () => addon.setRow(d, 0, f(H - 1)));
accepted('setRow H', () => addon.setRow(d, 0, f(H)));
refused('diagnostics out[4]', () => addon.diagnostics(d, 0, f(4)));
accepted('diagnostics out[5]', () => addon.diagnostics(d, 0, f(5)));
refused('getExt[7]', () => addon.getExt(d, f(7)));
accepted('getExt[8]', () => addon.getExt(d, f(8)));

// setTide: amp/omega/phase one per constituent, lat H, lon W
const m2 = Int32Array.from([2, 1]);
refused('setTide amp short', () => addon.setTide(d, m2, f(1), f(2), f(2), f(H), f(W)));
refused('setTide lat short', () => addon.setTide(d, m2, f(2), f(2), f(2), f(H - 1), f(W)));
refused('setTide lon short', () => addon.setTide(d, m2, f(2), f(2), f(2), f(H), f(W - 2)));
accepted('setTide', () => addon.setTide(d, m2, f(2), f(2), f(2), f(H), f(W)));

// harmonics: equal-length term arrays of at most 4
refused('setHarmonic mismatched terms', () => addon.setHarmonic(d, 0, 0, f(2), f(1), f(2), 0, null, null, null));
refused('setHarmonic 5 terms', () => addon.setHarmonic(d, 0, 0, f(5), f(5), f(5), 0, null, null, null));
accepted('setHarmonic', () => addon.setHarmonic(d, 0, 0, f(2), f(2), f(2), 0, f(1), f(1), f(1)));

// recorder: 3 gauges, capacity 10; the stand-in reports 5 samples -> 15 values
addon.setRecorder(d, Int32Array.from([30, 31, 32]), 1, 0, 0, 10);
refused('getRecord eta short', () => addon.getRecord(d, f(14), null), 1);
refused('getRecord times short', () => addon.getRecord(d, null, f(4)), 1);
accepted('getRecord', () => addon.getRecord(d, f(15), f(5)));
refused('setProjection coef short', () => addon.setProjection(d, f(ny - 1), 10));
addon.setProjection(d, f(ny), 10);
refused('getProjection vals short', () => addon.getProjection(d, f(4), null), 1);
accepted('getProjection', () => addon.getProjection(d, f(5), f(5)));

// the adjoint: 100 steps, gauges every 10 -> 10 samples per gauge
const adj = new Adjoint(sim, 100);
const a = adj._adj;
refused('adjSetGauges obs short', () => addon.adjSetGauges(a, Int32Array.from([30, 31]), 10, f(19)));
refused('adjSetGauges every 0', () => addon.adjSetGauges(a, Int32Array.from([30, 31]), 0, f(20)));
accepted('adjSetGauges', () => addon.adjSetGauges(a, Int32Array.from([30, 31]), 10, f(20)));
refused('adjForward h0 short', () => addon.adjForward(a, f(N - 1), null));
refused('adjForward n short', () => addon.adjForward(a, f(N), f(N - 1)));
refused('adjLoss h0 long', () => addon.adjLoss(a, f(N + 1), null));
accepted('adjForward', () => addon.adjForward(a, f(N), f(N)));
refused('adjBackward short', () => addon.adjBackward(a, f(N - 1), null));
refused('adjBackward n short', () => addon.adjBackward(a, null, f(1)));
accepted('adjBackward', () => addon.adjBackward(a, f(N), f(N)));
refused('adjSeries short', () => addon.adjSeries(a, f(19)));
accepted('adjSeries', () => addon.adjSeries(a, f(20)));

// create: every geometry row ny + 4 long
const rows = () => Array.from({ length: 9 }, () => f(H));
const desc = { nx, ny, kind: 0, order: 2, limiter: 1, g: 9.80665, minDepth: 1e-3, eps4: 1e-12, cfl: 0.45,
  coriolis: 0, manning: 0, R: Infinity };
refused('create row short', () => { const r = rows(); r[3] = f(H - 1); addon.create({ ...desc, rows: r }); });
refused('create nx 0', () => addon.create({ ...desc, nx: 0, rows: rows() }));
accepted('create', () => addon.create({ ...desc, rows: rows() }));

// the drop-in's own paths reach the library with full arrays
accepted('sim.step', () => sim.step(0.1));
console.log(`${fails ? fails + ' FAILURES' : 'ALL PASS'} -- ${checks - fails}/${checks} checks`);
process.exit(fails ? 1 : 0);
