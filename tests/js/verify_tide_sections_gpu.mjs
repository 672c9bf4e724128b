// Sections 3 and 3c of the reference's tools/verify-tide.mjs (a long-wave pulse
// leaving through a Flather boundary against a solid-wall control, and the same
// metric at five amplitudes split into its boundary part by a 140 km
// counterfactual domain), restated over the drop-in solver on the GPU, with every
// figure those sections print held to the digits the reference prints.  The
// expected strings are the reference's own output under Node 12.22.9; no
// friction enters, so the figures do not depend on the engine's Math.pow.
//
//   node tests/js/verify_tide_sections_gpu.mjs
import { ShallowWater, reflect, periodic, flather } from '../../water-simulator_amd/js/swe.mjs';

const G = 9.80665;
let fails = 0, checks = 0;
function same(label, got, want) {
  checks++;
  const ok = got === want;
  if (!ok) fails++;
  console.log(`${ok ? 'PASS' : 'FAIL'}  ${label}: ${got}${ok ? '' : `  (the reference prints ${want})`}`);
}

// what the reference prints for these runs
const WANT = {
  wall: '98.2',
  flather: '0.121',
  rows: ['0.0033 0.119 0.039 98.2', '0.0200 0.710 0.234 97.1', '0.0500 1.716 0.547 89.9',
    '0.1000 3.245 0.867 72.6', '0.2000 4.577 1.162 55.1'],
};

// a right-going Gaussian pulse (hu = c0 eta) on a 30 m deep channel, wall to the
// west, `bc` to the east
const h = 30, c0 = Math.sqrt(G * h), L = 60000, dx = 100, x0 = 20000, sig = 2500;
const T = 1.9 * (L - x0) / c0;
function channel(n, amp, bc) {
  const sim = new ShallowWater({ nx: n, ny: 1, dx, bed: () => -h, eta0: 0, manning: 0, cfl: 0.4 });
  for (let i = 0; i < n; i++) {
    const e = amp * Math.exp(-((((i + 0.5) * dx - x0) / sig) ** 2)), k = sim.idx(i, 0);
    sim.h[k] = h + e; sim.hu[k] = c0 * e;
  }
  sim.boundaries = { west: reflect, east: bc, south: periodic, north: periodic };
  return sim;
}
const absorber = () => flather(() => 0, () => 0);
// largest |eta| over the cells the measurement reads (10 .. nx - 41)
const peak = (f) => { let m = 0; for (let i = 10; i < 600 - 40; i++) m = Math.max(m, f(i)); return m; };

// ---- 3. the open boundary absorbs (tools/verify-tide.mjs:780-840) ---------------------------
{
  const returned = (bc) => {
    const sim = channel(600, 0.1, bc);
    while (sim.t < T) sim.step(Math.min(sim.maxDt(), T - sim.t));
    return peak((i) => Math.abs(sim.eta(i, 0))) / 0.1;
  };
  const rF = returned(absorber()), rW = returned(reflect);
  same('solid wall returns, %', (100 * rW).toFixed(1), WANT.wall);
  same('Flather returns, %', (100 * rF).toFixed(3), WANT.flather);
}

// ---- 3c. finite amplitude (tools/verify-tide.mjs:1002-1062) ---------------------------------
{
  const rows = [];
  for (const ratio of [0.0033, 0.02, 0.05, 0.10, 0.20]) {
    const amp = ratio * h;
    const S = channel(600, amp, absorber()), C = channel(1400, amp, absorber()), W = channel(600, amp, reflect);
    const dt = Math.min(S.maxDt(), C.maxDt()) * 0.98, steps = Math.ceil(T / dt);
    for (let n = 0; n < steps; n++) { S.step(dt); C.step(dt); W.step(dt); }
    const row = [ratio.toFixed(4),
      (100 * peak((i) => Math.abs(S.eta(i, 0))) / amp).toFixed(3),
      (100 * peak((i) => Math.abs(S.eta(i, 0) - C.eta(i, 0))) / amp).toFixed(3),
      (100 * peak((i) => Math.abs(W.eta(i, 0) - C.eta(i, 0))) / amp).toFixed(1)].join(' ');
    rows.push(row);
  }
  rows.forEach((r, k) => same(`A/h, suite metric %, boundary part %, wall % (row ${k + 1})`, r, WANT.rows[k]));
}

console.log(fails ? `${fails} FAILED of ${checks}` : `ALL PASS ${checks}/${checks}`);
process.exitCode = fails ? 1 : 0;
