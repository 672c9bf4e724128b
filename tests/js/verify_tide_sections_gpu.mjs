// Sections 3, 3b and 3c of the reference's tools/verify-tide.mjs (a long-wave
// pulse leaving through a Flather boundary against a solid-wall control, the
// same boundary met at 0-60 degrees of incidence in 2-D, and the metric at five
// amplitudes split into its boundary part by a 140 km counterfactual domain), restated over the drop-in solver on the GPU, with every
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

// ---- 3b. oblique incidence (tools/verify-tide.mjs:844-1000) ---------------------------------
{
  // a plane pulse at angle th to the east boundary in a channel periodic in y;
  // what the short domain's boundary returns is read against a long domain
  // whose boundary the wave has not reached, in the window 6-19 km
  const dx = 200, Ly = 12000, ny = Ly / dx, X0 = 4000, SIG = 2500, A = 0.1, iW0 = 30, iW1 = 95;
  const wave = (nx, th, bc) => {
    const ky = th === 0 ? 0 : 2 * Math.PI / Ly;
    const kx = th === 0 ? 2 * Math.PI / 6000 : ky / Math.tan(th);
    const k = Math.hypot(kx, ky);
    const sim = new ShallowWater({ nx, ny, dx, dy: dx, bed: () => -h, eta0: 0, manning: 0, cfl: 0.4 });
    for (let j = 0; j < ny; j++) for (let i = 0; i < nx; i++) {
      const x = (i + 0.5) * dx, y = (j + 0.5) * dx;
      const E = Math.exp(-(((x - X0) / SIG) ** 2)), e = A * E * Math.cos(kx * x) * Math.cos(ky * y), g = sim.idx(i, j);
      sim.h[g] = h + e;
      sim.hu[g] = c0 * (kx / k) * e;
      sim.hv[g] = -c0 * (ky / k) * A * E * Math.sin(kx * x) * Math.sin(ky * y);
    }
    sim.boundaries = { west: reflect, east: bc, south: periodic, north: periodic };
    return { sim, lam: 2 * Math.PI / k };
  };
  const inWindow = (f) => { let m = 0; for (let j = 0; j < ny; j++) for (let i = iW0; i < iW1; i++) m = Math.max(m, f(i, j)); return m; };
  const apart = (a, b) => inWindow((i, j) => Math.abs(a.eta(i, j) - b.eta(i, j)));
  const want = {
    0: '0 6000 30 0.022 0.076 79.2 84.8 0.089',
    30: '30 6000 30 12.978 5.655 65.3 89.9 6.288',
    45: '45 8485 42 8.961 7.317 50.7 75.3 9.721',
    60: '60 10392 52 24.529 9.262 43.1 70.3 13.168',
  };
  for (const deg of [0, 30, 45, 60]) {
    const th = deg * Math.PI / 180;
    const S = wave(100, th, absorber()), C = wave(240, th, absorber()), W = wave(100, th, reflect);
    const C2 = deg === 60 ? wave(320, th, absorber()) : null;
    const dt = Math.min(S.sim.maxDt(), C.sim.maxDt()) * 0.98;
    const steps = Math.ceil(((100 * dx - X0) + (100 - iW0) * dx) / (c0 * Math.cos(th)) / dt);
    let mB = 0, mW = 0, mB2 = 0, mInc = 0;
    for (let n = 0; n < steps; n++) {
      S.sim.step(dt); C.sim.step(dt); W.sim.step(dt);
      if (C2) C2.sim.step(dt);
      mB = Math.max(mB, apart(S.sim, C.sim));
      mW = Math.max(mW, apart(W.sim, C.sim));
      mInc = Math.max(mInc, inWindow((i, j) => Math.abs(C.sim.eta(i, j))));
      if (C2) mB2 = Math.max(mB2, apart(S.sim, C2.sim));
    }
    const naive = inWindow((i, j) => Math.abs(S.sim.eta(i, j))) / A;
    same(`angle, wavelength m, cells/L, naive %, boundary %, wall %, arrived %, returned % (${deg} deg)`,
      [deg, S.lam.toFixed(0), (S.lam / dx).toFixed(0), (100 * naive).toFixed(3), (100 * mB / A).toFixed(3),
        (100 * mW / A).toFixed(1), (100 * mInc / A).toFixed(1), (100 * mB / mInc).toFixed(3)].join(' '), want[deg]);
    if (C2) same('at 60 deg, the 48 km and 64 km controls, %', `${(100 * mB / A).toFixed(3)} ${(100 * mB2 / A).toFixed(3)}`, '9.262 9.262');
  }
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
