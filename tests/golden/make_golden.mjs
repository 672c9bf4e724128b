// Golden-vector writer: runs the REFERENCE solver (src/swe.mjs of
// Syntaxswine/water-simulator, read from REF_DIR, default /root/reference) on small
// seeded scenarios and writes inputs + outputs as JSON fixtures beside this file.
//
//   node tests/golden/make_golden.mjs [scenario ...]
//
// The fixtures are data only: initial arrays, geometry rows, per-step dt, boundary
// forcing values, massFloored, an FNV-1a-64 fingerprint of the raw bytes of h,hu,hv
// after every step, and full arrays at a few checkpoints.  Doubles are stored as
// 16-hex-digit IEEE-754 bit patterns and arrays as base64 of little-endian
// Float64Array bytes, so nothing is lost to decimal printing.
//
// This script and oracle/time_reference.mjs (the CPU timing of the reference) are
// the only places that execute the reference, and only in the build container;
// nothing that runs on a GPU machine reads /root/reference.
import fs from 'fs';
import zlib from 'zlib';
import path from 'path';
import url from 'url';

const REF_DIR = process.env.REF_DIR || '/root/reference';
const here = path.dirname(url.fileURLToPath(import.meta.url));
// Node 12 has no top-level await: the reference module is bound in main() below.
let ShallowWater, G, hllc, reflect, outflow, periodic, flather, makeSponge, BCNAME;

// ---- encoding helpers -------------------------------------------------------
const f64hex = (x) => { const b = Buffer.alloc(8); b.writeDoubleLE(x, 0); return Buffer.from(b).reverse().toString('hex'); };
const b64 = (arr) => Buffer.from(new Float64Array(arr).buffer).toString('base64');
function fnv1a64(arrays) {
  let h = 0xcbf29ce484222325n;
  const P = 0x100000001b3n, M = 0xffffffffffffffffn;
  for (const a of arrays) {
    const bytes = new Uint8Array(a.buffer, a.byteOffset, a.byteLength);
    for (let i = 0; i < bytes.length; i++) { h ^= BigInt(bytes[i]); h = (h * P) & M; }
  }
  return h.toString(16).padStart(16, '0');
}
// mulberry32: a small seeded 32-bit PRNG, used only to make scenario inputs
function rng(seed) {
  let s = seed >>> 0;
  return () => {
    s = (s + 0x6D2B79F5) >>> 0;
    let t = Math.imul(s ^ (s >>> 15), 1 | s);
    t = (t + Math.imul(t ^ (t >>> 7), 61 | t)) ^ t;
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
}
const ROWS = ['dxRow', 'dyRowN', 'dyRowS', 'dyCFL', 'bedPhi', 'tanPhi', 'fRow', 'geoCoef', 'area'];

function describe(sim, opts, bcs) {
  const g = sim.geom;
  const rows = {};
  for (const r of ROWS) rows[r] = b64(g[r]);
  return {
    node: process.version,
    G: f64hex(G),
    params: {
      nx: sim.nx, ny: sim.ny, dx: sim.dx === undefined ? null : f64hex(sim.dx), dy: f64hex(sim.dy),
      sphere: opts.sphere || null, order: sim.order, limiter: sim.limiterName,
      manning: f64hex(sim.manning), minDepth: f64hex(sim.minDepth), cfl: f64hex(sim.cfl),
      coriolis: f64hex(sim.coriolis), eps4: f64hex(sim.eps4), R: f64hex(g.R === undefined ? Infinity : g.R),
      kind: g.kind,
    },
    bc: bcs,
    rows,
  };
}

// Run a scenario: build(sim) sets up BCs/forcing; each step is self-timed
// (dt = null -> maxDt) unless fixedDt is given.
function runScenario(name, ctor, { steps = 50, fixedDt = null, setup = null, ext = null,
  sponge = null, checkpoints = [1, 10], hashOnly = false, mfield = null, until = null } = {}) {
  const sim = new ShallowWater(ctor);
  const bcs = ['reflect', 'reflect', 'reflect', 'reflect'];
  const sides = ['west', 'east', 'south', 'north'];
  if (setup) setup(sim);
  let extFns = [null, null, null, null];
  if (ext) {
    // ext: {side: [etaFn, uFn]} -> flather boundaries
    for (const [side, [ef, uf]] of Object.entries(ext)) {
      const si = sides.indexOf(side);
      sim.boundaries[side] = flather(ef, uf);
      extFns[si] = [ef, uf];
    }
  }
  for (let s = 0; s < 4; s++) {
    const f = sim.boundaries[sides[s]];
    bcs[s] = extFns[s] ? 'flather' : BCNAME.get(f);
    if (!bcs[s]) throw new Error(`${name}: unrecognised boundary on ${sides[s]}`);
  }
  if (mfield) sim.setManningField(mfield);
  let spongeRec = null;
  if (sponge) {
    // Re-derive the sponge's per-cell strengths exactly as makeSponge does by
    // probing the closure: run it on a scratch copy with h=1, hu=1, b=-big so
    // that hu -> hu - hu*f reveals f = str*dt*60 with dt=1/60 -> str (when < 1).
    const forcing = makeSponge(sim, sponge);
    const N = sim.W * sim.H;
    const probe = { t: 0, b: new Float64Array(N).fill(-1e6), h: new Float64Array(N).fill(1),
      hu: new Float64Array(N).fill(1), hv: new Float64Array(N).fill(1) };
    forcing(probe, 1 / 60);
    const str = new Float64Array(N);
    for (let k = 0; k < N; k++) str[k] = 1 - probe.hu[k];
    sim.forcing = forcing;
    spongeRec = { strength_probe: b64(str), width: sponge.width, strength: f64hex(sponge.strength),
      side: sponge.side, etaRef: f64hex(sponge.etaRef || 0) };
  }
  const rec = describe(sim, ctor, bcs);
  rec.name = name;
  rec.init = { b: b64(sim.b), h: b64(sim.h), hu: b64(sim.hu), hv: b64(sim.hv) };
  rec.manningField = sim.manningField ? b64(sim.manningField) : null;
  rec.sponge = spongeRec;
  rec.steps = [];
  rec.checkpoints = {};
  const vol0 = sim.volume();
  for (let n = 1; n <= steps; n++) {
    const e = [0, 0, 0, 0, 0, 0, 0, 0];
    for (let s = 0; s < 4; s++) if (extFns[s]) { e[2 * s] = extFns[s][0](sim.t); e[2 * s + 1] = extFns[s][1](sim.t); }
    let dtArg = fixedDt;
    if (until !== null) dtArg = Math.min(sim.maxDt(), until - sim.t);
    const t0 = sim.t;
    const dt = dtArg === null ? sim.step() : sim.step(dtArg);
    rec.steps.push({
      dt: f64hex(dt), dtArg: dtArg === null ? null : f64hex(dtArg), t0: f64hex(t0),
      ext: e.map(f64hex), mf: f64hex(sim.massFloored === undefined ? 0 : sim.massFloored),
      hash: fnv1a64([sim.h, sim.hu, sim.hv]), finite: sim.finite(),
    });
    if (!hashOnly && checkpoints.includes(n)) {
      rec.checkpoints[n] = { h: b64(sim.h), hu: b64(sim.hu), hv: b64(sim.hv) };
      if (n === 1) Object.assign(rec.checkpoints[n], { b: b64(sim.b), theta: b64(sim.theta),
        rh: b64(sim.rh), rhu: b64(sim.rhu), rhv: b64(sim.rhv) });
    }
    if (until !== null && sim.t >= until) break;
  }
  rec.final = {
    t: f64hex(sim.t), steps: sim.steps, volume: f64hex(sim.volume()), volume0: f64hex(vol0),
    maxSpeed: f64hex(sim.maxSpeed()), finite: sim.finite(),
    maxSurfaceDeviation0: f64hex(sim.maxSurfaceDeviation(0)), energy: f64hex(sim.energy()),
  };
  if (!hashOnly) rec.final.state = { h: b64(sim.h), hu: b64(sim.hu), hv: b64(sim.hv), b: b64(sim.b) };
  return rec;
}

// Direct residual probe, the verify-physics pattern: applyBC(); residual(dt).
function probeScenario(name, ctor, { setup = null, dts = [] } = {}) {
  const sim = new ShallowWater(ctor);
  if (setup) setup(sim);
  const bcs = ['west', 'east', 'south', 'north'].map((s) => BCNAME.get(sim.boundaries[s]));
  const rec = describe(sim, ctor, bcs);
  rec.name = name;
  rec.init = { b: b64(sim.b), h: b64(sim.h), hu: b64(sim.hu), hv: b64(sim.hv) };
  const maxDt = sim.maxDt();
  rec.maxDt = f64hex(maxDt);
  rec.probes = [];
  for (const mult of dts) {
    const s2 = new ShallowWater(ctor);
    if (setup) setup(s2);
    s2.applyBC();
    const dt = mult * maxDt;
    s2.residual(dt);
    rec.probes.push({ mult, dt: f64hex(dt), after_bc: { b: b64(s2.b), h: b64(s2.h), hu: b64(s2.hu), hv: b64(s2.hv) },
      rh: b64(s2.rh), rhu: b64(s2.rhu), rhv: b64(s2.rhv), theta: b64(s2.theta),
      fxM: b64(s2.fxM), fyM: b64(s2.fyM) });
  }
  return rec;
}

// ---- scenarios ---------------------------------------------------------------
const SC = {};
SC.lake_island = () => runScenario('lake_island', {
  nx: 40, ny: 24, dx: 10, manning: 0, eta0: 0,
  bed: (x, y) => -3 + 5 * Math.exp(-(((x - 200) ** 2 + (y - 120) ** 2) / 3600)) + 0.2 * Math.sin(x * 0.05),
}, { steps: 60 });

SC.dambreak_dry = () => runScenario('dambreak_dry', {
  nx: 48, ny: 6, dx: 2.5, manning: 0, bed: (x) => 0.01 * x, eta0: (x) => (x < 50 ? 2 : -1),
}, { steps: 120, setup: (s) => { s.boundaries = { west: reflect, east: outflow, south: periodic, north: periodic }; } });

SC.wavy_coriolis = () => runScenario('wavy_coriolis', {
  nx: 32, ny: 28, dx: 50, manning: 0, coriolis: 1e-3,
  bed: (x, y) => -20 + 4 * Math.sin(x * 0.004) * Math.cos(y * 0.005),
  eta0: (x, y) => 0.4 * Math.exp(-(((x - 700) ** 2 + (y - 600) ** 2) / 40000)),
}, { steps: 80 });

SC.order1_island = () => runScenario('order1_island', {
  nx: 36, ny: 20, dx: 5, manning: 0, order: 1,
  bed: (x, y) => -1 + 2.2 * Math.exp(-(((x - 90) ** 2 + (y - 50) ** 2) / 900)),
  eta0: (x, y) => (x < 40 ? 0.6 : 0),
}, { steps: 80 });

SC.minmod_beach = () => runScenario('minmod_beach', {
  nx: 44, ny: 12, dx: 4, manning: 0, limiter: 'minmod',
  bed: (x, y) => -2 + 0.02 * x + 0.1 * Math.cos(y * 0.2),
  eta0: (x) => 0.3 * Math.exp(-(((x - 40) / 15) ** 2)),
}, { steps: 100 });

SC.central_wave = () => runScenario('central_wave', {
  nx: 40, ny: 10, dx: 5, manning: 0, limiter: 'central',
  bed: () => -5, eta0: (x, y) => 0.05 * Math.sin(x * 0.03) * Math.cos(y * 0.1),
}, { steps: 60, setup: (s) => { s.boundaries = { west: periodic, east: periodic, south: reflect, north: reflect }; } });

SC.manning_scalar = () => runScenario('manning_scalar', {
  nx: 30, ny: 12, dx: 5, manning: 0.03, bed: (x, y) => -1.5 + 0.1 * Math.sin(y * 0.3), eta0: 0,
}, {
  steps: 60, fixedDt: 0.25,
  setup: (s) => {
    s.boundaries = { west: periodic, east: periodic, south: reflect, north: reflect };
    for (let k = 0; k < s.h.length; k++) { s.hu[k] = 1.2 * s.h[k]; s.hv[k] = 0.05 * s.h[k]; }
  },
  checkpoints: [1, 2, 10, 30, 60],
});

SC.manning_field = () => {
  const r = rng(7);
  const vals = [];
  for (let i = 0; i < 5000; i++) vals.push(r() < 0.15 ? 0 : 0.01 + 0.05 * r());
  return runScenario('manning_field', {
    nx: 40, ny: 16, dx: 3, manning: 0.025,
    bed: (x) => -1.2 + 0.012 * x, eta0: (x) => (x < 30 ? 0.4 : 0),
  }, {
    steps: 120,
    mfield: (x, y) => vals[(Math.floor(x * 7.3 + y * 13.1) % 5000 + 5000) % 5000],
    setup: (s) => { s.boundaries = { west: reflect, east: reflect, south: periodic, north: periodic }; },
    checkpoints: [1, 2, 10, 60, 120],
  });
};

SC.flather_tide = () => runScenario('flather_tide', {
  nx: 40, ny: 14, dx: 20, manning: 0,
  bed: (x, y) => -4 + 0.006 * x + 0.3 * Math.sin(y * 0.02), eta0: 0,
}, {
  steps: 120,
  ext: {
    west: [(t) => 0.5 * Math.cos(2 * Math.PI * t / 300), (t) => 0.05 * Math.sin(2 * Math.PI * t / 300)],
    north: [(t) => 0.1, () => 0],
  },
  setup: (s) => { s.boundaries = { west: reflect, east: reflect, south: outflow, north: reflect }; },
});

SC.periodic_all = () => runScenario('periodic_all', {
  nx: 28, ny: 24, dx: 8, manning: 0,
  bed: (x, y) => -6 + 2 * Math.exp(-(((x - 100) ** 2 + (y - 90) ** 2) / 2500)),
  eta0: (x, y) => 0.2 * Math.sin(x * 0.03),
}, {
  steps: 80,
  setup: (s) => {
    s.boundaries = { west: periodic, east: periodic, south: periodic, north: periodic };
    for (let k = 0; k < s.h.length; k++) { s.hu[k] = 0.7 * s.h[k]; s.hv[k] = -0.3 * s.h[k]; }
  },
});

SC.sponge_wave = () => runScenario('sponge_wave', {
  nx: 48, ny: 6, dx: 5, manning: 0, bed: () => -10, eta0: 0,
}, {
  steps: 100,
  ext: { west: [(t) => 0.1 * Math.sin(2 * Math.PI * t / 20), () => 0] },
  setup: (s) => { s.boundaries = { west: reflect, east: reflect, south: periodic, north: periodic }; },
  sponge: { side: 'east', width: 12, strength: 0.08, etaRef: 0 },
});

SC.sphere_pole = () => runScenario('sphere_pole', {
  nx: 36, ny: 18, manning: 0,
  sphere: { R: 6371e3, lat0: -90, lat1: 90, omega: 7.292115e-5 },
  bed: (lon, lat) => -4000 + 1500 * Math.sin(3 * lon * Math.PI / 180) * Math.cos(2 * lat * Math.PI / 180)
    + (lon > 20 && lon < 60 && lat > 0 && lat < 40 ? 4200 : 0),
  eta0: (lon, lat) => 2 * Math.exp(-(((lon + 60) / 25) ** 2 + ((lat - 20) / 20) ** 2)),
}, { steps: 40 });

SC.sphere_band_o1 = () => runScenario('sphere_band_o1', {
  nx: 40, ny: 26, manning: 0.02, order: 1,
  sphere: { R: 6371e3, lat0: -60, lat1: 70, omega: 7.292115e-5 },
  bed: (lon, lat) => -3000 + 800 * Math.cos(lon * Math.PI / 90),
  eta0: (lon, lat) => 1.5 * Math.exp(-(((lon - 10) / 30) ** 2 + ((lat - 5) / 15) ** 2)),
}, { steps: 40, checkpoints: [1, 2, 10, 40] });

SC.drying_flat = () => runScenario('drying_flat', {
  nx: 50, ny: 8, dx: 10, manning: 0.025,
  bed: (x, y) => -2 + 0.008 * x + 0.05 * Math.sin(y * 0.5), eta0: 0,
}, {
  steps: 200,
  ext: { west: [(t) => 1.2 * Math.sin(2 * Math.PI * t / 400), () => 0] },
  setup: (s) => { s.boundaries = { west: reflect, east: reflect, south: periodic, north: periodic }; },
  checkpoints: [1, 2, 20, 100, 200],
});

SC.blowup = () => runScenario('blowup', {
  nx: 24, ny: 10, dx: 2, manning: 0, cfl: 4.0,
  bed: () => 0, eta0: (x) => (x < 20 ? 3 : 0.5),
}, { steps: 30, hashOnly: true });

SC.probe_positivity = () => probeScenario('probe_positivity', {
  nx: 30, ny: 3, dx: 2, manning: 0, bed: () => 0, eta0: (x) => (x > 28 && x < 32 ? 1 : 0),
}, {
  dts: [0.5, 1, 5, 20],
  setup: (s) => { s.boundaries = { west: reflect, east: reflect, south: periodic, north: periodic }; },
});

SC.probe_sphere = () => probeScenario('probe_sphere', {
  nx: 24, ny: 12, manning: 0,
  sphere: { R: 6371e3, lat0: -90, lat1: 90, omega: 7.292115e-5 },
  bed: (lon, lat) => -3000 + 1000 * Math.cos(lat * Math.PI / 60) * Math.sin(lon * Math.PI / 45),
  eta0: (lon, lat) => 0.5 * Math.sin(lon * Math.PI / 90),
}, { dts: [1, 3] });

SC.hllc_probe = () => {
  const r = rng(11);
  const cases = [];
  for (let n = 0; n < 4000; n++) {
    const pick = () => { const u = r(); return u < 0.2 ? 0 : u < 0.3 ? 5e-5 * r() : u < 0.4 ? 1e-4 : 10 * r() ** 2; };
    const hL = pick(), hR = pick();
    const huL = hL * (r() - 0.5) * 8, hvL = hL * (r() - 0.5) * 4;
    const huR = hR * (r() - 0.5) * 8, hvR = hR * (r() - 0.5) * 4;
    const dry = r() < 0.5 ? 1e-4 : 1e-3;
    const out = [0, 0, 0];
    hllc(out, hL, huL, hvL, hR, huR, hvR, dry);
    cases.push([hL, huL, hvL, hR, huR, hvR, dry, out[0], out[1], out[2]]);
  }
  return { name: 'hllc_probe', node: process.version, G: f64hex(G), cases: b64(cases.flat()) };
};

SC.vel_probe = () => {
  const r = rng(5);
  const sim = new ShallowWater({ nx: 2, ny: 2, dx: 1, bed: () => 0 });
  const vals = [];
  for (let n = 0; n < 4000; n++) {
    const u = r();
    const h = u < 0.1 ? 0 : u < 0.3 ? 1e-3 * r() * 2 : u < 0.4 ? 1e-3 : 20 * r() ** 3;
    const hq = (r() - 0.5) * 10 * (h + 1e-4);
    vals.push([h, hq, sim.vel(hq, h)]);
  }
  return { name: 'vel_probe', node: process.version, eps4: f64hex(sim.eps4), cases: b64(vals.flat()) };
};

SC.math_probe = () => {
  // Math.hypot and Math.pow(x, 4/3) bit patterns from the engine, for the oracle's restatements
  const r = rng(3);
  const vals = [];
  for (let n = 0; n < 8000; n++) {
    const x = Math.pow(10, -3.5 + 8 * r());
    const a = (r() - 0.5) * Math.pow(10, -6 + 8 * r());
    const b = (r() - 0.5) * Math.pow(10, -6 + 8 * r());
    vals.push([x, Math.pow(x, 4 / 3), a, b, Math.hypot(a, b)]);
  }
  return { name: 'math_probe', node: process.version, cases: b64(vals.flat()) };
};

// Thacker radially symmetric bowl (Thacker 1981), 256x256, dx = 10 m, A = 0.5,
// run to two periods with the reference's self-timed step; g is the reference's own.
SC.thacker256 = () => {
  const a = 1000, h0 = 1, A = 0.5, n = 256, dx = 10, L = n * dx;
  const c = L / 2;
  const omega = Math.sqrt(8 * G * h0) / a;
  const T = 2 * Math.PI / omega;
  const eta0 = (x, y) => {
    const r2 = (x - c) ** 2 + (y - c) ** 2;
    return h0 + h0 * (Math.sqrt(1 - A * A) / (1 - A) - 1 - (r2 / (a * a)) * ((1 - A * A) / ((1 - A) ** 2) - 1));
  };
  const bed = (x, y) => h0 * (((x - c) ** 2 + (y - c) ** 2) / (a * a));
  const rec = runScenario('thacker256', { nx: n, ny: n, dx, manning: 0, bed, eta0 },
    { steps: 100000, until: 2 * T, hashOnly: true });
  rec.period = f64hex(T);
  return rec;
};

async function main() {
  const swe = await import(url.pathToFileURL(path.join(REF_DIR, 'src/swe.mjs')).href);
  ({ ShallowWater, G, hllc, reflect, outflow, periodic, flather, makeSponge } = swe);
  BCNAME = new Map([[reflect, 'reflect'], [outflow, 'outflow'], [periodic, 'periodic']]);
  const wanted = process.argv.slice(2);
  const names = wanted.length ? wanted : Object.keys(SC).filter((k) => k !== 'thacker256');
  for (const name of names) {
  if (!SC[name]) throw new Error(`unknown scenario ${name}`);
  const t0 = Date.now();
  const rec = SC[name]();
  const file = path.join(here, `${name}.json.gz`);
  fs.writeFileSync(file, zlib.gzipSync(Buffer.from(JSON.stringify(rec)), { level: 9 }));
  console.log(`${name}: ${(fs.statSync(file).size / 1024).toFixed(1)} KiB in ${((Date.now() - t0) / 1000).toFixed(1)} s`);
  }
}
main().catch((e) => { console.error(e); process.exit(1); });
