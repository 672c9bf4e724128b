// Sections 3-8 of the reference's tools/verify.mjs (Ritter's dam break,
// long-wave celerity, Thacker's bowl, Merian's seiche, the self-convergence study,
// numerical dissipation), restated over the drop-in solver on the GPU, and every figure
// those sections print held to the digits the reference prints for them.  The
// expected strings are the reference's own output under Node 12.22.9 (the
// engine the GPU host runs this script with); none of these runs touches
// Math.pow or Math.hypot, so they are engine-independent, and the README's
// claim table (README.md:76-87) quotes the same figures: Ritter 0.402-0.670%,
// celerity 0.011-0.307%, Thacker 0.004%, Merian 0.000%.
//
//   node tests/js/verify_sections_gpu.mjs
import { ShallowWater, reflect, periodic, outflow } from '../../water-simulator_amd/js/swe.mjs';

const G_REF = 9.80665;
let fails = 0, checks = 0;
// the reference's number format (tools/verify.mjs:57-61)
const fmt = (v) => {
  if (v === 0) return '0';
  const a = Math.abs(v);
  return (a < 1e-3 || a >= 1e5) ? v.toExponential(4) : v.toFixed(6);
};
// the reference's relative error (tools/verify.mjs:46), as it prints it
const relPct = (got, want) => (100 * (Math.abs(want) > 1e-12 ? Math.abs(got - want) / Math.abs(want)
  : Math.abs(got - want))).toFixed(3);
function same(label, got, want) {
  checks++;
  const ok = got === want;
  if (!ok) fails++;
  console.log(`${ok ? 'PASS' : 'FAIL'}  ${label}: ${got}${ok ? '' : `  (the reference prints ${want})`}`);
}
// the suites' stepping loop: step(min(maxDt, T - t)) until t reaches T
const runTo = (sim, T) => { while (sim.t < T) sim.step(Math.min(sim.maxDt(), T - sim.t)); };

// ---- 3. Ritter's dam break on a dry bed (tools/verify.mjs:195-232) --------------------------
{
  const h0 = 4, L = 2000, x0 = 1000, T = 20, c0 = Math.sqrt(G_REF * h0);
  const l1 = [];
  const want = { 0.5: ['116.875000', '117.663853', '0.670'], 0.2: ['165.625000', '166.496074', '0.523'],
    0.1: ['191.875000', '191.107451', '0.402'] };
  for (const nx of [800, 1600, 3200]) {
    const dx = L / nx;
    const sim = new ShallowWater({ nx, ny: 1, dx, bed: () => 0, eta0: (x) => (x < x0 ? h0 : 0), manning: 0 });
    sim.boundaries = { west: reflect, east: outflow, south: periodic, north: periodic };
    runTo(sim, T);
    let sum = 0;
    for (let i = 0; i < nx; i++) {
      const x = (i + 0.5) * dx - x0;
      const ex = x <= -c0 * T ? h0 : x >= 2 * c0 * T ? 0 : (2 * c0 - x / T) ** 2 / (9 * G_REF);
      sum += Math.abs(sim.depth(i, 0) - ex);
    }
    l1.push(sum / nx);
    if (nx !== 1600) continue;
    // the last cell centre deeper than the level, against the similarity solution's position
    const contour = (lvl) => { let p = 0; for (let i = 0; i < nx; i++) if (sim.depth(i, 0) > lvl) p = (i + 0.5) * dx; return p; };
    const exact = (lvl) => x0 + T * (2 * c0 - Math.sqrt(9 * G_REF * lvl));
    for (const lvl of [0.5, 0.2, 0.1]) {
      const got = contour(lvl) - x0, ex = exact(lvl) - x0;
      same(`Ritter h = ${lvl} m contour, m from the dam`, fmt(got), want[lvl][0]);
      same(`Ritter h = ${lvl} m contour, similarity solution`, fmt(ex), want[lvl][1]);
      same(`Ritter h = ${lvl} m contour, relative error %`, relPct(got, ex), want[lvl][2]);
    }
    const tip = 100 * (contour(0.01) - exact(0.01)) / (exact(0.01) - x0);
    same('Ritter h = 1 cm tip lag %', tip.toFixed(1), '-9.6');
  }
  same('Ritter mean |h - exact| at nx = 1600', fmt(l1[1]), '9.8898e-4');
  same('Ritter L1 under refinement', l1.map((e) => e.toExponential(2)).join(' -> '), '1.52e-3 -> 9.89e-4 -> 7.56e-4');
  same('Ritter convergence order', (Math.log2(l1[0] / l1[2]) / 2).toFixed(2), '0.50');
}

// ---- 4. long-wave celerity (tools/verify.mjs:252-300) ---------------------------------------
{
  const want = { 2: ['4.442280', '0.307'], 10: ['9.908797', '0.060'], 50: ['22.145787', '0.011'] };
  for (const h0 of [2, 10, 50]) {
    const c0 = Math.sqrt(G_REF * h0), nx = 1000, dx = 20000 / nx;
    const pulse = (x) => 0.01 * Math.exp(-(((x - 5000) / 700) ** 2));
    const sim = new ShallowWater({ nx, ny: 1, dx, bed: () => -h0, eta0: (x) => pulse(x), manning: 0, order: 2 });
    // a right-going simple wave: hu = c0 eta
    for (let i = 0; i < nx; i++) sim.hu[sim.idx(i, 0)] = c0 * pulse((i + 0.5) * dx);
    sim.boundaries = { west: outflow, east: outflow, south: periodic, north: periodic };
    const centroid = () => {
      let m = 0, mx = 0;
      for (let i = 0; i < nx; i++) { const w = Math.max(0, sim.eta(i, 0)); m += w; mx += w * (i + 0.5) * dx; }
      return mx / m;
    };
    const t1 = 1500 / c0, t2 = t1 + 9000 / c0;
    runTo(sim, t1);
    const xa = centroid(), ta = sim.t;
    runTo(sim, t2);
    const c = (centroid() - xa) / (sim.t - ta);
    same(`celerity at h = ${h0} m, m/s`, fmt(c), want[h0][0]);
    same(`celerity at h = ${h0} m, relative error %`, relPct(c, c0), want[h0][1]);
  }
}

// ---- 5. Thacker's planar oscillation in a parabolic bowl (tools/verify.mjs:318-399) ---------
{
  const h0 = 10, a = 3000, U = 5;
  const w = Math.sqrt(2 * G_REF * h0) / a, period = 2 * Math.PI / w, C = h0 - U * U / (4 * G_REF);
  const bed = (x) => h0 * ((x - 6000) / a) ** 2;
  const etaEx = (x, t) => C + (U * U / (4 * G_REF)) * Math.cos(2 * w * t) + (U * w / G_REF) * Math.sin(w * t) * (x - 6000);
  const nx = 1200, dx = 12000 / nx;
  const bowl = () => {
    const s = new ShallowWater({ nx, ny: 1, dx, bed: (x) => bed(x), eta0: (x) => etaEx(x, 0), manning: 0 });
    for (let i = 0; i < nx; i++) {
      const k = s.idx(i, 0);
      if (s.h[k] > s.minDepth) s.hu[k] = s.h[k] * U;   // u(0) = U cos 0
    }
    s.boundaries = { west: reflect, east: reflect, south: periodic, north: periodic };
    return s;
  };
  const sim = bowl();
  runTo(sim, 2 * period);
  let l1 = 0, cnt = 0, worst = 0;
  for (let i = 0; i < nx; i++) {
    const x = (i + 0.5) * dx, hx = etaEx(x, sim.t) - bed(x);
    if (hx < 0.15) continue;
    const e = Math.abs(sim.depth(i, 0) - hx);
    l1 += e; cnt++;
    if (e > worst) worst = e;
  }
  same('Thacker mean |h - exact| after 2 periods', fmt(l1 / cnt), '0.004043');
  same('Thacker cells compared', String(cnt), '596');
  same('Thacker worst |h - exact|', fmt(worst), '0.015972');
  // u at the bowl centre every 0.5 s over six periods; downward zero crossings
  const sim2 = bowl(), centre = Math.round(6000 / dx), hist = [];
  while (sim2.t < 6 * period) {
    let acc = 0;
    while (acc < 0.5 - 1e-12) acc += sim2.step(Math.min(sim2.maxDt(), 0.5 - acc));
    const k = sim2.idx(centre, 0);
    hist.push([sim2.t, sim2.vel(sim2.hu[k], sim2.h[k])]);
  }
  const zc = [];
  for (let n = 1; n < hist.length; n++) {
    const [ta, ua] = hist[n - 1], [tb, ub] = hist[n];
    if (ua > 0 && ub <= 0) zc.push(ta + (ua / (ua - ub)) * (tb - ta));
  }
  same('Thacker downward crossings in 6 periods', String(zc.length), '6');
  const measured = (zc[zc.length - 1] - zc[0]) / (zc.length - 1);
  same('Thacker period, s', fmt(measured), '1345.880856');
  same('Thacker closed-form period, s', fmt(period), '1345.940271');
  same('Thacker period relative error %', relPct(measured, period), '0.004');
}

// ---- 6. Merian's seiche in a closed basin (tools/verify.mjs:401-450) ------------------------
{
  const L = 5000, h = 20, nx = 500, dx = L / nx;
  const sim = new ShallowWater({ nx, ny: 1, dx, bed: () => -h, eta0: (x) => 0.05 * Math.cos(Math.PI * x / L), manning: 0 });
  sim.boundaries = { west: reflect, east: reflect, south: periodic, north: periodic };
  const period = 2 * L / Math.sqrt(G_REF * h), hist = [];
  // eta in the third cell every 1 s over 8.4 periods; downward zero crossings
  while (sim.t < 8.4 * period) {
    let acc = 0;
    while (acc < 1.0 - 1e-12) acc += sim.step(Math.min(sim.maxDt(), 1.0 - acc));
    hist.push([sim.t, sim.eta(2, 0)]);
  }
  const zc = [];
  for (let n = 1; n < hist.length; n++) {
    const [ta, ea] = hist[n - 1], [tb, eb] = hist[n];
    if (ea > 0 && eb <= 0) zc.push(ta + (ea / (ea - eb)) * (tb - ta));
  }
  const measured = (zc[zc.length - 1] - zc[0]) / (zc.length - 1);
  same('Merian: crossings in 8.4 periods', String(zc.length), '9');
  same('Merian: period, s', fmt(measured), '714.041145');
  same('Merian: 2L/sqrt(gh), s', fmt(period), '714.043491');
  same('Merian: relative error %', relPct(measured, period), '0.000');
}

// ---- 7. self-convergence against the same solver at nx = 1600 (tools/verify.mjs:453-491) ----
{
  const L = 1000, h0 = 5, T = 30;
  const build = (n) => {
    const s = new ShallowWater({ nx: n, ny: 1, dx: L / n, bed: () => -h0,
      eta0: (x) => 0.05 * Math.exp(-(((x - 300) / 60) ** 2)), manning: 0 });
    s.boundaries = { west: periodic, east: periodic, south: periodic, north: periodic };
    runTo(s, T);
    return s;
  };
  const fine = build(1600);
  const atFine = (x) => fine.eta(Math.min(1599, Math.floor(x / (L / 1600))), 0);
  const errs = [50, 100, 200].map((n) => {
    const s = build(n);
    let e = 0;
    for (let i = 0; i < n; i++) e += Math.abs(s.eta(i, 0) - atFine((i + 0.5) * (L / n)));
    return e / n;
  });
  same('self-convergence L1 errors', errs.map((e) => e.toExponential(3)).join('  '), '7.034e-4  2.126e-4  6.129e-5');
  same('self-convergence orders', `${Math.log2(errs[0] / errs[1]).toFixed(2)} and ${Math.log2(errs[1] / errs[2]).toFixed(2)}`,
    '1.73 and 1.79');
}

// ---- 8. numerical dissipation of a travelling wave (tools/verify.mjs:493-548) ---------------
{
  const h0 = 12, A = 0.02, T = 10, c0 = Math.sqrt(G_REF * h0), Lw = c0 * T;
  const wave = (nx, cpw, limiter) => {
    const dx = Lw / cpw;
    const opts = { nx, ny: 1, dx, bed: () => -h0, eta0: 0, manning: 0, cfl: 0.4 };
    if (limiter) opts.limiter = limiter;
    const s = new ShallowWater(opts);
    const k = 2 * Math.PI / Lw;
    for (let i = 0; i < nx; i++) {
      const e = A * Math.cos(k * ((i + 0.5) * dx)), kk = s.idx(i, 0);
      s.h[kk] = h0 + e; s.hu[kk] = c0 * e;
    }
    s.boundaries = { west: periodic, east: periodic, south: periodic, north: periodic };
    runTo(s, 15 * T);
    let m = 0;
    for (let i = 0; i < nx; i++) m = Math.max(m, Math.abs(s.eta(i, 0)));
    return m / A;
  };
  const want = { 20: '50.7', 40: '85.7', 80: '95.6' };
  const kept = [20, 40, 80].map((cpw) => {
    const r = wave(cpw * 4, cpw);
    same(`amplitude kept at ${cpw} cells/wavelength after 15 wavelengths, %`, (100 * r).toFixed(1), want[cpw]);
    return r;
  });
  same('amplitude kept at 40 cells/wavelength', fmt(kept[1]), '0.856888');
  const d = kept.map((r) => -Math.log(r));
  same('dissipation orders', `${Math.log2(d[0] / d[1]).toFixed(2)} and ${Math.log2(d[1] / d[2]).toFixed(2)}`, '2.14 and 1.78');
  same('minmod at 20 cells/wavelength, %', (100 * wave(80, 20, 'minmod')).toFixed(1), '3.9');
}

console.log(fails ? `${fails} FAILED of ${checks}` : `ALL PASS ${checks}/${checks}`);
process.exitCode = fails ? 1 : 0;
