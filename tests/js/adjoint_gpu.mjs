// The drop-in's Adjoint (forward/backward pair over the N-API addon) and its
// equilibrium tide, on the GPU: the gradient along a smooth direction against
// central finite differences of the replayed forward, for dJ/dh0 and dJ/dn.
//
//   node tests/js/adjoint_gpu.mjs
import { ShallowWater, Adjoint } from '../../water-simulator_amd/js/swe.mjs';

let fails = 0, checks = 0;
function report(label, ok, note) {
  checks++;
  if (!ok) fails++;
  console.log(`${ok ? 'PASS' : 'FAIL'}  ${label}${note ? '  -- ' + note : ''}`);
}
function xorshift(seed) {
  let s = seed >>> 0 || 1;
  return () => { s ^= s << 13; s >>>= 0; s ^= s >>> 17; s ^= s << 5; s >>>= 0; return s / 4294967296; };
}
const dot = (a, b) => { let s = 0; for (let i = 0; i < a.length; i++) s += a[i] * b[i]; return s; };

function gradientCheck(label, sim, nsteps, gauges, every, mkField) {
  const N = sim.W * sim.H;
  const h0 = Float64Array.from(sim.h);
  const nf = mkField(N);
  const adj = new Adjoint(sim, nsteps);
  adj.setGauges(gauges, every, new Float64Array(Math.floor(nsteps / every) * gauges.length));
  adj.forward(h0, nf);
  const obs = adj.series().map((v, i) => 0.5 * v + 0.01 * Math.sin(i));
  adj.setGauges(gauges, every, obs);
  const J = adj.forward(h0, nf);
  const { dh0, dn } = adj.backward();
  const dh = new Float64Array(N), dnf = new Float64Array(N);
  const r = xorshift(5);
  for (let j = 2; j < sim.H - 2; j++) {
    for (let i = 2; i < sim.W - 2; i++) {
      const k = j * sim.W + i;
      if (h0[k] > 0.2) { dh[k] = 0.01 * Math.sin(i / 5) * Math.cos(j / 4); dnf[k] = 1e-2 * (r() - 0.5); }
    }
  }
  const eps = 1e-3;
  const plus = h0.map((v, k) => v + eps * dh[k]), minus = h0.map((v, k) => v - eps * dh[k]);
  const fd = (adj.loss(plus, nf) - adj.loss(minus, nf)) / (2 * eps);
  const ad = dot(dh0, dh);
  report(`${label}: dJ/dh0 against central differences`, Math.abs(ad - fd) <= 1e-6 * Math.abs(fd),
    `${ad} vs ${fd}, J = ${J}`);
  const np = nf.map((v, k) => v + eps * dnf[k]), nm = nf.map((v, k) => v - eps * dnf[k]);
  const fdn = (adj.loss(h0, np) - adj.loss(h0, nm)) / (2 * eps);
  const adn = dot(dn, dnf);
  report(`${label}: dJ/dn against central differences`, Math.abs(adn - fdn) <= 1e-6 * Math.abs(fdn), `${adn} vs ${fdn}`);
  const t = adj.timing();
  report(`${label}: timings reported`, t.forwardMs > 0 && t.backwardMs > 0, `${t.forwardMs.toFixed(2)} / ${t.backwardMs.toFixed(2)} ms`);
}

{
  // a rough Cartesian basin, central slopes (no limiter kinks for the differences to cross)
  const r = xorshift(3);
  const sim = new ShallowWater({ nx: 40, ny: 30, dx: 25, manning: 0.025, limiter: 'central',
    bed: () => -4 + 0.1 * r(), eta0: (x, y) => 0.25 * Math.exp(-((x - 350) ** 2 + (y - 300) ** 2) / 200 ** 2) + 0.01 * r() });
  const W = sim.W;
  const gauges = [5 * W + 7, 12 * W + 20, 20 * W + 30, 25 * W + 9];
  gradientCheck('cartesian', sim, 40, gauges, 4, (N) => { const r2 = xorshift(9); return Float64Array.from({ length: N }, () => 0.015 + 0.02 * r2()); });
}
{
  // a band of a small planet under an M2 + K1 equilibrium tide
  const r = xorshift(4);
  const sim = new ShallowWater({ nx: 36, ny: 20, manning: 0.02, limiter: 'central',
    sphere: { R: 2e6, lat0: -50, lat1: 50, omega: 7.2921e-5 }, bed: () => -300 + 5 * r(), eta0: () => 0.05 * r() });
  sim.setTide([{ m: 2, amp: 2.0, omega: 1.405189025e-4, phase: 0.2 }, { m: 1, amp: 1.0, omega: 7.292115e-5, phase: 0.7 }]);
  const W = sim.W;
  const gauges = [6 * W + 5, 10 * W + 12, 14 * W + 19, 18 * W + 26, 20 * W + 33];
  gradientCheck('sphere + tide', sim, 40, gauges, 4, (N) => Float64Array.from({ length: N }, () => 0.02));
}

{
  // two-level checkpointing (outer checkpoints every 12 steps, segments of 4
  // recomputed from them) gives the one-level gradient
  const mk = () => {
    const r = xorshift(6);                      // the same bed for both runs
    return new ShallowWater({ nx: 32, ny: 24, dx: 25, manning: 0.025,
      bed: () => -4 + 0.1 * r(), eta0: (x, y) => 0.2 * Math.exp(-((x - 300) ** 2 + (y - 250) ** 2) / 150 ** 2) });
  };
  const grads = [];
  for (const opts of [{ checkpointEvery: 4 }, { checkpointEvery: 4, outerEvery: 12 }]) {
    const sim = mk();
    const W = sim.W;
    const gauges = [5 * W + 7, 12 * W + 20, 20 * W + 30];
    const adj = new Adjoint(sim, 48, opts);
    const h0 = Float64Array.from(sim.h);
    adj.setGauges(gauges, 4, new Float64Array(12 * gauges.length).fill(0.01));
    adj.forward(h0, null);
    grads.push(adj.backward());
  }
  let d = 0, m = 0;
  for (let k = 0; k < grads[0].dh0.length; k++) {
    d = Math.max(d, Math.abs(grads[0].dh0[k] - grads[1].dh0[k]));
    m = Math.max(m, Math.abs(grads[0].dh0[k]));
  }
  report('two-level checkpointing gives the one-level dJ/dh0', m > 0 && d <= 1e-12 * m, `max diff ${d} of ${m}`);
}

console.log(`\n${fails === 0 ? 'ALL PASS' : 'FAILURES'} ${checks - fails}/${checks}`);
process.exit(fails === 0 ? 0 : 1);
