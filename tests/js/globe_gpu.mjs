// The drop-in's globe harness (water-simulator_amd/js/globe.mjs) on the GPU:
// the reference's own checks of measureModePeriod and polarTimestepTax
// (src/globe.mjs:2891-2944), run through ShallowWater.projectedRun.
//
//   node tests/js/globe_gpu.mjs
import { ShallowWater } from '../../water-simulator_amd/js/swe.mjs';
import { measureModePeriod, polarTimestepTax, legendreP } from '../../water-simulator_amd/js/globe.mjs';

const G_REF = 9.80665, R_REF = 6371e3;
let fails = 0, checks = 0;
function report(label, ok, note) {
  checks++;
  if (!ok) fails++;
  console.log(`${ok ? 'PASS' : 'FAIL'}  ${label}${note ? '  -- ' + note : ''}`);
}

const plain = new ShallowWater({ nx: 120, ny: 60, sphere: { R: R_REF, lat0: -90, lat1: 90, omega: 0 },
  bed: () => -4000, eta0: 0, manning: 0 });
const tax = polarTimestepTax(plain);
report('the polar timestep tax is measured off the metric', tax.ratio > 5,
  `dt ${tax.pole.toFixed(2)} s at lat ${tax.poleLat.toFixed(1)} vs ${tax.eq.toFixed(2)} s at lat ${tax.eqLat.toFixed(1)}, `
  + `ratio ${tax.ratio.toFixed(1)}x at 3 deg`);

for (const deg of [6, 4]) {
  const t0 = Date.now();
  const r = measureModePeriod({ n: 2, deg, H: 4000, R: R_REF, gRef: G_REF, amp: 1, periods: 2 });
  report(`n = 2 period at ${deg} deg vs g H n(n+1)/R^2`, Math.abs(r.rel) < 0.01,
    `${(r.measured / 3600).toFixed(4)} h vs ${(r.closed / 3600).toFixed(4)} h, ${(100 * r.rel).toFixed(4)}%; `
    + `${r.note}; ${((Date.now() - t0) / 1000).toFixed(2)} s`);
  report(`n = 2 at ${deg} deg is late, not early`, r.rel < 0);
}
const r3 = measureModePeriod({ n: 3, deg: 6, H: 4000, R: R_REF, gRef: G_REF, amp: 1, periods: 2 });
report('n = 3 period at 6 deg', Math.abs(r3.rel) < 0.01, `${(100 * r3.rel).toFixed(4)}%`);
report('n = 3 is faster than n = 2 by sqrt(n(n+1))',
  r3.closed < 0.72 * 2 * Math.PI * R_REF / Math.sqrt(G_REF * 4000 * 2 * 3));
const r1 = measureModePeriod({ n: 2, deg: 1, H: 4000, R: R_REF, gRef: G_REF, amp: 1, periods: 2 });
report('at 1 deg the error falls below 1e-4', Math.abs(r1.rel) < 1e-4 && r1.rel < 0,
  `${(100 * r1.rel).toFixed(5)}%, ${r1.steps} steps`);
let threw = false;
try { measureModePeriod({ n: 2, deg: 12, R: R_REF }); } catch (e) { threw = true; }
report('the measurement refuses to build its own target from the solver\'s G', threw);
let threwN = false;
try { legendreP(1, 0.3); } catch (e) { threwN = true; }
report('n = 1 is refused (rigid translation of the shell)', threwN);

console.log(fails ? `${fails} FAILED of ${checks}` : `ALL PASS ${checks}/${checks}`);
process.exitCode = fails ? 1 : 0;
