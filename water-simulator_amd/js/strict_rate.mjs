// Throughput of the drop-in driven the way the reference's callers drive it:
// step() after step(), with sim.b/h/hu/hv public host arrays that the caller
// may read or edit between calls (every step uploads b, h, hu, hv and
// downloads h, hu, hv and b's ghost ring), against advance(n), which keeps the
// state on the device.  One JSON line per grid size.
//
//   node water-simulator_amd/js/strict_rate.mjs [n ...]     (default 256 1024 2048)
//
// Workload: an n x n basin at dx = 10 m -- a beach rising from 4 m deep in the
// west to 1 m above sea level in the east, 0.3 m ripples, a 0.5 m surface hump
// -- MC limiter, Manning 0.025, Coriolis 1e-4, outflow west, walls elsewhere,
// self-timed dt.
import { ShallowWater, reflect, outflow } from './swe.mjs';

function basin(n) {
  const L = 10 * n;
  const bed = (x, y) => -4 + 5 * x / L + 0.3 * Math.sin(x / 70) * Math.cos(y / 90);
  const eta0 = (x, y) => 0.5 * Math.exp(-(((x - 0.3 * L) ** 2 + (y - 0.5 * L) ** 2) / (0.01 * L * L)));
  const sim = new ShallowWater({ nx: n, ny: n, dx: 10, bed, eta0, manning: 0.025, coriolis: 1e-4 });
  sim.boundaries = { west: outflow, east: reflect, south: reflect, north: reflect };
  return sim;
}

const seconds = (t0) => Number(process.hrtime.bigint() - t0) * 1e-9;
const sizes = process.argv.slice(2).map(Number).filter((x) => x > 0);
for (const n of sizes.length ? sizes : [256, 1024, 2048]) {
  const sim = basin(n);
  for (let k = 0; k < 5; k++) sim.step();
  // about a second of strict stepping
  let steps = 0;
  let t0 = process.hrtime.bigint();
  while (steps < 10 || seconds(t0) < 1.0) { sim.step(); steps++; }
  const strict = n * n * steps / seconds(t0);
  const k = Math.max(400, 2 * steps);
  t0 = process.hrtime.bigint();
  const ran = sim.advance(k);
  const resident = n * n * ran / seconds(t0);
  console.log(JSON.stringify({ n, strict_steps: steps, strict_cell_updates_per_s: strict,
    strict_ms_per_step: 1e3 * n * n / strict, advance_steps: ran, advance_cell_updates_per_s: resident,
    finite: sim.finite() }));
}
