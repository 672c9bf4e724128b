// The drop-in's device diagnostics and device-recorded runs, on the GPU:
// deviceDiagnostics() against the host methods (the reference's loops) on the
// same state, and recordRun() against tools/waves.mjs's recording loop run
// through step() on the host -- times, samples and final state bit for bit.
//
//   node tests/js/record_gpu.mjs
import { ShallowWater, reflect, outflow, periodic } from '../../water-simulator_amd/js/swe.mjs';

let fails = 0, checks = 0;
function report(label, ok, note) {
  checks++;
  if (!ok) fails++;
  console.log(`${ok ? 'PASS' : 'FAIL'}  ${label}${note ? '  -- ' + note : ''}`);
}
const bitsEqual = (a, b) => a.length === b.length && a.every((x, i) => Object.is(x, b[i]));

function basin(extra = {}) {
  const bed = (x, y) => -3 + 4 * x / 480 + 0.3 * Math.sin(y / 35);
  const eta0 = (x, y) => 0.4 * Math.exp(-(((x - 150) ** 2 + (y - 120) ** 2) / 4000));
  const sim = new ShallowWater({ nx: 96, ny: 48, dx: 5, bed, eta0, manning: 0.025, ...extra });
  sim.boundaries = { west: outflow, east: reflect, south: periodic, north: periodic };
  return sim;
}

// ---- deviceDiagnostics against the host methods ----------------------------------
{
  const sim = basin();
  for (let n = 0; n < 30; n++) sim.step();
  const d = sim.deviceDiagnostics(0.1);
  report('device energy = host energy()', Object.is(d.energy, sim.energy()), `${d.energy}`);
  report('device maxSpeed = host maxSpeed()', Object.is(d.maxSpeed, sim.maxSpeed()));
  report('device maxSurfaceDeviation = host', Object.is(d.maxSurfaceDeviation, sim.maxSurfaceDeviation(0.1)));
  report('device finite = host finite()', d.finite === sim.finite());
  const rel = Math.abs(d.volume - sim.volume()) / sim.volume();
  report('device Cartesian volume within the host running sum\'s rounding', rel < 1e-13, `rel ${rel.toExponential(2)}`);

  const R = 6371e3, nx = 72, ny = 36;
  const sph = new ShallowWater({ nx, ny, sphere: { R, lat0: -90, lat1: 90, omega: 7.292115e-5 },
    bed: (lon, lat) => -4000 + 3000 * Math.sin(3 * lon * Math.PI / 180) * Math.cos(2 * lat * Math.PI / 180),
    eta0: (lon, lat) => 2 * Math.exp(-(((lat - 20) / 15) ** 2)), manning: 0.025 });
  sph.advance(20);
  const s = sph.deviceDiagnostics(0);
  report('sphere: device volume = host volume() bit for bit', Object.is(s.volume, sph.volume()), `${s.volume}`);
  report('sphere: device energy = host energy()', Object.is(s.energy, sph.energy()));
  report('sphere: device maxSpeed = host maxSpeed()', Object.is(s.maxSpeed, sph.maxSpeed()));
}

// ---- recordRun against waves.mjs's loop ------------------------------------------
{
  const cells = [];
  const probe = basin();
  for (let q = 0; q < 40; q++) cells.push(probe.idx((q * 7) % 96, (q * 5) % 48));
  const T0 = 0.9, dtRec = 0.125, T = 3.0;

  const host = basin();
  const times = [], eta = [];
  let next = T0;
  while (host.t < T) {                  // tools/waves.mjs:221-228
    host.step(Math.min(host.maxDt(), Math.max(1e-6, Math.min(T, next) - host.t)));
    if (host.t >= next && host.t >= T0) {
      times.push(host.t);
      for (const k of cells) eta.push(host.b[k] + host.h[k]);
      next += dtRec;
    } else if (host.t >= next) next += dtRec;
  }

  const dev = basin();
  const rec = dev.recordRun(cells, { t0: T0, dtRec, until: T });
  report('recordRun: sample times equal the host loop\'s', bitsEqual(Array.from(rec.times), times),
    `${rec.times.length} samples`);
  report('recordRun: samples equal the host loop\'s', bitsEqual(Array.from(rec.eta), eta));
  report('recordRun: same steps and end time', dev.steps === host.steps && Object.is(dev.t, host.t),
    `${dev.steps} steps, t = ${dev.t}`);
  report('recordRun: the final state is the host loop\'s', bitsEqual(dev.h, host.h) && bitsEqual(dev.hu, host.hu)
    && bitsEqual(dev.hv, host.hv));
}

console.log(fails ? `${fails} FAILED of ${checks}` : `ALL PASS ${checks}/${checks}`);
process.exitCode = fails ? 1 : 0;
