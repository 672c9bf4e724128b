// Times the REFERENCE solver (src/swe.mjs of Syntaxswine/water-simulator, read
// from REF_DIR, default /root/reference) single-threaded under Node, as shipped,
// on one of the benchmark's workloads: time_reference.py writes the bed and the
// initial surface (float64 files in the reference's W x H layout) and a JSON spec
// (grid, metric, Manning, Coriolis, CFL, boundaries, west Flather tide), and this
// script builds the reference's ShallowWater from them and times self-timed
// steps.  Prints one JSON line.
//
//   node oracle/time_reference.mjs SPEC.json
//
// Runs only in the build container (the reference does not travel); it is the
// CPU baseline's "reference" arm, timed the reference's way: the minimum over
// repeated runs of STEPS steps after one warm-up step.
import fs from 'fs';
import path from 'path';
import url from 'url';

function readF64(file) {
  const raw = fs.readFileSync(file);
  return new Float64Array(raw.buffer, raw.byteOffset, raw.byteLength / 8);
}

async function main() {
  const REF_DIR = process.env.REF_DIR || '/root/reference';
  const mod = await import(url.pathToFileURL(path.join(REF_DIR, 'src/swe.mjs')).href);
  const { ShallowWater, flather, reflect, outflow, periodic } = mod;
  const spec = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  const bed = readF64(spec.bed), eta = readF64(spec.eta);
  const kinds = { reflect, outflow, periodic };
  const times = [];
  let cells = 0, t = 0;
  for (let r = 0; r < spec.reps; r++) {
    // the constructor samples bed(x, y) and eta0(x, y) once per cell, row by row
    // (swe.mjs:221-230): hand it the prepared arrays in that order
    let kb = 0, ke = 0;
    const opts = { nx: spec.nx, ny: spec.ny, bed: () => bed[kb++], eta0: () => eta[ke++],
      manning: spec.manning, coriolis: spec.coriolis, cfl: spec.cfl };
    if (spec.sphere) opts.sphere = spec.sphere; else opts.dx = spec.dx;
    const sim = new ShallowWater(opts);
    if (kb !== sim.W * sim.H || ke !== sim.W * sim.H) throw new Error('unexpected sampling order');
    const sides = ['west', 'east', 'south', 'north'];
    spec.bc.forEach((k, s) => { if (k !== 'flather') sim.boundaries[sides[s]] = kinds[k]; });
    if (spec.flather) {
      const { amp, period } = spec.flather;
      sim.boundaries.west = flather((tt) => amp * Math.cos(2 * Math.PI * tt / period));
    }
    sim.step();
    const t0 = process.hrtime.bigint();
    for (let n = 0; n < spec.steps; n++) sim.step();
    times.push(Number(process.hrtime.bigint() - t0) * 1e-9);
    cells = spec.nx * spec.ny;
    t = sim.t;
    if (!sim.finite()) throw new Error('reference run went non-finite');
  }
  const best = Math.min(...times);
  console.log(JSON.stringify({
    workload: spec.workload, nx: spec.nx, ny: spec.ny, steps: spec.steps, repeats: spec.reps,
    seconds_best: best, seconds_all: times, cell_updates_per_s: cells * spec.steps / best,
    simulated_s: t, node: process.version, cores_used: 1,
  }));
}

main().catch((e) => { console.error(e); process.exit(1); });
