// Times the REFERENCE solver (src/swe.mjs of Syntaxswine/water-simulator, read
// from REF_DIR, default /root/reference) single-threaded under Node, as shipped,
// on a sample of the 4096x4096 large-grid recipe: the bed comes from a float64
// file written by time_reference.py (the same seeded SplitMix64 recipe the
// benchmark uses), west Flather M2 tide of 1 m, Manning 0.025, f = 1e-4,
// CFL 0.45, self-timed steps.  Prints one JSON line.
//
//   node oracle/time_reference.mjs BED_FILE NX NY STEPS REPEATS
//
// Runs only in the build container (the reference does not travel); it is the
// CPU baseline's "reference" arm, timed the reference's way: the minimum over
// repeated runs of STEPS steps after one warm-up step.
import fs from 'fs';
import path from 'path';
import url from 'url';

async function main() {
  const REF_DIR = process.env.REF_DIR || '/root/reference';
  const { ShallowWater, flather } = await import(url.pathToFileURL(path.join(REF_DIR, 'src/swe.mjs')).href);
  const [bedFile, nxS, nyS, stepsS, repS] = process.argv.slice(2);
  const nx = +nxS, ny = +nyS, steps = +stepsS, reps = +repS;
  const dx = 10;
  const raw = fs.readFileSync(bedFile);
  const bed = new Float64Array(raw.buffer, raw.byteOffset, raw.byteLength / 8);
  const W = nx + 4;
  // cellCentre(i, j) = ((i + 0.5) dx, (j + 0.5) dy) for interior indices from -2
  const bedAt = (x, y) => bed[(Math.round(y / dx - 0.5) + 2) * W + Math.round(x / dx - 0.5) + 2];
  const times = [];
  let dts = 0;
  for (let r = 0; r < reps; r++) {
    const sim = new ShallowWater({ nx, ny, dx, bed: bedAt, eta0: 0, manning: 0.025, coriolis: 1e-4, cfl: 0.45 });
    sim.boundaries.west = flather((t) => Math.cos(2 * Math.PI * t / 44712));
    sim.step();
    const t0 = process.hrtime.bigint();
    for (let n = 0; n < steps; n++) sim.step();
    times.push(Number(process.hrtime.bigint() - t0) * 1e-9);
    dts = sim.t;
    if (!sim.finite()) throw new Error('reference run went non-finite');
  }
  const best = Math.min(...times);
  console.log(JSON.stringify({ node: process.version, nx, ny, steps, reps, seconds: times, t: dts,
    cell_updates_per_s: nx * ny * steps / best }));
}

main().catch((e) => { console.error(e); process.exit(1); });
