// Waterbodies, invariants and the settle step of the drop-in's JS terrain
// adapter on grids given as JSON ({grids: [{W, H, bed: [...], surface: [...]}]}),
// written back as JSON, for comparison with the Python adapter and a
// brute-force flood fill.
//
//   node tests/js/gridlands_bodies.mjs IN.json OUT.json
import fs from 'fs';
import { waterbodies, verifyInvariants, settleLakes } from '../../water-simulator_amd/js/gridlands.mjs';

const [inp, out] = process.argv.slice(2);
const { grids } = JSON.parse(fs.readFileSync(inp, 'utf8'));
const res = grids.map(({ W, H, bed, surface }) => {
  const bedM = Float32Array.from(bed), surfaceM = Float32Array.from(surface);
  const bodies = waterbodies(bedM, surfaceM, W, H).map((b) => ({
    level: b.level, cells: Array.from(b.cells), spread: b.spread,
    ringMin: Number.isFinite(b.ringMin) ? b.ringMin : null, isOcean: b.isOcean,
  }));
  const inv = verifyInvariants({ bedM, surfaceM, W, H });
  for (const k of Object.keys(inv)) if (!Number.isFinite(inv[k])) inv[k] = null;
  const settled = settleLakes(bedM, surfaceM, W, H);
  return { bodies, inv, settledSurface: Array.from(surfaceM), moved: settled.bodiesMoved };
});
fs.writeFileSync(out, JSON.stringify(res));
