// Samples a gridlands-terrain export with the drop-in's JS adapter the way the
// solver's constructor does (every cell centre of the W*H grid, two ghost rings
// included, row by row: swe.mjs:222-231) and writes b, h and the Manning field
// as little-endian float64 to OUT, for comparison with the Python adapter.
//
//   node tests/js/gridlands_sample.mjs FIXTURE.json.gz OUT [settle]
import fs from 'fs';
import zlib from 'zlib';
import { gridlandsTerrain, manningSampler } from '../../water-simulator_amd/js/gridlands.mjs';

const [fixture, out, settle] = process.argv.slice(2);
const doc = JSON.parse(zlib.gunzipSync(fs.readFileSync(fixture)).toString('utf8'));
const t = gridlandsTerrain(doc, { settle: settle === 'settle' });
const ng = 2, Wt = t.domain.nx + 2 * ng, Ht = t.domain.ny + 2 * ng, N = Wt * Ht;
const res = new Float64Array(3 * N);
const nAt = manningSampler(t);
for (let j = 0; j < Ht; j++) {
  for (let i = 0; i < Wt; i++) {
    const k = j * Wt + i;
    const x = (i - ng + 0.5) * t.domain.dx, y = (j - ng + 0.5) * t.domain.dy;
    const b = t.bed(x, y);
    res[k] = b;
    res[N + k] = Math.max(0, t.eta0(x, y) - b);
    res[2 * N + k] = nAt(x, y);
  }
}
fs.writeFileSync(out, Buffer.from(res.buffer));
