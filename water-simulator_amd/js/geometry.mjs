// The metric rows the GPU step divides by: a transcription of src/geometry.mjs of
// Syntaxswine/water-simulator (cartesianGeometry :163-176, sphericalGeometry
// :190-299, cosLat :154-157, cellLonLat :312-315, interiorArea :325-329).
//
// Computed on the host in JavaScript, so the transcendentals are the engine's own
// Math.sin/cos/tan and every row is the same double the reference produces; the
// rows are then uploaded once (and again if a caller edits sim.geom).
// Cartesian rows are ASSIGNED the scalar spacings rather than derived from areas
// and face lengths, and spherical divisors stay divisors with Infinity at a pole
// face: both are what the step's bit-for-bit division requires.

export const R_EARTH = 6371e3;
export const OMEGA_EARTH = 7.292115e-5;

const HALF_PI = Math.PI / 2;
const clampLat = (p) => (p > HALF_PI ? HALF_PI : p < -HALF_PI ? -HALF_PI : p);
const reflectLat = (p) => (p > HALF_PI ? Math.PI - p : p < -HALF_PI ? -Math.PI - p : p);

// cos(latitude) through the co-latitude, exactly zero at a pole
function cosLat(p) {
  const co = HALF_PI - Math.abs(p);
  return co <= 0 ? 0 : Math.sin(co);
}

function rowsOf(kind, nx, ny, ng, H) {
  const f = () => new Float64Array(H);
  return {
    kind, nx, ny, ng, H,
    dxRow: f(), dyRowN: f(), dyRowS: f(), dyCFL: f(), area: f(),
    lx: f(), lyN: f(), lyS: f(), bedPhi: f(), phiC: f(), tanPhi: f(),
    fRow: f(), geoCoef: f(),
  };
}

export function cartesianGeometry({ nx, ny, dx, dy = dx, ng = 2, coriolis = 0 }) {
  const H = ny + 2 * ng;
  const g = rowsOf('cartesian', nx, ny, ng, H);
  g.dx = dx; g.dy = dy; g.R = Infinity;
  g.totalArea = nx * dx * ny * dy;
  for (let j = 0; j < H; j++) {
    g.dxRow[j] = dx; g.dyRowN[j] = dy; g.dyRowS[j] = dy; g.dyCFL[j] = dy;
    g.bedPhi[j] = dy;
    g.area[j] = dx * dy; g.lx[j] = dy; g.lyN[j] = dx; g.lyS[j] = dx;
    g.fRow[j] = coriolis;
  }
  return g;
}

export function sphericalGeometry({ nx, ny, R = R_EARTH, lat0 = -90, lat1 = 90, ng = 2, omega = 0 }) {
  if (!(lat1 > lat0)) throw new Error(`lat1 (${lat1}) must exceed lat0 (${lat0})`);
  if (lat0 < -90 || lat1 > 90) throw new Error(`latitudes must lie in [-90, 90]; got ${lat0}..${lat1}`);
  const reachesPole = lat0 <= -90 + 1e-12 || lat1 >= 90 - 1e-12;
  if (reachesPole && nx % 2 !== 0) {
    throw new Error(`nx must be even when the domain reaches a pole (got ${nx}): a ghost row `
      + 'beyond the pole is the row half a turn away in longitude.');
  }
  const H = ny + 2 * ng;
  const g = rowsOf('sphere', nx, ny, ng, H);
  const dlam = 2 * Math.PI / nx;
  const p0 = lat0 * Math.PI / 180, p1 = lat1 * Math.PI / 180;
  const dphi = (p1 - p0) / ny;
  g.R = R; g.dlam = dlam; g.dphi = dphi; g.lat0 = lat0; g.lat1 = lat1; g.omega = omega;
  g.dy = R * dphi;
  g.totalArea = 2 * Math.PI * R * R * (Math.sin(p1) - Math.sin(p0));
  // domain edges pinned exactly, interior faces accumulated
  const faceLat = (k) => (k === 0 ? p0 : k === ny ? p1 : p0 + k * dphi);
  for (let j = 0; j < H; j++) {
    const jj = j - ng;
    const rawS = faceLat(jj), rawN = faceLat(jj + 1);
    const pS = clampLat(rawS), pN = clampLat(rawN);
    const pC = 0.5 * (pS + pN);
    const half = 0.5 * (pN - pS);
    g.phiC[j] = pC;
    g.lyS[j] = R * dlam * cosLat(pS);
    g.lyN[j] = R * dlam * cosLat(pN);
    g.lx[j] = R * dphi;
    // product-form area; a row wholly beyond a pole takes its mirror's latitudes
    let A = 2 * R * R * dlam * cosLat(pC) * Math.sin(half);
    if (!(A > 0)) {
      const rS = reflectLat(rawS), rN = reflectLat(rawN);
      A = Math.abs(2 * R * R * dlam * cosLat(0.5 * (rS + rN)) * Math.sin(0.5 * (rN - rS)));
    }
    g.area[j] = A;
    g.dxRow[j] = A / g.lx[j];
    g.dyRowN[j] = g.lyN[j] > 0 ? A / g.lyN[j] : Infinity;
    g.dyRowS[j] = g.lyS[j] > 0 ? A / g.lyS[j] : Infinity;
    g.dyCFL[j] = Math.min(g.dyRowN[j], g.dyRowS[j], g.dy);
    g.bedPhi[j] = (g.lyN[j] + g.lyS[j]) > 0 ? 2 * A / (g.lyN[j] + g.lyS[j]) : Infinity;
    g.tanPhi[j] = Math.tan(pC);
    g.fRow[j] = 2 * omega * Math.sin(pC);
    g.geoCoef[j] = (g.lyN[j] - g.lyS[j]) / (4 * A);
  }
  for (let j = 0; j < H; j++) {
    if (!(g.area[j] > 0) || !isFinite(g.area[j])) throw new Error(`row ${j}: area ${g.area[j]}`);
    if (!(g.lyN[j] >= 0) || !(g.lyS[j] >= 0)) throw new Error(`row ${j}: negative face length`);
  }
  return g;
}

export function cellLonLat(g, i, j) {
  if (g.kind !== 'sphere') throw new Error('cellLonLat is spherical-only');
  return [(i + 0.5) * g.dlam * 180 / Math.PI - 180, g.phiC[j + g.ng] * 180 / Math.PI];
}

export function interiorArea(g) {
  let s = 0;
  for (let j = g.ng; j < g.H - g.ng; j++) s += g.area[j] * g.nx;
  return s;
}
