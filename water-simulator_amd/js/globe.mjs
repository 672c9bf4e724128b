// The globe harness of the reference (src/globe.mjs) over the drop-in solver:
// measureModePeriod (globe.mjs:1387-1456), polarTimestepTax (:1022-1037) and
// legendreP (:1088-1094), with the reference's signatures and results.  The
// reference module also holds a canvas renderer and does not parse on Node 12;
// these are the solver-facing parts.  measureModePeriod's loop runs
// device-resident (ShallowWater.projectedRun): the steps, and the projection of
// eta onto P_n after every one of them, stay on the GPU; the zero crossings are
// found here from the recorded series, as the reference finds them.
import { ShallowWater } from './swe.mjs';

const D2R = Math.PI / 180;
const R2D = 180 / Math.PI;

/** Legendre P_n(x) for n = 2, 3, 4. */
export function legendreP(n, x) {
  if (n === 2) return 0.5 * (3 * x * x - 1);
  if (n === 3) return 0.5 * (5 * x * x * x - 3 * x);
  if (n === 4) return (35 * x ** 4 - 30 * x * x + 3) / 8;
  throw new Error(`legendreP: n = ${n} is not one of 2, 3, 4 (n = 1 is a rigid translation of the shell, `
    + 'n = 0 a change of volume)');
}

/**
 * The CFL step a uniform depth-H ocean allows in the row nearest the equator
 * and in the row with the smallest step (the last one, capped or not), off the
 * sim's own metric.
 */
export function polarTimestepTax(sim, H = 4000) {
  const g = sim.geom, ng = g.ng === undefined ? sim.ng : g.ng;
  const c = Math.sqrt(9.80665 * H);
  const dtOf = (j) => sim.cfl / (c / g.dxRow[j] + c / g.dyCFL[j]);
  const rows = g.dxRow.length;
  let eqJ = ng, best = Infinity;
  for (let j = ng; j < rows - ng; j++) {
    const a = Math.abs(g.phiC[j]);
    if (a < best) { best = a; eqJ = j; }
  }
  let worstJ = ng;
  for (let j = ng; j < rows - ng; j++) if (dtOf(j) < dtOf(worstJ)) worstJ = j;
  const eq = dtOf(eqJ), pole = dtOf(worstJ);
  return { eq, pole, ratio: eq / pole, poleLat: g.phiC[worstJ] * R2D, eqLat: g.phiC[eqJ] * R2D };
}

/**
 * The period of the zonal mode P_n(sin lat) on a flat aquaplanet against
 * 2 pi R / sqrt(gRef H n(n+1)).  gRef is required: a target built from the
 * solver's own g moves with any error in it.
 */
export function measureModePeriod({
  n = 2, deg = 6, H = 4000, R = 6371e3, gRef, amp = 1, periods = 2, cfl = 0.45,
  lat0 = -90, lat1 = 90, order = 2, onProgress = null,
}) {
  if (!(gRef > 0)) {
    throw new Error('measureModePeriod: gRef is required; a closed form built from the solver\'s own G '
      + 'can never detect an error in it');
  }
  const nx = Math.round(360 / deg), ny = Math.round((lat1 - lat0) / deg);
  const closed = 2 * Math.PI * R / Math.sqrt(gRef * H * n * (n + 1));
  const sim = new ShallowWater({
    nx, ny, bed: () => -H, manning: 0, cfl, order,
    eta0: (lon, lat) => amp * legendreP(n, Math.sin(lat * D2R)),
    sphere: { R, lat0, lat1, omega: 0 },
  });
  // projection weights: cell area x P_n at the row's latitude
  const g = sim.geom, ng = sim.ng;
  const wRow = new Float64Array(ny), pRow = new Float64Array(ny), coef = new Float64Array(ny);
  let norm = 0;
  for (let j = 0; j < ny; j++) {
    const P = legendreP(n, Math.sin(g.phiC[j + ng]));
    pRow[j] = P; wRow[j] = g.area[j + ng] * nx;
    norm += wRow[j] * P * P;
    coef[j] = (wRow[j] / nx) * pRow[j];
  }
  const project = () => {
    let s = 0;
    for (let j = 0; j < ny; j++) {
      let row = 0;
      for (let i = 0; i < nx; i++) row += sim.b[sim.idx(i, j)] + sim.h[sim.idx(i, j)];
      s += coef[j] * row;
    }
    return s / norm;
  };
  const tEnd = periods * closed;
  const tax = polarTimestepTax(sim, H);
  const c0 = project();
  const { times, vals } = sim.projectedRun(coef, tEnd, Math.ceil(1.25 * tEnd / tax.pole) + 64);
  const cross = [];
  let prevT = 0, prevC = c0;
  for (let k = 0; k < vals.length; k++) {
    const c = vals[k] / norm, t = times[k];
    if ((prevC < 0) !== (c < 0)) cross.push(prevT + (t - prevT) * (0 - prevC) / (c - prevC));
    prevT = t; prevC = c;
  }
  if (onProgress) onProgress(1);
  const steps = vals.length;
  if (cross.length < 2) {
    return { measured: NaN, closed, rel: NaN, steps, crossings: cross.length, amp0: c0, ampEnd: prevC,
      note: 'fewer than two zero crossings: nothing to measure' };
  }
  const measured = 2 * (cross[cross.length - 1] - cross[0]) / (cross.length - 1);
  return {
    measured, closed, rel: (measured - closed) / closed, steps, crossings: cross.length,
    amp0: c0, ampEnd: prevC, nx, ny, deg,
    note: `${cross.length} crossings, ${steps} steps, amplitude ${c0.toFixed(4)} -> ${prevC.toFixed(4)} m`,
  };
}
