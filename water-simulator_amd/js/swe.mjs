// Drop-in for src/swe.mjs of Syntaxswine/water-simulator whose time step runs on
// an MI355X through the N-API addon (../napi/swe_napi.node) over the C ABI of
// include/swe_amd.h.  Same exports, same constructor options, same public fields
// and methods, same numbers: the GPU kernels reproduce the reference's step bit
// for bit (Manning friction to the last ulp of Math.pow).
//
// Coherence: callers read and write sim.h, sim.hu, sim.hv and sim.b between
// calls, so every device call uploads the host arrays it needs first and
// downloads what it changed afterwards (strict mode): a step uploads b, h, hu,
// hv and downloads h, hu, hv and the ghost ring of b (the only part of b a step
// changes).  sim.rh, sim.rhu, sim.rhv and sim.theta hold what the reference
// leaves in them (the last residual: stage 2's after a step) and are downloaded
// when first read.  advance(n) runs n steps device-resident for throughput work
// and downloads once at the end.
//
// Boundaries are recognised by identity (reflect, outflow, periodic exported
// here) or by the tag a flather() closure carries; the flather closure's
// etaExt(t)/uExt(t) are evaluated here on the host once per step at sim.t and
// handed to the device, exactly the values the reference's per-ghost-cell calls
// produce (t does not move inside a step).  Any other boundary function runs
// on the host, as the reference runs it: applyBC's loop over the host arrays
// before each stage, the step split around the second one.  Device-resident
// runs take a flather side's values from harmonicFlather()'s description,
// evaluated on the device every step.  A makeSponge() forcing is recognised
// the same way and runs inside the step on the device; any other forcing
// callback runs on the host after the step, where the reference calls it.

import { createRequire } from 'module';
import { cartesianGeometry, sphericalGeometry } from './geometry.mjs';

const require = createRequire(import.meta.url);
const addon = require('../napi/swe_napi.node');

export const G = 9.80665;          // standard gravity [m/s^2]

const F = { B: 0, H: 1, HU: 2, HV: 3, MANNING: 4, RH: 5, RHU: 6, RHV: 7, THETA: 8, SPONGE: 9 };
const S = { MANNING: 0, CORIOLIS: 1, CFL: 2, T: 3, SPONGE_REF: 4, STEPS: 5 };
const ROWS = ['dxRow', 'dyRowN', 'dyRowS', 'dyCFL', 'bedPhi', 'tanPhi', 'fRow', 'geoCoef', 'area'];
const LIMITERS = { minmod: 0, mc: 1, central: 2 };
const BC = { REFLECT: 0, OUTFLOW: 1, PERIODIC: 2, FLATHER: 3, STORED: 4 };
const NO_EXT = new Float64Array(8);
const RESIDUAL = ['rh', 'rhu', 'rhv', 'theta'];
const SIDES = ['west', 'east', 'south', 'north'];
const FLATHER_TAG = Symbol('swe.flather');
const SPONGE_TAG = Symbol('swe.sponge');

export class ShallowWater {
  constructor({
    nx, ny, dx, dy = dx, bed, eta0 = 0, manning = 0.025,
    minDepth = 1e-3, cfl = 0.45, coriolis = 0, order = 2, limiter = 'mc',
    sphere = null,
  }) {
    this.nx = nx; this.ny = ny; this.dx = dx; this.dy = dy;
    this.geom = sphere
      ? sphericalGeometry({ nx, ny, ng: 2, ...sphere })
      : cartesianGeometry({ nx, ny, dx, dy, ng: 2, coriolis });
    if (sphere) { this.dx = undefined; this.dy = this.geom.dy; }
    this.manning = manning; this.minDepth = minDepth; this.cfl = cfl;
    this.eps4 = Math.max(minDepth, 1e-12) ** 4;
    this.coriolis = coriolis; this.order = order;
    this.manningField = null;
    this.limiterName = limiter;
    if (!(limiter in LIMITERS)) throw new Error(`unknown limiter ${limiter}; have ${Object.keys(LIMITERS)}`);
    this.ng = 2;
    this.W = nx + 2 * this.ng;
    this.H = ny + 2 * this.ng;
    const N = this.W * this.H;
    // the public fields are ordinary Float64Arrays over page-locked host memory, so
    // the per-call uploads and downloads of strict mode run at full DMA speed (an
    // array a caller substitutes for one of them works too, at pageable speed)
    const f64 = (n) => addon.allocPinned(n);
    this.b = f64(N);
    this.h = f64(N);
    this.hu = f64(N);
    this.hv = f64(N);
    // rh, rhu, rhv, theta: the last residual's, fetched from the device on first read
    this._resid = { rh: f64(N), rhu: f64(N), rhv: f64(N), theta: f64(N).fill(1) };
    this._residStale = false;
    for (const name of RESIDUAL) {
      Object.defineProperty(this, name, {
        enumerable: true,
        get() { if (this._residStale) this._pullResidual(); return this._resid[name]; },
        set(v) { this._resid[name] = v; },
      });
    }
    for (let j = 0; j < this.H; j++) {
      for (let i = 0; i < this.W; i++) {
        const k = j * this.W + i;
        const [x, y] = this.geom.kind === 'sphere'
          ? this.cellLonLat(i - this.ng, j - this.ng)
          : this.cellCentre(i - this.ng, j - this.ng);
        this.b[k] = bed(x, y);
        const e = typeof eta0 === 'function' ? eta0(x, y) : eta0;
        this.h[k] = Math.max(0, e - this.b[k]);
      }
    }
    this.t = 0;
    this.steps = 0;
    this.massFloored = 0;
    this.boundaries = this.geom.kind === 'sphere'
      ? { west: periodic, east: periodic, south: reflect, north: reflect }
      : { west: reflect, east: reflect, south: reflect, north: reflect };
    this.forcing = null;
    this.stats = {};

    this._dev = addon.create({
      nx, ny, kind: sphere ? 1 : 0, order, limiter: LIMITERS[limiter], g: G, minDepth,
      eps4: this.eps4, cfl, coriolis: sphere ? 0 : coriolis, manning,
      R: this.geom.R, rows: ROWS.map((r) => this.geom[r]),
    });
    this._rowsSent = ROWS.map((r) => Float64Array.from(this.geom[r]));
    this._spongeSent = null;
    this._mfieldSent = null;
    // every device step leaves stage 2's residual and theta on the device
    addon.keepResidual(this._dev, true);
  }

  cellCentre(i, j) {
    if (this.geom.kind === 'sphere') {
      throw new Error('cellCentre is Cartesian-only; a spherical cell has no single '
        + 'x in metres. Use cellLonLat(i, j).');
    }
    return [(i + 0.5) * this.dx, (j + 0.5) * this.dy];
  }

  cellLonLat(i, j) {
    const g = this.geom;
    if (g.kind !== 'sphere') throw new Error('cellLonLat is spherical-only');
    const lon = (i + 0.5) * g.dlam * 180 / Math.PI;
    return [((lon % 360) + 360) % 360 - 180, g.phiC[j + this.ng] * 180 / Math.PI];
  }

  idx(i, j) { return (j + this.ng) * this.W + (i + this.ng); }
  eta(i, j) { const k = this.idx(i, j); return this.b[k] + this.h[k]; }
  depth(i, j) { return this.h[this.idx(i, j)]; }

  /** Desingularised velocity, the same expression the device uses. */
  vel(hq, h) {
    const h4 = h * h * h * h;
    return (Math.SQRT2 * h * hq) / Math.sqrt(h4 + Math.max(h4, this.eps4));
  }

  // ---- host <-> device coherence ------------------------------------------
  _syncParams() {
    const d = this._dev;
    addon.setScalar(d, S.MANNING, this.manning);
    addon.setScalar(d, S.CORIOLIS, this.geom.kind === 'sphere' ? 0 : this.coriolis);
    addon.setScalar(d, S.CFL, this.cfl);
    for (let r = 0; r < ROWS.length; r++) {
      const now = this.geom[ROWS[r]], sent = this._rowsSent[r];
      let same = now.length === sent.length;
      for (let j = 0; same && j < now.length; j++) same = Object.is(now[j], sent[j]);
      if (!same) { addon.setRow(d, r, now); this._rowsSent[r] = Float64Array.from(now); }
    }
    if (this.manningField !== this._mfieldSent) {
      addon.setField(d, F.MANNING, this.manningField);
      this._mfieldSent = this.manningField;
    } else if (this.manningField) {
      addon.setField(d, F.MANNING, this.manningField);
    }
  }

  _upload(fields) { for (const f of fields) addon.setField(this._dev, F[f], this[f.toLowerCase()]); }
  _download(fields) { for (const f of fields) addon.getField(this._dev, F[f], this[f.toLowerCase()]); }
  // after a device step or applyBC: h, hu, hv whole, and of b only the ghost ring
  _downloadState() {
    this._download(['H', 'HU', 'HV']);
    addon.getGhosts(this._dev, F.B, this.b);
  }

  _pullResidual() {
    this._residStale = false;
    for (const name of RESIDUAL) addon.getField(this._dev, F[name.toUpperCase()], this._resid[name]);
  }

  // Boundary kinds and the Flather values at the current t.  Returns null when a
  // side is a function the device does not implement: then every side is filled
  // on the host (_hostApplyBC) and the device keeps the ghosts it is given.
  _boundaries() {
    const ext = new Float64Array(8);
    const kinds = [];
    for (let s = 0; s < 4; s++) {
      const fn = this.boundaries[SIDES[s]];
      let kind = null;
      if (fn === reflect) kind = BC.REFLECT;
      else if (fn === outflow) kind = BC.OUTFLOW;
      else if (fn === periodic) kind = BC.PERIODIC;
      else if (fn && fn[FLATHER_TAG]) kind = BC.FLATHER;
      kinds.push(kind);
    }
    const host = kinds.includes(null);
    for (let s = 0; s < 4; s++) {
      addon.setBc(this._dev, s, host ? BC.STORED : kinds[s]);
      if (!host && kinds[s] === BC.FLATHER) {
        const tag = this.boundaries[SIDES[s]][FLATHER_TAG];
        ext[2 * s] = tag.etaExt(this.t);
        ext[2 * s + 1] = tag.uExt(this.t);
      }
    }
    return host ? null : ext;
  }

  // applyBC's two passes (swe.mjs:289-303) over the host arrays with the
  // boundary functions themselves: the west and east ghost columns over every
  // row, then the south and north ghost rows over every column, corners included
  _hostApplyBC() {
    const { H, ng, nx, ny } = this;
    const bc = this.boundaries;
    for (let j = 0; j < H; j++) {
      for (let g = 0; g < ng; g++) {
        bc.west(this, -1 - g, j - ng, 0, g);
        bc.east(this, nx + g, j - ng, 1, g);
      }
    }
    for (let i = -ng; i < nx + ng; i++) {
      for (let g = 0; g < ng; g++) {
        bc.south(this, i, -1 - g, 2, g);
        bc.north(this, i, ny + g, 3, g);
      }
    }
  }

  // device-resident runs: every side a device kind, a flather side described by harmonicFlather()
  _deviceBoundaries(what) {
    if (!this._boundaries()) {
      throw new Error(`${what} runs device-resident: every boundary must be reflect, outflow, periodic or flather`);
    }
    for (let s = 0; s < 4; s++) {
      const tag = this.boundaries[SIDES[s]][FLATHER_TAG];
      if (!tag) continue;
      const hm = tag.harmonic;
      if (!hm) {
        throw new Error(`${what} runs device-resident: the flather boundary on ${SIDES[s]} needs its forcing `
          + 'described with harmonicFlather() (a flather(etaExt, uExt) closure is evaluated by step() only)');
      }
      const col = (terms, k) => Float64Array.from(terms.map((x) => (k === 'phase' ? (x.phase || 0) : x[k])));
      const uTerms = hm.uScale == null ? hm.u : [];
      addon.setHarmonic(this._dev, s, hm.msl, col(hm.eta, 'amp'), col(hm.eta, 'omega'), col(hm.eta, 'phase'),
        hm.u0, col(uTerms, 'amp'), col(uTerms, 'omega'), col(uTerms, 'phase'));
      addon.setHarmonicRamp(this._dev, s, hm.ramp, hm.uScale != null, hm.uScale == null ? 0 : hm.uScale);
    }
  }

  // a makeSponge() forcing runs on the device: its strengths and reference level
  // go with the step; returns its tag (null when the forcing is none or another)
  _syncSponge() {
    const sp = this.forcing && this.forcing[SPONGE_TAG];
    if (sp) {
      if (this._spongeSent !== sp.strength) { addon.setField(this._dev, F.SPONGE, sp.strength); this._spongeSent = sp.strength; }
      addon.setScalar(this._dev, S.SPONGE_REF, typeof sp.etaRef === 'function' ? sp.etaRef(this.t) : sp.etaRef);
    } else if (this._spongeSent) {
      addon.setField(this._dev, F.SPONGE, null);
      this._spongeSent = null;
    }
    return sp || null;
  }

  applyBC() {
    const ext = this._boundaries();
    if (!ext) { this._hostApplyBC(); return; }
    this._upload(['B', 'H', 'HU', 'HV']);
    addon.applyBc(this._dev, ext);
    for (const f of ['B', 'H', 'HU', 'HV']) addon.getGhosts(this._dev, F[f], this[f.toLowerCase()]);
  }

  residual(dt) {
    this._syncParams();
    this._upload(['B', 'H', 'HU', 'HV']);
    addon.residual(this._dev, dt);
    this._residStale = true;
  }

  maxDt() {
    this._syncParams();
    this._upload(['H', 'HU', 'HV']);
    return addon.maxDt(this._dev);
  }

  applyFriction(dt) {
    this._syncParams();
    this._upload(['H', 'HU', 'HV']);
    addon.applyFriction(this._dev, dt);
    this._download(['HU', 'HV']);
  }

  setManningField(fn) {
    if (!fn) { this.manningField = null; return this; }
    const N = this.W * this.H, out = new Float64Array(N);
    const sph = this.geom.kind === 'sphere';
    for (let j = 0; j < this.H; j++) {
      for (let i = 0; i < this.W; i++) {
        const [x, y] = sph
          ? this.cellLonLat(i - this.ng, j - this.ng)
          : this.cellCentre(i - this.ng, j - this.ng);
        const v = fn(x, y);
        if (!(v >= 0) || !isFinite(v)) {
          throw new Error(`setManningField: n = ${v} at (${x}, ${y}); Manning n must be finite and >= 0`);
        }
        out[j * this.W + i] = v;
      }
    }
    this.manningField = out;
    return this;
  }

  dryClean() {
    this._upload(['H', 'HU', 'HV']);
    const added = addon.dryClean(this._dev);
    this._download(['H', 'HU', 'HV']);
    return added;
  }

  /** Advance by dt (or by the CFL step if dt is omitted). SSP-RK2 on the GPU. */
  step(dt = null) {
    const ext = this._boundaries();
    this._syncParams();
    const sp = this._syncSponge();
    addon.setScalar(this._dev, S.T, this.t);
    const dtArg = dt == null ? NaN : dt;
    let taken, mf;
    if (ext) {
      this._upload(['B', 'H', 'HU', 'HV']);
      [taken, mf] = addon.step(this._dev, dtArg, ext, true, true);
      this._downloadState();
    } else {
      // a boundary function of the caller's: applyBC on the host before each
      // stage (swe.mjs:784,798), the device step split between them
      this._hostApplyBC();
      this._upload(['B', 'H', 'HU', 'HV']);
      taken = addon.stepBegin(this._dev, dtArg, NO_EXT);
      if (taken > 0) {
        this._download(['H', 'HU', 'HV']);
        this._hostApplyBC();
        this._upload(['B', 'H', 'HU', 'HV']);
        mf = addon.stepEnd(this._dev, true, true);
        this._download(['H', 'HU', 'HV']);
      }
    }
    if (!(taken > 0)) return 0;
    this._residStale = true;
    this.massFloored = mf;
    if (this.forcing && !sp) this.forcing(this, taken);
    this.t += taken;
    this.steps++;
    return taken;
  }

  /** n self-timed steps device-resident (no per-step host round trip), then one download. */
  advance(n) {
    this._deviceBoundaries('advance(n)');
    this._syncParams();
    if (this.forcing) throw new Error('advance(n) runs without a host forcing hook');
    addon.setScalar(this._dev, S.T, this.t);
    this._upload(['B', 'H', 'HU', 'HV']);
    const taken = addon.advance(this._dev, n, 0, 0, Infinity);
    this._downloadState();
    if (taken > 0) this._residStale = true;
    this.t = addon.getScalar(this._dev, S.T);
    this.steps += taken;
    return taken;
  }

  /**
   * Equilibrium-tide body force (no counterpart in the reference, which leaves
   * such forcing to `sim.forcing`): constituents [{m, amp, omega, phase}], m = 2
   * semidiurnal (amp cos^2(lat) cos(omega t + 2 lon + phase)), m = 1 diurnal
   * (amp sin(2 lat) cos(omega t + lon + phase)); both RK stages add
   * g h grad(eta_eq) on the device.  Spherical grids only; [] removes it.
   */
  setTide(constituents) {
    const g = this.geom;
    if (g.kind !== 'sphere') throw new Error('setTide: the equilibrium tide needs a spherical grid');
    const lon = new Float64Array(this.W);
    for (let i = 0; i < this.W; i++) lon[i] = (i - this.ng + 0.5) * g.dlam - Math.PI;
    const pick = (k) => Float64Array.from(constituents.map((c) => c[k]));
    addon.setTide(this._dev, Int32Array.from(constituents.map((c) => c.m)), pick('amp'), pick('omega'),
      pick('phase'), Float64Array.from(g.phiC), lon);
    return this;
  }

  /**
   * tools/waves.mjs's recording loop (waves.mjs:211-232) run on the device, with
   * no counterpart in the reference: from sim.t until `until`, each step is
   * min(maxDt, max(1e-6, min(until, next) - t)), and eta = b + h at `cells`
   * (indices into W*H) is sampled after the first step at or past each sample
   * time t0, t0 + dtRec, ...  Returns { times, eta } (eta: samples x cells) and
   * leaves the state as the host loop would.
   */
  recordRun(cells, { t0 = this.t, dtRec, until, capacity = 0 }) {
    this._deviceBoundaries('recordRun');
    this._syncParams();
    if (this.forcing) throw new Error('recordRun runs without a host forcing hook');
    const c = Int32Array.from(cells);
    const cap = capacity > 0 ? capacity : Math.ceil((until - t0) / dtRec) + 2;
    addon.setScalar(this._dev, S.T, this.t);
    this._upload(['B', 'H', 'HU', 'HV']);
    addon.setRecorder(this._dev, c, 0, t0, dtRec, cap);
    try {
      const taken = addon.advance(this._dev, Number.MAX_SAFE_INTEGER, 3, 0, until);
      const n = addon.getRecord(this._dev, null, null);
      const eta = new Float64Array(n * c.length), times = new Float64Array(n);
      addon.getRecord(this._dev, eta, times);
      this._downloadState();
      if (taken > 0) this._residStale = true;
      this.t = addon.getScalar(this._dev, S.T);
      this.steps += taken;
      return { times, eta };
    } finally {
      addon.setRecorder(this._dev, new Int32Array(0), 0, 0, 0, 0);
    }
  }

  /**
   * Run device-resident until `until` (steps min(maxDt, until - t)), recording
   * after every step s = sum_j coef[j] * (sum over row j of b + h), both sums in
   * the reference's order -- the mode projection of src/globe.mjs:1417-1424 --
   * with no counterpart in swe.mjs.  coef: ny values.  Returns { times, vals }.
   */
  projectedRun(coef, until, capacity) {
    this._deviceBoundaries('projectedRun');
    this._syncParams();
    if (this.forcing) throw new Error('projectedRun runs without a host forcing hook');
    addon.setScalar(this._dev, S.T, this.t);
    this._upload(['B', 'H', 'HU', 'HV']);
    addon.setProjection(this._dev, Float64Array.from(coef), capacity);
    try {
      const taken = addon.advance(this._dev, Number.MAX_SAFE_INTEGER, 2, 0, until);
      const n = addon.getProjection(this._dev, null, null);
      const vals = new Float64Array(n), times = new Float64Array(n);
      addon.getProjection(this._dev, vals, times);
      this._downloadState();
      if (taken > 0) this._residStale = true;
      this.t = addon.getScalar(this._dev, S.T);
      this.steps += taken;
      return { times, vals };
    } finally {
      addon.setProjection(this._dev, null, 0);
    }
  }

  /**
   * volume, energy, maxSurfaceDeviation(ref), maxSpeed and finite of the device
   * state, computed on the device (no counterpart in the reference, whose
   * methods below loop over the arrays).  The device state is the arrays'
   * state after step(), advance() and recordRun(); writes to sim.h etc. reach
   * the device with the next step.  Energy, the maxima, finite and the
   * spherical volume equal the host methods bit for bit; the Cartesian volume
   * is the exactly rounded sum (the host method's running sum rounds at each
   * cell).
   */
  deviceDiagnostics(ref = 0) {
    const out = new Float64Array(5);
    addon.diagnostics(this._dev, ref, out);
    return { volume: out[0], energy: out[1], maxSurfaceDeviation: out[2], maxSpeed: out[3], finite: out[4] === 1 };
  }

  // ---- diagnostics (host, over the coherent mirror) ---------------------------
  volume() {
    const { W, ng, nx, ny } = this;
    if (this.geom.kind !== 'sphere') {
      let v = 0;
      for (let j = 0; j < ny; j++) for (let i = 0; i < nx; i++) v += this.h[(j + ng) * W + (i + ng)];
      return v * this.dx * this.dy;
    }
    const area = this.geom.area;
    let v = 0;
    for (let j = 0; j < ny; j++) {
      let row = 0;
      for (let i = 0; i < nx; i++) row += this.h[(j + ng) * W + (i + ng)];
      v += row * area[j + ng];
    }
    return v;
  }

  energy() {
    const { W, ng, nx, ny } = this;
    const sph = this.geom.kind === 'sphere', area = this.geom.area;
    let e = 0;
    for (let j = 0; j < ny; j++) {
      let row = 0;
      for (let i = 0; i < nx; i++) {
        const k = (j + ng) * W + (i + ng), hk = this.h[k];
        if (hk <= 0) continue;
        const u = this.vel(this.hu[k], hk), v = this.vel(this.hv[k], hk);
        row += 0.5 * hk * (u * u + v * v) + 0.5 * G * hk * (hk + 2 * this.b[k]);
      }
      if (sph) e += row * area[j + ng]; else e += row;
    }
    return sph ? e : e * this.dx * this.dy;
  }

  maxSurfaceDeviation(ref) {
    const { W, ng, nx, ny } = this;
    let m = 0;
    for (let j = 0; j < ny; j++) {
      for (let i = 0; i < nx; i++) {
        const k = (j + ng) * W + (i + ng);
        if (this.h[k] <= this.minDepth) continue;
        const d = Math.abs(this.b[k] + this.h[k] - ref);
        if (d > m) m = d;
      }
    }
    return m;
  }

  maxSpeed() {
    const { W, ng, nx, ny } = this;
    let m = 0;
    for (let j = 0; j < ny; j++) {
      for (let i = 0; i < nx; i++) {
        const k = (j + ng) * W + (i + ng), hk = this.h[k];
        if (hk <= this.minDepth) continue;
        const s = Math.hypot(this.vel(this.hu[k], hk), this.vel(this.hv[k], hk));
        if (s > m) m = s;
      }
    }
    return m;
  }

  finite() {
    for (let k = 0; k < this.h.length; k++) {
      if (!Number.isFinite(this.h[k]) || !Number.isFinite(this.hu[k]) || !Number.isFinite(this.hv[k])) return false;
    }
    return true;
  }
}

/** HLLC interface flux with Einfeldt speeds and dry states: the kernels' own code. */
/**
 * Discrete adjoint of `nsteps` self-timed steps of `sim` (no counterpart in the
 * reference): the forward/backward pair of a differentiable function, the JS
 * analogue of a torch.autograd.Function.
 *
 *   const adj = new Adjoint(sim, 4000);
 *   adj.setGauges(cells, 10, obs);        // obs: (nsteps/10) x cells.length
 *   const J = adj.forward(h0, nField);    // runs the steps, records dt, checkpoints
 *   const { dh0, dn } = adj.backward();   // dJ/dh0, dJ/dn over W*H cells
 *
 * J = sum over samples and gauges of (eta - obs)^2.  The dt sequence is the
 * forward's (held fixed); branches take the subgradient the forward took.
 * Boundaries reflect/outflow/periodic, and flather described by
 * harmonicFlather() (the forward runs from t = 0); a makeSponge() forcing is
 * differentiated too.  The adjoint drives the solver's device state, so
 * `sim.h` etc. are re-uploaded by its next call.
 */
export class Adjoint {
  // checkpointEvery: steps per replayed segment (0: chosen by the library);
  // outerEvery: steps between the checkpoints kept for the whole run, a multiple of
  // checkpointEvery (0: one level)
  constructor(sim, nsteps, { checkpointEvery = 0, outerEvery = 0 } = {}) {
    this.sim = sim;
    this.nsteps = nsteps;
    this.ng = 0;
    this.every = 1;
    this._prepare();
    this._adj = addon.adjCreate(sim._dev, nsteps, checkpointEvery, outerEvery);
  }

  _prepare() {
    const sim = this.sim;
    sim._deviceBoundaries('Adjoint');
    if (sim.forcing && !sim._syncSponge()) throw new Error('Adjoint: the only forcing it differentiates is makeSponge()');
    sim._syncParams();
    sim._upload(['B']);
  }

  setGauges(cells, every, obs) {
    this.ng = cells.length;
    this.every = every;
    addon.adjSetGauges(this._adj, Int32Array.from(cells), every, Float64Array.from(obs));
    return this;
  }

  forward(h0, manningField = null) {
    this._prepare();
    return addon.adjForward(this._adj, Float64Array.from(h0), manningField ? Float64Array.from(manningField) : null);
  }

  /** J of the forward that replays the recorded dt sequence (for finite-difference checks). */
  loss(h0, manningField = null) {
    return addon.adjLoss(this._adj, Float64Array.from(h0), manningField ? Float64Array.from(manningField) : null);
  }

  backward() {
    const N = this.sim.W * this.sim.H;
    const dh0 = new Float64Array(N), dn = new Float64Array(N);
    addon.adjBackward(this._adj, dh0, dn);
    return { dh0, dn };
  }

  series() {
    const out = new Float64Array(Math.floor(this.nsteps / this.every) * this.ng);
    addon.adjSeries(this._adj, out);
    return out;
  }

  timing() {
    const [forwardMs, backwardMs] = addon.adjTiming(this._adj);
    return { forwardMs, backwardMs };
  }
}

export function hllc(out, hL, huL, hvL, hR, huR, hvR, dry = 1e-4) {
  const f = addon.hllc(hL, huL, hvL, hR, huR, hvR, dry, G);
  out[0] = f[0]; out[1] = f[1]; out[2] = f[2];
  return out;
}

// ---- boundary conditions: host bodies for direct calls, device kinds by identity
function clampIdx(v, hi) { return v < 0 ? 0 : v > hi ? hi : v; }

export function reflect(sim, i, j, side) {
  let si = i, sj = j;
  if (side === 0) si = -1 - i; else if (side === 1) si = 2 * sim.nx - 1 - i;
  else if (side === 2) sj = -1 - j; else sj = 2 * sim.ny - 1 - j;
  si = clampIdx(si, sim.nx - 1); sj = clampIdx(sj, sim.ny - 1);
  const g = sim.idx(i, j), s = sim.idx(si, sj);
  sim.b[g] = sim.b[s]; sim.h[g] = sim.h[s];
  const flipX = side === 0 || side === 1;
  sim.hu[g] = flipX ? -sim.hu[s] : sim.hu[s];
  sim.hv[g] = flipX ? sim.hv[s] : -sim.hv[s];
}

export function outflow(sim, i, j) {
  const g = sim.idx(i, j), s = sim.idx(clampIdx(i, sim.nx - 1), clampIdx(j, sim.ny - 1));
  sim.b[g] = sim.b[s]; sim.h[g] = sim.h[s]; sim.hu[g] = sim.hu[s]; sim.hv[g] = sim.hv[s];
}

export function periodic(sim, i, j, side) {
  let si = i, sj = j;
  if (side === 0) si = i + sim.nx; else if (side === 1) si = i - sim.nx;
  else if (side === 2) sj = j + sim.ny; else sj = j - sim.ny;
  const g = sim.idx(i, j), s = sim.idx(si, sj);
  sim.b[g] = sim.b[s]; sim.h[g] = sim.h[s]; sim.hu[g] = sim.hu[s]; sim.hv[g] = sim.hv[s];
}

/** Flather radiation boundary; the device evaluates it from the tagged etaExt/uExt. */
export function flather(etaExt, uExt = () => 0) {
  const fn = (sim, i, j, side) => {
    const g = sim.idx(i, j), s = sim.idx(clampIdx(i, sim.nx - 1), clampIdx(j, sim.ny - 1));
    sim.b[g] = sim.b[s];
    const bed = sim.b[g], hInt = sim.h[s];
    const e0 = etaExt(sim.t), u0 = uExt(sim.t);
    if (hInt <= sim.minDepth) { sim.h[g] = Math.max(0, e0 - bed); sim.hu[g] = 0; sim.hv[g] = 0; return; }
    const etaInt = bed + hInt;
    const r = Math.sqrt(G / hInt);
    const sgn = (side === 0 || side === 2) ? 1 : -1;
    const uInt = (side === 0 || side === 1) ? sim.vel(sim.hu[s], hInt) : sim.vel(sim.hv[s], hInt);
    const Rin = u0 + sgn * e0 * r;
    const Rout = uInt - sgn * etaInt * r;
    const un = 0.5 * (Rin + Rout);
    const en = sgn * (Rin - Rout) / (2 * r);
    sim.h[g] = Math.max(0, en - bed);
    if (side === 0 || side === 1) { sim.hu[g] = sim.h[g] * un; sim.hv[g] = sim.hv[s]; }
    else { sim.hv[g] = sim.h[g] * un; sim.hu[g] = sim.hu[s]; }
  };
  fn[FLATHER_TAG] = { etaExt, uExt };
  return fn;
}

/**
 * A flather() boundary whose far field is a sum of at most four cosines (no
 * counterpart in src/swe.mjs; Tide.eta of src/tide.mjs:103-107 and regularWave
 * of src/waves.mjs:163-175 are such sums):
 *   etaExt(t) = msl + sum_k a_k cos(omega_k t - phase_k),
 *   uExt(t)   = u0 + the same over `u`, or uScale * (etaExt(t) - msl),
 * with a_k = amp_k, or amp_k * min(1, t / ramp) when ramp > 0.  step()
 * evaluates the closures on the host like any flather(); advance(),
 * recordRun() and projectedRun() evaluate the same sums on the device at each
 * step's t (the device cos may differ from Math.cos in the last bit).
 */
export function harmonicFlather({ msl = 0, eta = [], u0 = 0, u = [], ramp = 0, uScale = null } = {}) {
  if (eta.length > 4 || u.length > 4) throw new Error('harmonicFlather: at most 4 terms each for eta and u');
  const sum = (c0, terms, t) => {
    let e = c0;
    const m = ramp > 0 ? Math.min(1, t / ramp) : 1;
    for (const x of terms) e += (ramp > 0 ? m * x.amp : x.amp) * Math.cos(x.omega * t - (x.phase || 0));
    return e;
  };
  const etaExt = (t) => sum(msl, eta, t);
  const uExt = uScale == null ? (t) => sum(u0, u, t) : (t) => uScale * (etaExt(t) - msl);
  const fn = flather(etaExt, uExt);
  fn[FLATHER_TAG].harmonic = { msl, eta, u0, u, ramp, uScale };
  return fn;
}

/** Sponge layer forcing: relax toward max(0, etaRef - b) over a cos^2-tapered band. */
export function makeSponge(sim, { side = 'east', width = 20, strength = 0.06, etaRef = 0 }) {
  const { nx, ny, W, ng } = sim;
  const str = new Float64Array(W * sim.H);
  const cells = [];
  for (let j = 0; j < ny; j++) {
    for (let i = 0; i < nx; i++) {
      let d;
      if (side === 'east') d = nx - 1 - i; else if (side === 'west') d = i;
      else if (side === 'north') d = ny - 1 - j; else d = j;
      if (d >= width) continue;
      const s = strength * Math.cos(0.5 * Math.PI * d / width) ** 2;
      const k = (j + ng) * W + (i + ng);
      str[k] = s;
      cells.push([k, s]);
    }
  }
  const fn = (s2, dt) => {
    const ref = typeof etaRef === 'function' ? etaRef(s2.t) : etaRef;
    for (const [k, s] of cells) {
      const f = Math.min(1, s * dt * 60);
      const target = Math.max(0, ref - s2.b[k]);
      s2.h[k] += (target - s2.h[k]) * f;
      s2.hu[k] -= s2.hu[k] * f;
      s2.hv[k] -= s2.hv[k] * f;
    }
  };
  fn[SPONGE_TAG] = { strength: str, etaRef };
  return fn;
}
