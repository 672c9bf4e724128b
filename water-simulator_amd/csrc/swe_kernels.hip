// HIP kernels (gfx950) of the shallow-water time step of src/swe.mjs.
//
// k_apply_bc runs before each stage and leaves the ghost ring of the stage's
// input exactly as the reference's applyBC() does; the sweeps then read ghost
// cells like any other cell.
//
// k_sweep: one fused stage of SSP-RK2 (swe.mjs:754-814 stage bodies, with
//   residual() :308-629, dryClean() :734-745, and in stage 2 applyFriction()
//   :659-704, the sponge forcing :1188-1197 and the next step's maxDt()
//   :632-647).  One wavefront owns a strip of 64 columns (58 outputs + 3 halo
//   columns each side) and marches down a chunk of rows.  Along y every value a
//   cell needs stays in that lane's registers from one row to the next; along x
//   the lanes trade neighbour values with cross-lane shuffles.  Nothing but the
//   state is read from or written to HBM: slopes, face states, HLLC fluxes,
//   positivity factors and residuals never leave registers.  Each row's
//   (b, h, hu, hv) is read from memory once: a four-row ring in LDS hands it
//   to the cell update three rows later.
//
// Row pipeline of iteration r (row r is loaded):
//   slopes + face states at s=r-1, y-face (s-1|s), x-faces at s,
//   theta at q=r-2, x-face contributions at q, y-face (q-1|q) contributions,
//   and the cell update of row o=r-3.
#include "swe_device.hpp"
#include "swe_kernels.hpp"

namespace swe {

struct Cell { double b, h, hu, hv; };

__device__ __forceinline__ Cell ld(const StateP &S, const double *b, int k) {
  Cell c; c.b = b[k]; c.h = S.h[k]; c.hu = S.hu[k]; c.hv = S.hv[k]; return c;
}

__device__ __forceinline__ int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Source cell (interior coordinates) of ghost (i,j) on `side` for BC `kind`:
// reflect swe.mjs:1026-1034, outflow :1042-1045, flather :1086-1089, periodic :1119-1124.
__device__ __forceinline__ void bc_src(int kind, int side, int i, int j, int nx, int ny, int &si, int &sj) {
  if (kind == BC_REFLECT) {
    si = i; sj = j;
    if (side == 0) si = -1 - i; else if (side == 1) si = 2 * nx - 1 - i;
    else if (side == 2) sj = -1 - j; else sj = 2 * ny - 1 - j;
    si = iclamp(si, 0, nx - 1); sj = iclamp(sj, 0, ny - 1);
  } else if (kind == BC_PERIODIC) {
    // i+nx / i-nx (j likewise); when nx or ny is 1 the reference's second ghost
    // layer reads the first one it has just written, which is the same as
    // wrapping the index modulo the interior size
    si = i; sj = j;
    if (side < 2) si = ((i % nx) + nx) % nx;
    else sj = ((j % ny) + ny) % ny;
  } else {
    si = iclamp(i, 0, nx - 1); sj = iclamp(j, 0, ny - 1);
  }
}

// Ghost value from its source cell (the bodies of reflect/outflow/periodic/flather).
__device__ __forceinline__ Cell bc_apply(int kind, int side, const Cell &s, const Params &P, const double *ext) {
  Cell g;
  if (kind == BC_REFLECT) {
    const bool flipX = side < 2;
    g.b = s.b; g.h = s.h;
    g.hu = flipX ? -s.hu : s.hu;
    g.hv = flipX ? s.hv : -s.hv;
  } else if (kind == BC_FLATHER) {                     // swe.mjs:1084-1116
    g.b = s.b;
    const double bed = s.b, hInt = s.h;
    const double e0 = ext[2 * side], u0 = ext[2 * side + 1];
    if (hInt <= P.min_depth) { g.h = jmax0(e0 - bed); g.hu = 0.0; g.hv = 0.0; return g; }
    const double etaInt = bed + hInt;
    const double r = __builtin_sqrt(P.g / hInt);
    const double sgn = (side == 0 || side == 2) ? 1.0 : -1.0;
    const double uInt = side < 2 ? vel1(s.hu, hInt, P.eps4) : vel1(s.hv, hInt, P.eps4);
    const double Rin = u0 + sgn * e0 * r;
    const double Rout = uInt - sgn * etaInt * r;
    const double un = 0.5 * (Rin + Rout);
    const double en = sgn * (Rin - Rout) / (2 * r);
    g.h = jmax0(en - bed);
    if (side < 2) { g.hu = g.h * un; g.hv = s.hv; }
    else { g.hv = g.h * un; g.hu = s.hu; }
  } else {
    g = s;
  }
  return g;
}

// Ghost of a W/E column at an interior row, as applyBC's first pass leaves it.
__device__ __forceinline__ Cell ghost_we(const Params &P, const StateP &S, const double *b, const double *ext,
                                      int gi, int gj) {
  const int side = gi < 2 ? 0 : 1;
  const int kind = P.bc[side];
  if (kind == BC_STORED) return ld(S, b, gj * P.W + gi);
  int si, sj;
  bc_src(kind, side, gi - 2, gj - 2, P.nx, P.ny, si, sj);
  const Cell s = ld(S, b, (sj + 2) * P.W + (si + 2));
  return bc_apply(kind, side, s, P, ext);
}

// Any cell after applyBC: interior as stored, W/E ghosts from the first pass,
// S/N ghost rows (corners included) from the second pass, whose periodic source
// may itself be a W/E ghost.
__device__ __forceinline__ Cell ghost_cell(const Params &P, const StateP &S, const double *b, const double *ext,
                                        int gi, int gj) {
  if (gj >= 2 && gj < P.H - 2) return ghost_we(P, S, b, ext, gi, gj);
  const int side = gj < 2 ? 2 : 3;
  const int kind = P.bc[side];
  if (kind == BC_STORED) return ld(S, b, gj * P.W + gi);
  int si, sj;
  bc_src(kind, side, gi - 2, gj - 2, P.nx, P.ny, si, sj);
  const int sgi = si + 2, sgj = sj + 2;
  Cell s;
  if (sgi >= 2 && sgi < P.W - 2) s = ld(S, b, sgj * P.W + sgi);
  else s = ghost_we(P, S, b, ext, sgi, sgj);
  return bc_apply(kind, side, s, P, ext);
}

struct Row { double h, b, eta, U, V; };          // what the slope stencil needs of a cell
struct Face4 { double eta, b, u, v; };           // reconstructed state on one side of a face
struct FluxF { double M, NL, NR, T; };           // face flux: mass, normal+pL, normal+pR, transverse
struct Part3 { double h, hu, hv; };

// Whole-wavefront shift by one lane through DPP (wave_shl:1 = 0x130 brings lane
// i+1 to lane i, wave_shr:1 = 0x138 brings lane i-1): two v_mov_b32_dpp per
// double, no LDS round trip.  The end lane, which has no source, reads zero
// (bound_ctrl): it is a halo lane whose values reach no output lane.
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = (int)(unsigned)u, hi = (int)(unsigned)(u >> 32);
  const int rlo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xf, 0xf, true);
  const int rhi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xf, 0xf, true);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)rhi << 32) | (unsigned)rlo);
}
__device__ __forceinline__ double shdn(double v) { return dpp_shift<0x130>(v); }
__device__ __forceinline__ double shup(double v) { return dpp_shift<0x138>(v); }

// Interface flux + hydrostatic pressure corrections (swe.mjs:413-445 body, and
// :448-475 with normal/tangential swapped by the caller).
__device__ __forceinline__ FluxF face_flux(const Face4 &L, const Face4 &R, const HllcConst &C, double half_g) {
  const double hLf = jmax0(L.eta - L.b);
  const double hRf = jmax0(R.eta - R.b);
  // Math.max(L.b, R.b): where the beds are equal zeros of either sign any of
  // them serves, since bStar is only subtracted from a surface and the
  // difference taken through Math.max(0, .), which maps -0 to +0
  const double bStar = (L.b != L.b || R.b != R.b) ? __builtin_nan("") : __builtin_fmax(L.b, R.b);
  const double hsL = jmax0(L.eta - bStar);
  const double hsR = jmax0(R.eta - bStar);
  const Flux3 f = hllc<true>(hsL, hsL * L.u, hsL * L.v, hsR, hsR * R.u, hsR * R.v, C);
  const double pL = half_g * (hLf * hLf - hsL * hsL);
  const double pR = half_g * (hRf * hRf - hsR * hsR);
  FluxF o;
  o.M = f.m; o.NL = f.n + pL; o.NR = f.n + pR; o.T = f.t;
  return o;
}

// One wavefront's share of a stage: strip `strip` (58 output columns) and rows
// [chunk*rows_per_chunk, +rows_per_chunk) of the basin described by A.  wave_id
// indexes the massFloored partial this wavefront owns.
// REC: compile in the optional recording for the adjoint replay (A.recf);
// LIM: the slope limiter (Params::limiter), fixed at compile time;
// UNI: every row's metric values are the same (Params::geo_uni), read once.
// LEAN (stage 2 of a plain launch): no per-cell Manning field, no sponge, and
// neither theta nor the residual written out, so none of them is compiled in.
// sweep_ring: the wavefront's LDS (one per workgroup of one wavefront, however
// many bodies a kernel instantiates): the last four rows' (b, h, hu, hv) of its
// columns, which the cell update of row r-3 reads instead of memory
__shared__ double sweep_ring[4][4][SWEEP_LANES];
// sweep_tide: the wavefront's columns' equilibrium-tide factors of the stage,
// per constituent (ts cl + tc sl, tc cl - ts sl): functions of the column and
// the stage time only, formed once before the rows are marched (per-row-metric
// bodies only; a grid with uniform rows has no tide)
__shared__ double sweep_tide[MAX_TIDE][2][SWEEP_LANES];
template <int MODE, bool REC, int LIM, bool UNI = false, bool LEAN = false>
__device__ __forceinline__ void sweep_body(const SweepArgs &A, int strip, int chunk, int wave_id) {
  double (&ring)[4][4][SWEEP_LANES] = sweep_ring;
  const Params &P = A.P;
  const int W = P.W, H = P.H;
  const int lane = threadIdx.x & (SWEEP_LANES - 1);
  const int gi = strip * SWEEP_OUT - SWEEP_HALO + lane;
  const bool colIn = gi >= 0 && gi < W;
  const bool outLane = colIn && lane >= SWEEP_HALO && lane < SWEEP_LANES - SWEEP_HALO;
  const int y0 = chunk * A.rows_per_chunk;
  const int y1 = min(H, y0 + A.rows_per_chunk);
  const Ctl *ctl = A.ctl;
  const double dt = ctl->dt;
  const int fs = P.fstride;
  const double *inp = A.in.h;          // in.hu = inp + fs, in.hv = inp + 2 fs
  const double *h0p = A.h0.h;
  double *outp = A.out.h;
  if (MODE != MODE_PROBE && (!(dt > 0) || __builtin_isinf(dt))) {
    // a refused step (swe.mjs:784) leaves the state as it was: stage 2 hands the
    // step-start state through to its output buffer unchanged
    if (MODE == MODE_STAGE2 && colIn && outLane) {
      for (int o = y0; o < y1; ++o) {
        const int k = o * W + gi;
        outp[k] = h0p[k]; outp[k + fs] = h0p[k + fs]; outp[k + 2 * fs] = h0p[k + 2 * fs];
      }
    }
    if (lane == 0) A.mf_part[wave_id] = 0.0;
    return;
  }
  const bool posOn = dt > 0;
  const double *tcs = &ctl->tide_cs[(MODE == MODE_STAGE2 || (MODE == MODE_PROBE && P.probe_stage2)) ? 1 : 0][0][0];
  const double *b = A.b;
  const HllcConst HC = {P.g, P.half_g, P.rg, P.min_depth, P.g * P.min_depth >= SQRT_N_MIN};
  const double md = P.min_depth;
  const bool o2 = P.order >= 2;
  // a sphere's rows differ (set_uniform_rows), so UNI instances are Cartesian: no
  // geometric or spherical Coriolis terms, no equilibrium tide
  const bool sph = !UNI && P.sphere;
  const bool bedOn = o2 || sph;
  // the per-row values, interleaved per row (Params::geoA).  They are
  // wave-uniform and unchanged while a stage runs, so they are read through the
  // constant address space: scalar loads into SGPRs, waited on by their own
  // counter, instead of vector loads whose in-order completion holds up the
  // wave's other memory traffic.
  typedef const double __attribute__((address_space(4))) const_double;
  const const_double *RA = (const const_double *)P.geoA;
  // (with UNI every row equals row 2, whose values are loaded once before the row loop)
  auto rowv = [&](int which, int j) -> double { return RA[(UNI ? 2 : j) * GEO_AOS + which]; };
  enum { DX, RDX, DYN, RDYN, DYS, RDYS, DYCFL, RDYCFL, BED, RBED, TAN, FROW, GEOC };

  auto load_row = [&](int r) -> Row {
    Row c;
    if (r >= 0 && r < H && colIn) {
      const int k = r * W + gi;
      const double xh = inp[k], xb = b[k], xhu = inp[k + fs], xhv = inp[k + 2 * fs];
      ring[r & 3][0][lane] = xb; ring[r & 3][1][lane] = xh;
      ring[r & 3][2][lane] = xhu; ring[r & 3][3][lane] = xhv;
      c.h = xh; c.b = xb; c.eta = xb + xh;
      const Vel2 uv = vel2(xhu, xhv, xh, P.eps4);
      c.U = uv.u; c.V = uv.v;
    } else {
      c.h = 0.0; c.b = 0.0; c.eta = 0.0; c.U = 0.0; c.V = 0.0;
    }
    return c;
  };

  if constexpr (!UNI) {
    for (int c = 0; c < P.tide_n; ++c) {
      double X = 0.0, Y = 0.0;
      if (colIn) {
        const double *T = P.tide + (size_t)c * (2 * H + 2 * W);
        const double cl = T[2 * H + gi], sl = T[2 * H + W + gi];
        const double tc = tcs[2 * c], ts = tcs[2 * c + 1];
        X = ts * cl + tc * sl;
        Y = tc * cl - ts * sl;
      }
      sweep_tide[c][0][lane] = X;
      sweep_tide[c][1][lane] = Y;
    }
  }

  Row cM = load_row(y0 - 3);
  Row c0 = load_row(y0 - 2);
  Face4 nPrev = {0.0, 0.0, 0.0, 0.0};
  FluxF fxPrev = {0.0, 0.0, 0.0, 0.0}, fyPrev = {0.0, 0.0, 0.0, 0.0};
  double thPrev = 1.0;
  Part3 xpPrev = {0.0, 0.0, 0.0}, ysPrev = {0.0, 0.0, 0.0};
  Part3 bedPrev = {0.0, 0.0, 0.0}, bedPrev2 = {0.0, 0.0, 0.0};
  double Uo = 0.0;
  double mf = 0.0;
  unsigned long long cflb = 0ull;
  bool bad = false;

  const bool xin = gi >= 2 && gi <= W - 3;             // interior column
  const bool slopeCol = gi >= 1 && gi <= W - 2;
  const bool fxCol = gi >= 1 && gi <= W - 3;           // owns x-face (gi|gi+1)
  const bool fxColW = gi >= 2 && gi <= W - 2;          // x-face (gi-1|gi) exists
  const bool fyCol = gi >= 2 && gi <= W - 3;

  for (int r = y0 - 1; r <= y1 + 2; ++r) {
    const Row cP = load_row(r);
    const int s = r - 1;
    // ---- slopes at row s (swe.mjs:364-411) ---------------------------------
    const double hW = shup(c0.h), hE = shdn(c0.h);
    const double eW = shup(c0.eta), eE = shdn(c0.eta);
    const double bW = shup(c0.b), bE = shdn(c0.b);
    const double UW = shup(c0.U), UE = shdn(c0.U);
    const double VW = shup(c0.V), VE = shdn(c0.V);
    double sEx = 0.0, sBx = 0.0, sUx = 0.0, sVx = 0.0, sEy = 0.0, sBy = 0.0, sUy = 0.0, sVy = 0.0;
    if (o2 && slopeCol && s >= 1 && s <= H - 2) {
      if (!(c0.h <= md || hE <= md || hW <= md)) {
        sEx = limit(LIM, c0.eta - eW, eE - c0.eta);
        sBx = limit(LIM, c0.b - bW, bE - c0.b);
        sUx = limit(LIM, c0.U - UW, UE - c0.U);
        sVx = limit(LIM, c0.V - VW, VE - c0.V);
      }
      if (!(c0.h <= md || cP.h <= md || cM.h <= md)) {
        sEy = limit(LIM, c0.eta - cM.eta, cP.eta - c0.eta);
        sBy = limit(LIM, c0.b - cM.b, cP.b - c0.b);
        sUy = limit(LIM, c0.U - cM.U, cP.U - c0.U);
        sVy = limit(LIM, c0.V - cM.V, cP.V - c0.V);
      }
    }
    if (REC && A.recf && outLane && s >= y0 && s < y1) {
      // the cell fields of (gi, s), recorded for the adjoint replay
      double *R = A.recf;
      const int k = s * W + gi;
      R[k] = c0.U; R[k + fs] = c0.V; R[k + 2 * fs] = c0.eta;
      R[k + 3 * fs] = sEx; R[k + 4 * fs] = sEy; R[k + 5 * fs] = sUx; R[k + 6 * fs] = sUy;
      R[k + 7 * fs] = sVx; R[k + 8 * fs] = sVy; R[k + 9 * fs] = sBx; R[k + 10 * fs] = sBy;
    }
    // reconstructed face states of cell (gi,s) (swe.mjs:418-427)
    const Face4 fE = {c0.eta + 0.5 * sEx, c0.b + 0.5 * sBx, c0.U + 0.5 * sUx, c0.V + 0.5 * sVx};
    const Face4 fWst = {c0.eta - 0.5 * sEx, c0.b - 0.5 * sBx, c0.U - 0.5 * sUx, c0.V - 0.5 * sVx};
    const Face4 fN = {c0.eta + 0.5 * sEy, c0.b + 0.5 * sBy, c0.U + 0.5 * sUy, c0.V + 0.5 * sVy};
    const Face4 fS = {c0.eta - 0.5 * sEy, c0.b - 0.5 * sBy, c0.U - 0.5 * sUy, c0.V - 0.5 * sVy};
    // centred bed term and spherical geometric source of cell (gi,s) (swe.mjs:570-598)
    Part3 bedCur = {0.0, 0.0, 0.0};
    if (bedOn && xin && s >= 2 && s <= H - 3) {
      const double hPx = jmax0(fE.eta - fE.b), hMx = jmax0(fWst.eta - fWst.b);
      const double hPy = jmax0(fN.eta - fN.b), hMy = jmax0(fS.eta - fS.b);
      bedCur.hu = div_r(-P.half_g * (hPx + hMx) * sBx, rowv(DX, s), rowv(RDX, s));
      bedCur.h = div_r(-P.half_g * (hPy + hMy) * sBy, rowv(BED, s), rowv(RBED, s));   // rhv bed term
      bedCur.hv = sph ? rowv(GEOC, s) * P.g * (hPy * hPy + hMy * hMy) : 0.0;        // rhv geometric term
    }
    // ---- y-face (s-1|s): v normal, u tangential (swe.mjs:448-475) ----------
    FluxF fyCur = {0.0, 0.0, 0.0, 0.0};
    if (fyCol && s - 1 >= 1 && s - 1 <= H - 3) {
      const Face4 Ly = {nPrev.eta, nPrev.b, nPrev.v, nPrev.u};
      const Face4 Ry = {fS.eta, fS.b, fS.v, fS.u};
      fyCur = face_flux(Ly, Ry, HC, P.half_g);
    }
    // ---- x-face (gi|gi+1) at row s -----------------------------------------
    const Face4 fWR = {shdn(fWst.eta), shdn(fWst.b), shdn(fWst.u), shdn(fWst.v)};
    FluxF fxCur = {0.0, 0.0, 0.0, 0.0};
    if (fxCol && s >= 2 && s <= H - 3) fxCur = face_flux(fE, fWR, HC, P.half_g);

    // ---- positivity factor at row q = s-1 (swe.mjs:507-529) -----------------
    const int q = s - 1;
    const double hq = cM.h;
    double th = 1.0, out = 0.0;
    {
      const double eM = fxPrev.M, wM = shup(fxPrev.M);
      if (posOn && slopeCol && q >= 1 && q <= H - 2) {
        if (eM > 0) out += div_r(eM, rowv(DX, q), rowv(RDX, q));
        if (wM < 0) out -= div_r(wM, rowv(DX, q), rowv(RDX, q));
        if (fyCur.M > 0) out += div_r(fyCur.M, rowv(DYN, q), rowv(RDYN, q));
        if (fyPrev.M < 0) out -= div_r(fyPrev.M, rowv(DYS, q), rowv(RDYS, q));
        if (out * dt > hq) th = hq > 0 ? hq / (out * dt) : 0.0;
      }
    }
    if (REC && A.recf && outLane) {
      // each value once, by the chunk that owns its row: the x-face (gi|gi+1) of
      // row s, the y-face (s-1|s) of its owner row s-1, theta and the outflow sum
      // of row q (the cell fields of row s went out with the slopes)
      double *R = A.recf;
      if (s >= y0 && s < y1) {
        const int k = s * W + gi;
        R[k + 11 * fs] = fxCur.M; R[k + 12 * fs] = fxCur.NL; R[k + 13 * fs] = fxCur.NR; R[k + 14 * fs] = fxCur.T;
      }
      if (s - 1 >= y0 && s - 1 < y1) {
        const int k = (s - 1) * W + gi;
        R[k + 15 * fs] = fyCur.M; R[k + 16 * fs] = fyCur.NL; R[k + 17 * fs] = fyCur.NR; R[k + 18 * fs] = fyCur.T;
      }
      if (q >= y0 && q < y1) {
        const int k = q * W + gi;
        R[k + 19 * fs] = th; R[k + 20 * fs] = out;
      }
    }
    // ---- x-face contributions at row q (swe.mjs:530-540) ---------------------
    Part3 xpCur;
    {
      const double thE = shdn(th);
      const double the = fxPrev.M >= 0 ? th : thE;
      double inv = 0.0;
      if (q >= 0 && q < H) inv = (the == 1.0) ? rowv(RDX, q) : div_r(the, rowv(DX, q), rowv(RDX, q));
      const double eMc = fxPrev.M * inv, eNL = fxPrev.NL * inv, eNR = fxPrev.NR * inv, eT = fxPrev.T * inv;
      const double wMc = shup(eMc), wNR = shup(eNR), wT = shup(eT);
      const bool rowX = q >= 2 && q <= H - 3;
      double ah = 0.0, ahu = 0.0, ahv = 0.0;
      if (rowX && fxColW) { ah += wMc; ahu += wNR; ahv += wT; }
      if (rowX && fxCol) { ah -= eMc; ahu -= eNL; ahv -= eT; }
      xpCur.h = ah; xpCur.hu = ahu; xpCur.hv = ahv;
    }
    // ---- y-face (q-1|q) contributions (swe.mjs:541-555) ----------------------
    Part3 ysCur = {0.0, 0.0, 0.0}, yn = {0.0, 0.0, 0.0};
    const bool fyPrevOn = fyCol && q - 1 >= 1 && q - 1 <= H - 3;
    if (fyPrevOn) {
      const double thy = fyPrev.M >= 0 ? thPrev : th;
      const double invL = (thy == 1.0) ? rowv(RDYN, q - 1) : div_r(thy, rowv(DYN, q - 1), rowv(RDYN, q - 1));
      const double invR = (thy == 1.0) ? rowv(RDYS, q) : div_r(thy, rowv(DYS, q), rowv(RDYS, q));
      yn.h = fyPrev.M * invL; yn.hu = fyPrev.T * invL; yn.hv = fyPrev.NL * invL;
      ysCur.h = fyPrev.M * invR; ysCur.hu = fyPrev.T * invR; ysCur.hv = fyPrev.NR * invR;
    }

    // ---- cell update of row o = q-1 ------------------------------------------
    const int o = q - 1;
    if (outLane && o >= y0 && o < y1) {
      const int k = o * W + gi;
      const bool fySouth = fyCol && o - 1 >= 1 && o - 1 <= H - 3;
      double rh = xpPrev.h, rhu = xpPrev.hu, rhv = xpPrev.hv;
      if (fySouth) { rh += ysPrev.h; rhu += ysPrev.hu; rhv += ysPrev.hv; }
      if (fyPrevOn) { rh -= yn.h; rhu -= yn.hu; rhv -= yn.hv; }
      const bool interior = xin && o >= 2 && o <= H - 3;
      const Cell cur = {ring[o & 3][0][lane], ring[o & 3][1][lane], ring[o & 3][2][lane], ring[o & 3][3][lane]};
      if (interior) {
        if (bedOn) {
          rhu += bedPrev2.hu;
          rhv += bedPrev2.h;
          if (sph) rhv += bedPrev2.hv;
        }
        if (sph) {                                        // swe.mjs:600-619
          const double fe = rowv(FROW, o) + Uo * rowv(GEO_TANR, o);   // tanPhi / R of the row
          rhu += fe * cur.hv;
          rhv -= fe * cur.hu;
        } else if (P.coriolis_on) {                       // swe.mjs:620-628
          rhu += P.coriolis * cur.hv;
          rhv -= P.coriolis * cur.hu;
        }
        // equilibrium tide: g h grad(eta_eq), eta_eq = sum A cos^2(phi) cos(w t + 2 lon + ph)
        // (semidiurnal) or A sin(2 phi) cos(w t + lon + ph) (diurnal); no reference counterpart
        // (the column factors from sweep_tide, formed before the row loop; the
        // row factors, wave-uniform, through the constant address space as the
        // row metric is read)
        for (int c = 0; c < (UNI ? 0 : P.tide_n); ++c) {
          const const_double *T = (const const_double *)P.tide + (size_t)c * (2 * H + 2 * W);
          rhu += cur.h * T[o] * sweep_tide[c][0][lane];
          rhv += cur.h * T[H + o] * sweep_tide[c][1][lane];
        }
      }
      if (MODE == MODE_PROBE) {
        A.rh[k] = rh; A.rh[k + fs] = rhu; A.rh[k + 2 * fs] = rhv; A.theta[k] = thPrev;
      } else if (MODE == MODE_STAGE1) {
        // swe.mjs:790-796: every cell, ghosts too (the next applyBC rewrites the
        // ghosts; a caller's own boundary function sees them as the reference
        // leaves them), then dryClean of the interior
        double nh = cur.h + dt * rh, nhu = cur.hu + dt * rhu, nhv = cur.hv + dt * rhv;
        if (interior) {
          if (nh < 0) { mf -= nh; nh = 0.0; }
          if (nh <= md) { nhu = 0.0; nhv = 0.0; }
        }
        outp[k] = nh; outp[k + fs] = nhu; outp[k + 2 * fs] = nhv;
      } else {                                            // stage 2, swe.mjs:799-812
        const double h0 = h0p[k], hu0 = h0p[k + fs], hv0 = h0p[k + 2 * fs];
        double nh = 0.5 * (h0 + cur.h + dt * rh);
        double nhu = 0.5 * (hu0 + cur.hu + dt * rhu);
        double nhv = 0.5 * (hv0 + cur.hv + dt * rhv);
        if (interior) {
          if (nh < 0) { mf -= nh; nh = 0.0; }
          if (nh <= md) { nhu = 0.0; nhv = 0.0; }
          if (REC && A.rec) { A.rec[k] = nh; A.rec[k + fs] = nhu; A.rec[k + 2 * fs] = nhv; }
          // applyFriction (swe.mjs:659-704)
          if (P.friction) {
            if (nh <= md) { nhu = 0.0; nhv = 0.0; }
            else {
              const double n = (!LEAN && P.friction == 2) ? A.mfield[k] : P.manning;
              if (n != 0.0 && n == n) {
                const Vel2 uv = vel2(nhu, nhv, nh, P.eps4);
                const double speed = js_hypot(uv.u, uv.v);
                if (!(speed < 1e-14)) {
                  const double denom = 1 + dt * P.g * n * n * speed / js_pow43(nh);
                  nhu = nhu / denom;
                  nhv = nhv / denom;
                }
              }
            }
          }
          // makeSponge forcing body (swe.mjs:1188-1197)
          if (!LEAN && P.sponge) {
            const double str = A.sponge[k];
            if (str > 0) {
              const double f = jmin(1.0, str * dt * 60);
              const double target = jmax0(P.sponge_ref - b[k]);
              nh += (target - nh) * f;
              nhu -= nhu * f;
              nhv -= nhv * f;
            }
          }
          // next step's maxDt (swe.mjs:632-647)
          if (nh > md) {
            const double c = sqrt_gh(P.g * nh, HC.normal);   // nh > minDepth
            const Vel2 uv = vel2(nhu, nhv, nh, P.eps4);
            const double sp = div_r(__builtin_fabs(uv.u) + c, rowv(DX, o), rowv(RDX, o)) +
                              div_r(__builtin_fabs(uv.v) + c, rowv(DYCFL, o), rowv(RDYCFL, o));
            if (sp > 0) {
              const unsigned long long bits = (unsigned long long)__double_as_longlong(sp);
              if (bits > cflb) cflb = bits;
            }
          }
        }
        if (!(__builtin_isfinite(nh) && __builtin_isfinite(nhu) && __builtin_isfinite(nhv))) bad = true;
        outp[k] = nh; outp[k + fs] = nhu; outp[k + 2 * fs] = nhv;
        // (an adjoint replay, REC, never asks for theta or the residual)
        if (!LEAN && !REC && A.P.want_theta) A.theta[k] = thPrev;
        if (!LEAN && !REC && A.P.want_rh) { A.rh[k] = rh; A.rh[k + fs] = rhu; A.rh[k + 2 * fs] = rhv; }
      }
    }

    // ---- rotate the pipeline ---------------------------------------------------
    Uo = cM.U;
    cM = c0; c0 = cP;
    nPrev = fN;
    fxPrev = fxCur; fyPrev = fyCur;
    thPrev = th;
    xpPrev = xpCur; ysPrev = ysCur;
    bedPrev2 = bedPrev; bedPrev = bedCur;
  }

  if (MODE == MODE_PROBE) return;
  // wave reductions: massFloored partial (fixed order), max wave speed, non-finite flag
  for (int off = 32; off >= 1; off >>= 1) mf += __shfl_down(mf, off, SWEEP_LANES);
  if (lane == 0) A.mf_part[wave_id] = mf;
  if (MODE == MODE_STAGE2) {
    for (int off = 32; off >= 1; off >>= 1) {
      const unsigned long long o2b = __shfl_down(cflb, off, SWEEP_LANES);
      if (o2b > cflb) cflb = o2b;
    }
    if (lane == 0 && cflb) atomicMax(&A.ctl->cfl_bits, cflb);
    if (__any(bad) && lane == 0) atomicOr(&A.ctl->nonfinite, 1);
  }
}

template <int MODE, int LIM, bool REC, bool UNI, bool LEAN = false>
__global__ __launch_bounds__(SWEEP_LANES) void k_sweep(SweepArgs A) {
  sweep_body<MODE, REC, LIM, UNI, LEAN>(A, blockIdx.x, blockIdx.y, blockIdx.y * gridDim.x + blockIdx.x);
}

// Batched ensembles: a persistent grid of wavefronts pulls (basin, strip, chunk)
// tasks from a queue until it is empty, so basins of very different sizes share
// one launch and the device stays full to the last task.  LIMS >= 0: every
// basin uses that slope limiter (and with UNI every basin's rows are alike), so
// the kernel holds that one body (LEAN: at stage 2, no basin has a per-cell
// Manning field or a sponge); LIMS < 0: each task's limiter is switched on in the
// kernel (three bodies, per-row metric).
template <int MODE, int LIMS, bool UNI, bool LEAN = false>
__global__ __launch_bounds__(SWEEP_LANES, 3) void k_sweep_batch(const BasinDev *__restrict__ basins,
                                                             const Task *__restrict__ tasks, int ntasks,
                                                             unsigned *qctr, int parity) {
  for (;;) {
    unsigned t = 0;
    if (threadIdx.x == 0) t = atomicAdd(qctr, 1u);
    // wave-uniform from here on: task and basin fields load through the scalar unit
    t = (unsigned)__builtin_amdgcn_readfirstlane((int)__shfl(t, 0, SWEEP_LANES));
    if (t >= (unsigned)ntasks) break;
    const int basin = __builtin_amdgcn_readfirstlane(tasks[t].basin);
    const int strip = __builtin_amdgcn_readfirstlane(tasks[t].strip);
    const int chunk = __builtin_amdgcn_readfirstlane(tasks[t].chunk);
    const int wave = __builtin_amdgcn_readfirstlane(tasks[t].wave);
    const SweepArgs &A = basins[basin].sw[parity][MODE - 1];
    if constexpr (LIMS >= 0) {
      sweep_body<MODE, false, LIMS, UNI, LEAN>(A, strip, chunk, wave);
    } else {
      const int lim = __builtin_amdgcn_readfirstlane(A.P.limiter);
      if (lim == LIM_MC) sweep_body<MODE, false, LIM_MC>(A, strip, chunk, wave);
      else if (lim == LIM_MINMOD) sweep_body<MODE, false, LIM_MINMOD>(A, strip, chunk, wave);
      else sweep_body<MODE, false, LIM_CENTRAL>(A, strip, chunk, wave);
    }
  }
}

// applyBC on stored arrays (swe.mjs:289-303): ghost cell t of the ring
// rewritten from the interior.  Every ghost value depends only on interior cells
// (a periodic S/N source in a ghost column is recomputed, not read), so the
// threads' writes race with no read.
__device__ __forceinline__ void apply_bc_cell(const Params &P, StateW S, double *b, const double *ext, int t) {
  int gi, gj;
  if (t < 4 * P.W) { const int rr = t / P.W; gi = t % P.W; gj = rr < 2 ? rr : P.H - 4 + rr; }
  else { const int u = t - 4 * P.W; const int cc = u / (P.H - 4); gj = 2 + u % (P.H - 4); gi = cc < 2 ? cc : P.W - 4 + cc; }
  const StateP R = {S.h, S.hu, S.hv};
  const Cell c = ghost_cell(P, R, b, ext, gi, gj);
  const int k = gj * P.W + gi;
  b[k] = c.b; S.h[k] = c.h; S.hu[k] = c.hu; S.hv[k] = c.hv;
}

__global__ void k_apply_bc(Params P, StateW S, double *b, const Ctl *ctl) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * P.W + 4 * (P.H - 4)) return;
  // the steps a sampled run launches past its end are not steps of the run:
  // they leave even the ghost ring alone
  if ((ctl->dt_mode == DT_SAMPLED || ctl->dt_mode == DT_UNTIL) && ctl->idle) return;
  apply_bc_cell(P, S, b, ctl->ext, t);
}

// maxDt() over the stored state (swe.mjs:632-647): the largest wave-speed sum of
// this thread's cells as a bit pattern (non-negative doubles order like integers).
__device__ __forceinline__ unsigned long long maxdt_cells(const Params &P, StateP S, int t0, int stride) {
  const int n = P.nx * P.ny;
  unsigned long long best = 0ull;
  for (int t = t0; t < n; t += stride) {
    const int j = 2 + t / P.nx, i = 2 + t % P.nx;
    const int k = j * P.W + i;
    const double hk = S.h[k];
    if (hk <= P.min_depth) continue;
    const double c = __builtin_sqrt(P.g * hk);
    const Vel2 uv = vel2(S.hu[k], S.hv[k], hk, P.eps4);
    const double sp = div_r(__builtin_fabs(uv.u) + c, P.geo.dx[j], P.geo.rdx[j]) +
                      div_r(__builtin_fabs(uv.v) + c, P.geo.dyCFL[j], P.geo.rdyCFL[j]);
    if (sp > 0) {
      const unsigned long long bits = (unsigned long long)__double_as_longlong(sp);
      if (bits > best) best = bits;
    }
  }
  for (int off = 32; off >= 1; off >>= 1) {
    const unsigned long long o = __shfl_down(best, off, 64);
    if (o > best) best = o;
  }
  return best;
}

__global__ void k_maxdt(Params P, StateP S, Ctl *ctl) {
  const unsigned long long best = maxdt_cells(P, S, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
  if ((threadIdx.x & 63) == 0 && best) atomicMax(&ctl->cfl_bits, best);
}

// applyFriction(dt) on the stored state (swe.mjs:659-704), interior cells.
__global__ void k_friction(Params P, StateW S, const double *mfield, double dt) {
  const int n = P.nx * P.ny;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int k = (2 + t / P.nx) * P.W + 2 + t % P.nx;
    const double hk = S.h[k];
    if (hk <= P.min_depth) { S.hu[k] = 0.0; S.hv[k] = 0.0; continue; }
    const double nk = P.friction == 2 ? mfield[k] : P.manning;
    if (nk == 0.0 || nk != nk) continue;
    double hu = S.hu[k], hv = S.hv[k];
    const Vel2 uv = vel2(hu, hv, hk, P.eps4);
    const double speed = js_hypot(uv.u, uv.v);
    if (speed < 1e-14) continue;
    const double denom = 1 + dt * P.g * nk * nk * speed / js_pow_pos(hk, 4.0 / 3.0);
    S.hu[k] = hu / denom;
    S.hv[k] = hv / denom;
  }
}

// dryClean() on the stored state (swe.mjs:734-745); the floored mass of each
// block goes to part[block] (summed in block order by the host).
__global__ void k_dry_clean(Params P, StateW S, double *part) {
  const int n = P.nx * P.ny;
  double added = 0.0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int k = (2 + t / P.nx) * P.W + 2 + t % P.nx;
    double h = S.h[k];
    if (h < 0) { added -= h; h = 0.0; S.h[k] = 0.0; }
    if (h <= P.min_depth) { S.hu[k] = 0.0; S.hv[k] = 0.0; }
  }
  for (int off = 32; off >= 1; off >>= 1) added += __shfl_down(added, off, 64);
  __shared__ double w[16];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = added;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += w[i];
    part[blockIdx.x] = s;
  }
}

hipError_t launch_friction(const Params &P, StateW S, const double *mfield, double dt, hipStream_t st) {
  int blocks = (P.nx * P.ny + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_friction, dim3(blocks), dim3(256), 0, st, P, S, mfield, dt);
  return hipGetLastError();
}

hipError_t launch_dry_clean(const Params &P, StateW S, double *part, int blocks, hipStream_t st) {
  hipLaunchKernelGGL(k_dry_clean, dim3(blocks), dim3(256), 0, st, P, S, part);
  return hipGetLastError();
}

__global__ void k_tide_times(Ctl *ctl) {
  (void)tide_times_lane(ctl, ctl->t, ctl->dt, threadIdx.x);
}

hipError_t launch_tide_times(Ctl *ctl, hipStream_t st) {
  hipLaunchKernelGGL(k_tide_times, dim3(1), dim3(64), 0, st, ctl);
  return hipGetLastError();
}

// c0 + sum_k a_k cos(om_k t - ph_k), a_k = amp_k, or amp_k min(1, t / ramp)
// when ramp > 0 (src/waves.mjs:166-167: msl + ramp(t) * a * cos(w t))
__device__ __forceinline__ double harm_eval(int n, double c0, const double *amp, const double *om,
                                            const double *ph, double t, double ramp) {
  double e = c0;
  if (ramp > 0) {
    const double m = fmin(1.0, t / ramp);
    for (int k = 0; k < n; ++k) e += (m * amp[k]) * cos(om[k] * t - ph[k]);
  } else {
    for (int k = 0; k < n; ++k) e += amp[k] * cos(om[k] * t - ph[k]);
  }
  return e;
}

// Turn the max-reduction into a step size and set up the next step (one
// wavefront per basin).
//   phase 0: after k_maxdt -- dt for the step about to run
//   phase 1: after a step -- massFloored, t, steps, dt_next and next dt
// s1, s2: the two stages' massFloored sums (phase 1); the body runs on one lane.
// rec: the step's gauge recorder is armed (device-resident runs only).  The
// step's forcing follows on the wavefront's lanes (forcing_lanes).
__device__ __forceinline__ void finalize_ctl(Ctl *ctl, double s1, double s2, double cfl, int phase, bool rec = false) {
  const unsigned long long bits = ctl->cfl_bits;
  const double inv = __longlong_as_double((long long)bits);
  const double dtn = bits ? cfl / inv : __builtin_inf();
  ctl->cfl_bits = 0ull;
  ctl->dt_next = dtn;
  if (phase == 1) {
    const double dt = ctl->dt;
    ctl->rec_flag = 0;
    if ((dt > 0) && !__builtin_isinf(dt)) {
      ctl->mf = s1 + s2;
      ctl->t = ctl->t + dt;
      ctl->steps = ctl->steps + 1;
      ctl->dt_last = dt;
      if (rec) {
        bool take = false;
        if (ctl->rec_every > 0) {
          if (++ctl->rec_since >= ctl->rec_every) { ctl->rec_since = 0; take = true; }
        } else if (ctl->t >= ctl->rec_next) {
          // tools/waves.mjs:224-227: a sample once t reaches the next sample
          // time (from the analysis start on), and the next time moves on
          take = ctl->t >= ctl->rec_t0;
          ctl->rec_next = ctl->rec_next + ctl->rec_dt;
        }
        if (take) { ctl->rec_row = ctl->rec_count; ctl->rec_count = ctl->rec_count + 1; ctl->rec_flag = 1; }
      }
    } else {
      ctl->dt_last = 0.0;
    }
  }
  if (ctl->dt_mode == DT_SELF) ctl->dt = dtn;
  else if (ctl->dt_mode == DT_FIXED) ctl->dt = ctl->dt_fixed;
  else if (ctl->dt_mode == DT_UNTIL) ctl->dt = fmin(dtn, ctl->t_end - ctl->t);
  else if (ctl->dt_mode == DT_SAMPLED)
    // tools/waves.mjs:222-223: while (t < T) step(min(maxDt, max(1e-6, min(T, next) - t)));
    // at t >= T the step is refused (dt 0), which leaves the state as it is
    ctl->dt = ctl->t < ctl->t_end ? jmin(dtn, jmax(1e-6, jmin(ctl->t_end, ctl->rec_next) - ctl->t)) : 0.0;
  ctl->idle = (ctl->dt_mode == DT_SAMPLED || ctl->dt_mode == DT_UNTIL) && !(ctl->t < ctl->t_end);
}

// The forcing of the step about to run, after finalize_ctl has set t and dt
// (passed in, the same on every lane): the tide factors on lanes 0-15, and the
// harmonic Flather values on lanes 32-39, two per side (eta on the even lane, u
// on the odd one, which evaluates eta as well when u follows it).
__device__ __forceinline__ void forcing_lanes(Ctl *ctl, double t, double dt, int lane) {
  (void)tide_times_lane(ctl, t, dt, lane);
  if (ctl->harm_on && lane >= 32 && lane < 40) {
    const int side = (lane - 32) >> 1;
    const Harm &h = ctl->harm[side];
    const double e = harm_eval(h.n_eta, h.msl, h.amp, h.omega, h.phase, t, h.ramp);
    if ((lane & 1) == 0) ctl->ext[2 * side] = e;
    else ctl->ext[2 * side + 1] = h.u_from_eta ? h.u_scale * (e - h.msl)
                                               : harm_eval(h.n_u, h.u0, h.uamp, h.uomega, h.uphase, t, h.ramp);
  }
}

// massFloored partials of both stages summed in a fixed order by one wavefront
// (batched basins: few partials each)
__device__ __forceinline__ void finalize_body(Ctl *ctl, const double *mf_part, int nparts, double cfl, int phase) {
  const int lane = threadIdx.x & 63;
  double s1 = 0.0, s2 = 0.0;
  if (phase == 1) {
    for (int i = lane; i < nparts; i += 64) { s1 += mf_part[i]; s2 += mf_part[nparts + i]; }
    for (int off = 32; off >= 1; off >>= 1) { s1 += __shfl_down(s1, off, 64); s2 += __shfl_down(s2, off, 64); }
  }
  if (lane == 0) finalize_ctl(ctl, s1, s2, cfl, phase);
  __syncthreads();
  forcing_lanes(ctl, ctl->t, ctl->dt, lane);
}

// One grid's partials (thousands per stage) by FIN_THREADS threads: strided
// per-thread sums, a shuffle tree per wavefront, then the wavefronts' sums in
// order, so the result is the same on every run.
constexpr int FIN_THREADS = 1024;
// With a recorder (r.cells set, phase 1), the step's gauge sample is written by
// the whole block once the control lane has decided on it.
__global__ __launch_bounds__(FIN_THREADS) void k_finalize(Ctl *ctl, const double *mf_part, int nparts, double cfl,
                                                          int phase, RecArgs r) {
  __shared__ double w1[FIN_THREADS / 64], w2[FIN_THREADS / 64];
  __shared__ long long rrow;
  __shared__ double rt, rdt;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const bool rec = r.cells != nullptr && phase == 1;
  double s1 = 0.0, s2 = 0.0;
  if (phase == 1) {
    for (int i = t; i < nparts; i += FIN_THREADS) { s1 += mf_part[i]; s2 += mf_part[nparts + i]; }
    for (int off = 32; off >= 1; off >>= 1) { s1 += __shfl_down(s1, off, 64); s2 += __shfl_down(s2, off, 64); }
    if (lane == 0) { w1[wv] = s1; w2[wv] = s2; }
    __syncthreads();
    if (t == 0) {
      s1 = 0.0; s2 = 0.0;
      for (int q = 0; q < FIN_THREADS / 64; ++q) { s1 += w1[q]; s2 += w2[q]; }
    }
  }
  if (t == 0) {
    finalize_ctl(ctl, s1, s2, cfl, phase, rec);
    rrow = (rec && ctl->rec_flag) ? ctl->rec_row : -1;
    rt = ctl->t;
    rdt = ctl->dt;
  }
  __syncthreads();
  if (t < 64) forcing_lanes(ctl, rt, rdt, t);
  if (!rec) return;
  const long long row = rrow;
  if (row < 0 || row >= r.cap) return;
  // tools/waves.mjs:225: sim.eta(i, j) = b + h
  for (int g = t; g < r.n; g += FIN_THREADS) {
    const int k = r.cells[g];
    r.series[row * r.n + g] = r.b[k] + r.h[k];
  }
  if (t == 0) r.times[row] = rt;
}

// ---- batched ensembles: one launch per phase for all basins -----------------
__global__ void k_apply_bc_batch(const BasinDev *__restrict__ basins, int parity, int stage) {
  const BasinDev &B = basins[blockIdx.y];
  const Params &P = B.sw[parity][0].P;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * P.W + 4 * (P.H - 4)) return;
  apply_bc_cell(P, B.bc_in[parity][stage], B.sw[parity][0].b, B.ctl->ext, t);
}

__global__ void k_maxdt_batch(const BasinDev *__restrict__ basins, int parity) {
  const BasinDev &B = basins[blockIdx.y];
  const SweepArgs &A = B.sw[parity][0];
  const unsigned long long best = maxdt_cells(A.P, A.in, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
  if ((threadIdx.x & 63) == 0 && best) atomicMax(&B.ctl->cfl_bits, best);
}

__global__ void k_finalize_batch(const BasinDev *__restrict__ basins, int nb, unsigned *qctr, int phase) {
  const BasinDev &B = basins[blockIdx.x];
  finalize_body(B.ctl, B.sw[0][0].mf_part, B.nparts, B.sw[0][0].P.cfl, phase);
  // the next step's stage launches start their task queues from zero
  if (blockIdx.x == 0 && threadIdx.x == 0) { qctr[0] = 0u; qctr[1] = 0u; }
}

// ---- launchers (host side, C++ linkage inside the library) ------------------
// the stage kernel for a limiter; the recording variant only when the launch
// records (the residual probe never does), the uniform-row variant for grids
// whose rows are all alike when it does not, and at stage 2 the lean one of
// either when the launch has none of the per-cell Manning field, the sponge and
// the theta / residual outputs.  (A captured step graph holds the kernel chosen at
// capture: every setter that changes one of these conditions drops the graphs.)
template <int MODE, int LIM>
static const void *sweep_fn(bool rec, bool uni, bool lean) {
  if (MODE != MODE_PROBE && rec) return reinterpret_cast<const void *>(&k_sweep<MODE, LIM, MODE != MODE_PROBE, false>);
  if (MODE == MODE_STAGE2 && lean)
    return uni ? reinterpret_cast<const void *>(&k_sweep<MODE, LIM, false, true, true>)
               : reinterpret_cast<const void *>(&k_sweep<MODE, LIM, false, false, true>);
  if (uni) return reinterpret_cast<const void *>(&k_sweep<MODE, LIM, false, true>);
  return reinterpret_cast<const void *>(&k_sweep<MODE, LIM, false, false>);
}
template <int MODE>
static const void *sweep_fn(int lim, bool rec, bool uni, bool lean = false) {
  if (lim == LIM_MINMOD) return sweep_fn<MODE, LIM_MINMOD>(rec, uni, lean);
  if (lim == LIM_CENTRAL) return sweep_fn<MODE, LIM_CENTRAL>(rec, uni, lean);
  return sweep_fn<MODE, LIM_MC>(rec, uni, lean);
}

hipError_t launch_sweep(int mode, const SweepArgs &a, hipStream_t st) {
  dim3 grid((a.P.W + SWEEP_OUT - 1) / SWEEP_OUT, (a.P.H + a.rows_per_chunk - 1) / a.rows_per_chunk);
  dim3 block(SWEEP_LANES);
  const bool rec = a.recf != nullptr, uni = a.P.geo_uni != 0;
  const bool lean = a.P.friction != 2 && !a.P.sponge && !a.P.want_theta && !a.P.want_rh;
  const void *fn = mode == MODE_STAGE1   ? sweep_fn<MODE_STAGE1>(a.P.limiter, rec, uni)
                   : mode == MODE_STAGE2 ? sweep_fn<MODE_STAGE2>(a.P.limiter, rec, uni, lean)
                                         : sweep_fn<MODE_PROBE>(a.P.limiter, false, uni);
  SweepArgs arg = a;
  void *args[] = {&arg};
  return hipLaunchKernel(fn, grid, block, args, 0, st);
}

const void *k_sweep_stage1_ptr() { return sweep_fn<MODE_STAGE1>(LIM_MC, false, false); }

int sweep_waves(const Params &P, int rows_per_chunk) {
  return ((P.W + SWEEP_OUT - 1) / SWEEP_OUT) * ((P.H + rows_per_chunk - 1) / rows_per_chunk);
}

// the batch stage kernel for the ensemble's common limiter (lim_all < 0: mixed),
// whether all its basins have uniform rows, and (stage 2) whether none has a
// per-cell Manning field or a sponge
template <int MODE, int LIM>
static const void *sweep_batch_one(bool uni_all, bool lean_all) {
  if (MODE == MODE_STAGE2 && lean_all)
    return uni_all ? reinterpret_cast<const void *>(&k_sweep_batch<MODE, LIM, true, true>)
                   : reinterpret_cast<const void *>(&k_sweep_batch<MODE, LIM, false, true>);
  return uni_all ? reinterpret_cast<const void *>(&k_sweep_batch<MODE, LIM, true>)
                 : reinterpret_cast<const void *>(&k_sweep_batch<MODE, LIM, false>);
}
template <int MODE>
static const void *sweep_batch_fn(int lim_all, bool uni_all, bool lean_all) {
  switch (lim_all) {
    case LIM_MC: return sweep_batch_one<MODE, LIM_MC>(uni_all, lean_all);
    case LIM_MINMOD: return sweep_batch_one<MODE, LIM_MINMOD>(uni_all, lean_all);
    case LIM_CENTRAL: return sweep_batch_one<MODE, LIM_CENTRAL>(uni_all, lean_all);
    default: return reinterpret_cast<const void *>(&k_sweep_batch<MODE, -1, false>);
  }
}

hipError_t launch_sweep_batch(int mode, const BasinDev *basins, const Task *tasks, int ntasks, unsigned *qctr,
                              int parity, int nwaves, int lim_all, bool uni_all, bool lean_all, hipStream_t st) {
  const void *fn = mode == MODE_STAGE1 ? sweep_batch_fn<MODE_STAGE1>(lim_all, uni_all, false)
                                       : sweep_batch_fn<MODE_STAGE2>(lim_all, uni_all, lean_all);
  unsigned *q = mode == MODE_STAGE1 ? qctr : qctr + 1;
  void *args[] = {&basins, &tasks, &ntasks, &q, &parity};
  return hipLaunchKernel(fn, dim3(nwaves), dim3(SWEEP_LANES), args, 0, st);
}

const void *k_sweep_batch_ptr() { return sweep_batch_fn<MODE_STAGE1>(-1, false, false); }

hipError_t launch_apply_bc_batch(const BasinDev *basins, int nb, int max_ghost, int parity, int stage, hipStream_t st) {
  hipLaunchKernelGGL(k_apply_bc_batch, dim3((max_ghost + 255) / 256, nb), dim3(256), 0, st, basins, parity, stage);
  return hipGetLastError();
}

hipError_t launch_maxdt_batch(const BasinDev *basins, int nb, int max_cells, int parity, hipStream_t st) {
  int bx = (max_cells + 255) / 256;
  if (bx > 64) bx = 64;
  hipLaunchKernelGGL(k_maxdt_batch, dim3(bx, nb), dim3(256), 0, st, basins, parity);
  return hipGetLastError();
}

hipError_t launch_finalize_batch(const BasinDev *basins, int nb, unsigned *qctr, int phase, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize_batch, dim3(nb), dim3(64), 0, st, basins, nb, qctr, phase);
  return hipGetLastError();
}

hipError_t launch_apply_bc(const Params &P, StateW S, double *b, const Ctl *ctl, hipStream_t st) {
  const int nG = 4 * P.W + 4 * (P.H - 4);
  hipLaunchKernelGGL(k_apply_bc, dim3((nG + 255) / 256), dim3(256), 0, st, P, S, b, ctl);
  return hipGetLastError();
}

hipError_t launch_maxdt(const Params &P, StateP S, Ctl *ctl, hipStream_t st) {
  const int n = P.nx * P.ny;
  int blocks = (n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_maxdt, dim3(blocks), dim3(256), 0, st, P, S, ctl);
  return hipGetLastError();
}

hipError_t launch_finalize(Ctl *ctl, const double *mf_part, int nparts, double cfl, int phase, hipStream_t st,
                           const RecArgs *rec) {
  RecArgs r{};
  if (rec) r = *rec;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, st, ctl, mf_part, nparts, cfl, phase, r);
  return hipGetLastError();
}

}  // namespace swe
