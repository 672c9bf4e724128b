// Reverse-mode (adjoint) pieces of the shallow-water step: for each local
// function of the forward step, given the forward inputs and the adjoint of its
// outputs, the adjoint of its inputs.  Branches and kinks take the subgradient of
// the branch the forward took (the limiter's selected slope, HLLC's wave
// pattern, Audusse's max(0, .), the wet/dry thresholds).  The dt sequence is not
// differentiated: it is held at the forward's values.
//
// Everything here is __host__ __device__: the kernels run it one thread per
// cell, and the library's host entry points (swe_adj_host_*) run the same code
// in loops so that it can be checked on a machine without a GPU.
#pragma once
#include "swe_device.hpp"
#include "swe_kernels.hpp"

namespace swe {

// d vel(hq, h) / d hq and / d h  (vel = SQRT2 h hq / sqrt(h^4 + max(h^4, eps4)))
struct VelD { double dhq, dh; };
__host__ __device__ __forceinline__ VelD vel_deriv(double hq, double h, double eps4) {
  const double h4 = h * h * h * h;
  const bool self = h4 > eps4;                       // max(h4, eps4) picks h4
  const double D = __builtin_sqrt(h4 + (self ? h4 : eps4));
  const double rD = 1.0 / D;
  const double dD = (4.0 * h * h * h * (self ? 2.0 : 1.0)) * (0.5 * rD);
  VelD r;
  r.dhq = SQRT2 * h * rD;
  r.dh = SQRT2 * hq * rD - SQRT2 * h * hq * dD * (rD * rD);
  return r;
}

// (d lim / d a, d lim / d b) on the branch the limiter takes
struct LimD { double da, db; };
__host__ __device__ __forceinline__ LimD limit_deriv(int lim, double a, double b) {
  LimD r = {0.0, 0.0};
  if (lim == LIM_MC) {
    if (a * b <= 0) return r;
    const double c = 0.5 * (a + b);
    const double ac = __builtin_fabs(c), a2 = 2 * __builtin_fabs(a), b2 = 2 * __builtin_fabs(b);
    // the same selection jmin(jmin(|c|, 2|a|), 2|b|) makes
    const double m1 = ac < a2 ? ac : a2;
    if (!(b2 < m1)) {
      if (ac < a2) { r.da = 0.5; r.db = 0.5; }
      else { r.da = 2.0; }
    } else {
      r.db = 2.0;
    }
    return r;
  }
  if (lim == LIM_MINMOD) {
    if (a > 0 && b > 0) { if (a < b) r.da = 1.0; else r.db = 1.0; }
    else if (a < 0 && b < 0) { if (a > b) r.da = 1.0; else r.db = 1.0; }
    return r;
  }
  r.da = 0.5; r.db = 0.5;
  return r;
}

// 1/x to about an ulp, for factors that only scale derivatives (on the device
// the hardware estimate and one Newton-Raphson step; on the host 1/x): never for
// a value a branch of the forward's recomputation compares
__host__ __device__ __forceinline__ double rcp_d(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rcp(x);
  return __builtin_fma(__builtin_fma(-x, y, 1.0), y, y);
#else
  return 1.0 / x;
#endif
}

// Reverse of hllc(hL, huL, hvL, hR, huR, hvR) (swe.mjs:974-1011).  The forward's
// wave pattern (velocities, celerities, hStar, qL, qR, sL, sR, sM) is recomputed
// with the forward's exact operations, so every branch goes the way the forward
// went; the reciprocals that only scale derivatives are rcp_d.
struct HllcIn { double hL, huL, hvL, hR, huR, hvR; };
__host__ __device__ inline HllcIn hllc_rev(double hL, double huL, double hvL, double hR, double huR, double hvR,
                                           const HllcConst &C, double lm, double ln, double lt) {
  HllcIn g = {0, 0, 0, 0, 0, 0};
  const bool wetL = hL > C.dry, wetR = hR > C.dry;
  if (!wetL && !wetR) return g;
  const double G = C.g;
  double uL = 0.0, vL = 0.0, uR = 0.0, vR = 0.0;
  if (wetL) { uL = huL / hL; vL = hvL / hL; }
  if (wetR) { uR = huR / hR; vR = hvR / hR; }
  const double cL = sqrt_gh(G * jmax(hL, 0.0), C.normal), cR = sqrt_gh(G * jmax(hR, 0.0), C.normal);
  double sL, sR, a = 0, hStar = 0, qL = 1, qR = 1;
  if (!wetL) { sL = uR - 2 * cR; sR = uR + cR; }
  else if (!wetR) { sL = uL - cL; sR = uL + 2 * cL; }
  else {
    a = (cL + cR) * 0.5 + (uL - uR) * 0.25;
    hStar = div_r(a * a, G, C.rg);
    qL = hStar > hL ? sqrt_n(0.5 * (hStar + hL) * hStar / (hL * hL)) : 1.0;
    qR = hStar > hR ? sqrt_n(0.5 * (hStar + hR) * hStar / (hR * hR)) : 1.0;
    sL = uL - cL * qL;
    sR = uR + cR * qR;
  }
  // adjoints of u, v, c (velocities and celerities of the two states)
  double luL = 0, lvL = 0, luR = 0, lvR = 0, lcL = 0, lcR = 0;
  if (sL >= 0) {                                   // supercritical to the right: F = f(L)
    g.huL += lm + ln * uL;
    luL += ln * huL + lt * hvL;
    g.hL += ln * G * hL;
    g.hvL += lt * uL;
  } else if (sR <= 0) {                            // supercritical to the left: F = f(R)
    g.huR += lm + ln * uR;
    luR += ln * huR + lt * hvR;
    g.hR += ln * G * hR;
    g.hvR += lt * uR;
  } else {
    const double num = sL * hR * (uR - sR) - sR * hL * (uL - sL);
    const double den = hR * (uR - sR) - hL * (uL - sL);
    const double sM = __builtin_fabs(den) > 1e-30 ? num / den : 0.5 * (sL + sR);
    const double FL0 = huL, FL1 = huL * uL + 0.5 * G * hL * hL;
    const double FR0 = huR, FR1 = huR * uR + 0.5 * G * hR * hR;
    const double inv = 1.0 / (sR - sL);
    const double dh = hR - hL, dq = huR - huL;
    const double Nm = sR * FL0 - sL * FR0 + sL * sR * dh;
    const double Nn = sR * FL1 - sL * FR1 + sL * sR * dq;
    const double m = Nm * inv;
    const bool left = sM >= 0;
    // t = m * vsel
    const double lm2 = lm + lt * (left ? vL : vR);
    if (left) lvL += lt * m; else lvR += lt * m;
    const double lNm = lm2 * inv, lNn = ln * inv;
    const double linv = lm2 * Nm + ln * Nn;
    double lsL = linv * inv * inv, lsR = -linv * inv * inv;
    lsR += lNm * (FL0 + sL * dh) + lNn * (FL1 + sL * dq);
    lsL += lNm * (-FR0 + sR * dh) + lNn * (-FR1 + sR * dq);
    const double lFL0 = lNm * sR, lFR0 = -lNm * sL, lFL1 = lNn * sR, lFR1 = -lNn * sL;
    g.hR += lNm * sL * sR; g.hL -= lNm * sL * sR;
    g.huR += lNn * sL * sR; g.huL -= lNn * sL * sR;
    g.huL += lFL0 + lFL1 * uL; luL += lFL1 * huL; g.hL += lFL1 * G * hL;
    g.huR += lFR0 + lFR1 * uR; luR += lFR1 * huR; g.hR += lFR1 * G * hR;
    // wave speeds
    if (!wetL) { luR += lsL + lsR; lcR += -2 * lsL + lsR; }
    else if (!wetR) { luL += lsL + lsR; lcL += -lsL + 2 * lsR; }
    else {
      luL += lsL; lcL -= lsL * qL; const double lqL = -lsL * cL;
      luR += lsR; lcR += lsR * qR; const double lqR = lsR * cR;
      // X = (hStar + h) hStar / (2 h^2) and qL = sqrt(X) (qR likewise)
      double lhStar = 0.0;
      if (hStar > hL) {
        const double r = rcp_d(hL), r2 = r * r;
        const double X = 0.5 * (hStar + hL) * hStar * r2;
        const double lX = lqL * (0.5 * rcp_d(qL));
        lhStar += lX * 0.5 * (2.0 * hStar + hL) * r2;
        g.hL += lX * (0.5 * hStar * r2 - 2.0 * X * r);
      }
      if (hStar > hR) {
        const double r = rcp_d(hR), r2 = r * r;
        const double X = 0.5 * (hStar + hR) * hStar * r2;
        const double lX = lqR * (0.5 * rcp_d(qR));
        lhStar += lX * 0.5 * (2.0 * hStar + hR) * r2;
        g.hR += lX * (0.5 * hStar * r2 - 2.0 * X * r);
      }
      const double la = lhStar * 2.0 * a * C.rg;
      lcL += 0.5 * la; lcR += 0.5 * la; luL += 0.25 * la; luR -= 0.25 * la;
    }
  }
  // u = hu / h, c = sqrt(G h)
  if (wetL) {
    const double r = rcp_d(hL);
    g.hL += lcL * G * (0.5 * rcp_d(cL));
    g.huL += luL * r; g.hvL += lvL * r;
    g.hL -= (luL * uL + lvL * vL) * r;
  }
  if (wetR) {
    const double r = rcp_d(hR);
    g.hR += lcR * G * (0.5 * rcp_d(cR));
    g.huR += luR * r; g.hvR += lvR * r;
    g.hR -= (luR * uR + lvR * vR) * r;
  }
  return g;
}

// Reverse of the interface flux with hydrostatic reconstruction (face_flux):
// inputs eta, u (normal), v (tangential) of both sides (beds are constants).
struct FaceIn { double etaL, uL, vL, etaR, uR, vR; };
__host__ __device__ inline FaceIn face_rev(double etaL, double bL, double uL, double vL, double etaR, double bR,
                                           double uR, double vR, const HllcConst &C, double lM, double lNL,
                                           double lNR, double lT) {
  const double hLf = jmax0(etaL - bL), hRf = jmax0(etaR - bR);
  const double bStar = jmax(bL, bR);
  const double hsL = jmax0(etaL - bStar), hsR = jmax0(etaR - bStar);
  const double G = C.g;
  const double ln = lNL + lNR;
  const HllcIn gh = hllc_rev(hsL, hsL * uL, hsL * vL, hsR, hsR * uR, hsR * vR, C, lM, ln, lT);
  // pL = G/2 (hLf^2 - hsL^2), NL = n + pL
  const double lhLf = lNL * G * hLf, lhRf = lNR * G * hRf;
  const double lhsL = gh.hL + gh.huL * uL + gh.hvL * vL - lNL * G * hsL;
  const double lhsR = gh.hR + gh.huR * uR + gh.hvR * vR - lNR * G * hsR;
  FaceIn f;
  f.uL = gh.huL * hsL; f.vL = gh.hvL * hsL;
  f.uR = gh.huR * hsR; f.vR = gh.hvR * hsR;
  f.etaL = (etaL - bL > 0 ? lhLf : 0.0) + (etaL - bStar > 0 ? lhsL : 0.0);
  f.etaR = (etaR - bR > 0 ? lhRf : 0.0) + (etaR - bStar > 0 ? lhsR : 0.0);
  return f;
}

// ---- residual^T as per-cell passes (swe.mjs:308-629) ---------------------------
// Every pass is a function of one cell (i, j) of the W x H grid that writes only
// that cell's slots, so the passes run as one-thread-per-cell kernels on the
// device and as plain loops on the host.
enum Slot {
  // forward recompute: velocities, surface, slopes (x, y), faces, positivity
  S_U, S_V, S_ETA, S_SEX, S_SEY, S_SUX, S_SUY, S_SVX, S_SVY, S_SBX, S_SBY,
  S_FXM, S_FXNL, S_FXNR, S_FXT, S_FYM, S_FYNL, S_FYNR, S_FYT, S_TH, S_OUT,
  // adjoints: face flux terms per face, theta^T per cell
  S_LXM, S_LXNL, S_LXNR, S_LXT, S_LYM, S_LYNL, S_LYNR, S_LYT, S_LOUT, S_LHTH,
  // adjoints of the reconstructed face states (grid frame: eta, u, v; L and R)
  S_LXETAL, S_LXUL, S_LXVL, S_LXETAR, S_LXUR, S_LXVR,
  S_LYETAL, S_LYUL, S_LYVL, S_LYETAR, S_LYUR, S_LYVR,
  // adjoints per cell: eta, U, V, slopes, Coriolis into hu / hv
  S_LETA, S_LU, S_LV, S_LSEX, S_LSEY, S_LSUX, S_LSUY, S_LSVX, S_LSVY, S_LHUC, S_LHVC,
  S_COUNT
};
struct Work { double *s[S_COUNT]; };

// T[idx] for an idx that is the same on every active lane: on the device through
// the constant address space with the index taken from the first active lane (a
// scalar load, waited on apart from the vector loads).  The device callers of
// adj_faces_rev_eval and cells_rev run one grid row per wavefront (the tiled
// passes' own cells), so a row's index qualifies.
__host__ __device__ inline double row_uniform(const double *T, int idx) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const double __attribute__((address_space(4))) const_double;
  return ((const const_double *)T)[__builtin_amdgcn_readfirstlane(idx)];
#else
  return T[idx];
#endif
}


__host__ __device__ inline bool xface_on(const Params &P, int i, int j) {
  return j >= 2 && j <= P.H - 3 && i >= 1 && i <= P.W - 3;
}
__host__ __device__ inline bool yface_on(const Params &P, int i, int j) {
  return j >= 1 && j <= P.H - 3 && i >= 2 && i <= P.W - 3;
}

// pass 1: velocities, surface and limited slopes (swe.mjs:358-411)
__host__ __device__ inline void adj_cells_at(const Params &P, const StateP &S, const double *b, const Work &w,
                                             int i, int j) {
  const int k = j * P.W + i;
  const double h = S.h[k];
  const Vel2 uv = vel2(S.hu[k], S.hv[k], h, P.eps4);
  w.s[S_U][k] = uv.u; w.s[S_V][k] = uv.v; w.s[S_ETA][k] = b[k] + h;
  double sl[8] = {0, 0, 0, 0, 0, 0, 0, 0};        // sEx sEy sUx sUy sVx sVy sBx sBy
  const bool in = P.order >= 2 && i >= 1 && i <= P.W - 2 && j >= 1 && j <= P.H - 2;
  const double bk = b[k];
  for (int d = 0; d < 2; ++d) {
    // neighbours loaded before the branches (cell k stands in off the grid)
    const int kp = in ? (d == 0 ? k + 1 : k + P.W) : k, km = in ? (d == 0 ? k - 1 : k - P.W) : k;
    const double hp = S.h[kp], hm = S.h[km], bp = b[kp], bm = b[km];
    const double hup = S.hu[kp], hvp = S.hv[kp], hum = S.hu[km], hvm = S.hv[km];
    if (!in || h <= P.min_depth || hp <= P.min_depth || hm <= P.min_depth) continue;
    const double eC = bk + h, eP = bp + hp, eM = bm + hm;
    const Vel2 up = vel2(hup, hvp, hp, P.eps4), um = vel2(hum, hvm, hm, P.eps4);
    sl[0 + d] = limit(P.limiter, eC - eM, eP - eC);
    sl[6 + d] = limit(P.limiter, bk - bm, bp - bk);
    sl[2 + d] = limit(P.limiter, uv.u - um.u, up.u - uv.u);
    sl[4 + d] = limit(P.limiter, uv.v - um.v, up.v - uv.v);
  }
  for (int q = 0; q < 8; ++q) w.s[S_SEX + q][k] = sl[q];
}

struct FaceSide { double eta, b, u, v; };
// the reconstructed state of cell k on its + (east/north) or - side along d
__host__ __device__ inline FaceSide face_side(const Work &w, const double *b, int k, int d, double sgn) {
  FaceSide f;
  f.eta = w.s[S_ETA][k] + sgn * 0.5 * w.s[S_SEX + d][k];
  f.b = b[k] + sgn * 0.5 * w.s[S_SBX + d][k];
  f.u = w.s[S_U][k] + sgn * 0.5 * w.s[S_SUX + d][k];
  f.v = w.s[S_V][k] + sgn * 0.5 * w.s[S_SVX + d][k];
  return f;
}

// pass 2: the east x-face and the north y-face of cell (i, j) (swe.mjs:413-475);
// the slots hold M, N + pL, N + pR, T
__host__ __device__ inline void adj_faces_at(const Params &P, const double *b, const Work &w, int i, int j) {
  const int k = j * P.W + i;
  const HllcConst HC = {P.g, P.half_g, P.rg, P.min_depth, P.g * P.min_depth >= SQRT_N_MIN};
  double f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int d = 0; d < 2; ++d) {
    // both sides loaded before the branch (cell k stands in for a face off the grid)
    const bool on = d == 0 ? xface_on(P, i, j) : yface_on(P, i, j);
    const FaceSide L = face_side(w, b, k, d, 1.0), R = face_side(w, b, on ? (d == 0 ? k + 1 : k + P.W) : k, d, -1.0);
    if (!on) continue;
    const double nL = d == 0 ? L.u : L.v, tL = d == 0 ? L.v : L.u;
    const double nR = d == 0 ? R.u : R.v, tR = d == 0 ? R.v : R.u;
    const double hLf = jmax0(L.eta - L.b), hRf = jmax0(R.eta - R.b);
    const double bS = jmax(L.b, R.b);
    const double hsL = jmax0(L.eta - bS), hsR = jmax0(R.eta - bS);
    const Flux3 fl = hllc(hsL, hsL * nL, hsL * tL, hsR, hsR * nR, hsR * tR, HC);
    f[4 * d + 0] = fl.m;
    f[4 * d + 1] = fl.n + P.half_g * (hLf * hLf - hsL * hsL);
    f[4 * d + 2] = fl.n + P.half_g * (hRf * hRf - hsR * hsR);
    f[4 * d + 3] = fl.t;
  }
  for (int q = 0; q < 8; ++q) w.s[S_FXM + q][k] = f[q];
}

// pass 3: outgoing flux and positivity factor (swe.mjs:507-529)
__host__ __device__ inline void adj_theta_at(const Params &P, const StateP &S, double dt, const Work &w, int i, int j) {
  const int k = j * P.W + i;
  double th = 1.0, out = 0.0;
  const bool in = dt > 0 && i >= 1 && i <= P.W - 2 && j >= 1 && j <= P.H - 2;
  const double *fxM = w.s[S_FXM], *fyM = w.s[S_FYM];
  const double fe = fxM[k], fw = fxM[in ? k - 1 : k], fn = fyM[k], fs = fyM[in ? k - P.W : k];
  const double avail = S.h[k];
  if (in) {
    if (fe > 0) out += fe / P.geo.dx[j];
    if (fw < 0) out -= fw / P.geo.dx[j];
    if (fn > 0) out += fn / P.geo.dyN[j];
    if (fs < 0) out -= fs / P.geo.dyS[j];
    if (out * dt > avail) th = avail > 0 ? avail / (out * dt) : 0.0;
  }
  w.s[S_TH][k] = th;
  w.s[S_OUT][k] = out;
}

// The positivity-factor adjoint of one face: d/dtheta of its contributions
// (the face's flux terms scaled by theta over the row divisors), given the
// adjoint lR of the residual.  x-face owned by cell kf in row j:
__host__ __device__ inline double face_lth_x(const Params &P, const StateP &lR, const Work &w, int kf, int j) {
  const int kL = kf, kR = kf + 1;
  const double M = w.s[S_FXM][kf], NL = w.s[S_FXNL][kf], NR = w.s[S_FXNR][kf], T = w.s[S_FXT][kf];
  const double dh = lR.h[kR] - lR.h[kL], dv = lR.hv[kR] - lR.hv[kL];
  return (M * dh + NR * lR.hu[kR] - NL * lR.hu[kL] + T * dv) / P.geo.dx[j];
}
// y-face owned by cell kf in row j (between rows j and j + 1)
__host__ __device__ inline double face_lth_y(const Params &P, const StateP &lR, const Work &w, int kf, int j) {
  const int kL = kf, kR = kf + P.W;
  const double M = w.s[S_FYM][kf], NL = w.s[S_FYNL][kf], NR = w.s[S_FYNR][kf], T = w.s[S_FYT][kf];
  return (M * lR.h[kR] + NR * lR.hv[kR] + T * lR.hu[kR]) / P.geo.dyS[j + 1] -
         (M * lR.h[kL] + NL * lR.hv[kL] + T * lR.hu[kL]) / P.geo.dyN[j];
}

// pass 4: accumulation^T per face (swe.mjs:530-555) and positivity^T per cell
// (swe.mjs:507-529).  The adjoints of the flux terms of the cell's east and north
// faces; then the positivity factor's adjoint gathered from the four faces it
// scaled as their upwind cell (the west and south faces' shares recomputed from
// their owners' data) and taken through theta = h / (out dt).
struct AccumOut { double lx[4], ly[4], lh, lout; };   // east face, north face; theta^T
__host__ __device__ inline AccumOut adj_accum_eval(const Params &P, const StateP &S, double dt, const StateP &lR,
                                                   const Work &w, int i, int j) {
  const int k = j * P.W + i;
  double lx[4] = {0, 0, 0, 0}, ly[4] = {0, 0, 0, 0}, lthE = 0.0, lthN = 0.0, lthW = 0.0, lthS = 0.0;
  const double *th = w.s[S_TH];
  // everything loaded before the branches (cell k stands in for a face off the grid)
  const double lhk = lR.h[k], lhuk = lR.hu[k], lhvk = lR.hv[k], thk = th[k];
  const bool xon = xface_on(P, i, j), yon = yface_on(P, i, j);
  const int kx = xon ? k + 1 : k, ky = yon ? k + P.W : k;
  const double xM = w.s[S_FXM][k], xNL = w.s[S_FXNL][k], xNR = w.s[S_FXNR][k], xT = w.s[S_FXT][k];
  const double yM = w.s[S_FYM][k], yNL = w.s[S_FYNL][k], yNR = w.s[S_FYNR][k], yT = w.s[S_FYT][k];
  const double xth = th[kx], xlh = lR.h[kx], xlhu = lR.hu[kx], xlhv = lR.hv[kx];
  const double yth = th[ky], ylh = lR.h[ky], ylhu = lR.hu[ky], ylhv = lR.hv[ky];
  const double dxj = P.geo.dx[j], dyNj = P.geo.dyN[j], dySj = P.geo.dyS[yon ? j + 1 : j];
  const bool won = i >= 1 && xface_on(P, i - 1, j), son = j >= 1 && yface_on(P, i, j - 1);
  const double fw = w.s[S_FXM][i >= 1 ? k - 1 : k], fs = w.s[S_FYM][j >= 1 ? k - P.W : k];
  const double thc = w.s[S_TH][k], out = w.s[S_OUT][k], h = S.h[k];
  if (xon) {
    const double M = xM, NL = xNL, NR = xNR, T = xT;
    const double t = M >= 0 ? thk : xth;
    const double inv = t / dxj;
    const double dh = xlh - lhk, dv = xlhv - lhvk;
    lx[0] = inv * dh; lx[1] = -inv * lhuk; lx[2] = inv * xlhu; lx[3] = inv * dv;
    lthE = (M * dh + NR * xlhu - NL * lhuk + T * dv) / dxj;
  }
  if (yon) {
    const double M = yM, NL = yNL, NR = yNR, T = yT;
    const double t = M >= 0 ? thk : yth;
    const double iL = t / dyNj, iR = t / dySj;
    ly[0] = ylh * iR - lhk * iL;
    ly[1] = -lhvk * iL;
    ly[2] = ylhv * iR;
    ly[3] = ylhu * iR - lhuk * iL;
    lthN = (M * ylh + NR * ylhv + T * ylhu) / dySj -
           (M * lhk + NL * lhvk + T * lhuk) / dyNj;
  }
  if (won) lthW = face_lth_x(P, lR, w, k - 1, j);
  if (son) lthS = face_lth_y(P, lR, w, k - P.W, j - 1);
  AccumOut r;
  for (int q = 0; q < 4; ++q) { r.lx[q] = lx[q]; r.ly[q] = ly[q]; }
  // a face off the grid has zero flux and zero share
  double lth = 0.0;
  if (xM >= 0) lth += lthE;                    // east face
  if (i >= 1 && fw < 0) lth += lthW;           // west face
  if (yM >= 0) lth += lthN;                    // north face
  if (j >= 1 && fs < 0) lth += lthS;           // south face
  r.lh = 0.0;
  r.lout = 0.0;
  if (thc < 1.0 && h > 0 && out * dt > h) {
    r.lh = lth / (out * dt);
    r.lout = -lth * thc / out;
  }
  return r;
}
__host__ __device__ inline void adj_accum_at(const Params &P, const StateP &S, double dt, const StateP &lR,
                                             const Work &w, int i, int j) {
  const int k = j * P.W + i;
  const AccumOut a = adj_accum_eval(P, S, dt, lR, w, i, j);
  w.s[S_LXM][k] = a.lx[0]; w.s[S_LXNL][k] = a.lx[1]; w.s[S_LXNR][k] = a.lx[2]; w.s[S_LXT][k] = a.lx[3];
  w.s[S_LYM][k] = a.ly[0]; w.s[S_LYNL][k] = a.ly[1]; w.s[S_LYNR][k] = a.ly[2]; w.s[S_LYT][k] = a.ly[3];
  w.s[S_LHTH][k] = a.lh;
  w.s[S_LOUT][k] = a.lout;
}

// pass 5: face^T: the adjoints of both sides' reconstructed states of the
// east x-face and north y-face of cell (i, j), given pass 4's flux-term adjoints
// of the two faces (lf: M, N + pL, N + pR, T of the x-face, then of the y-face)
// and theta^T's lout of the cell and of its east and north neighbours (read
// only for a face on the grid); r: eta, u, v of the x-face's L and R sides, then
// the y-face's
__host__ __device__ inline void adj_faces_rev_eval(const Params &P, const double *b, const Work &w, int i, int j,
                                                   const double (&lf)[8], double loutK, double loutE, double loutN,
                                                   double (&r)[12]) {
  const int k = j * P.W + i;
  const HllcConst HC = {P.g, P.half_g, P.rg, P.min_depth, P.g * P.min_depth >= SQRT_N_MIN};
  for (int q = 0; q < 12; ++q) r[q] = 0.0;
  for (int d = 0; d < 2; ++d) {
    // everything loaded before the branch (cell k stands in for a face off the grid)
    const bool on = d == 0 ? xface_on(P, i, j) : yface_on(P, i, j);
    const int kR = on ? (d == 0 ? k + 1 : k + P.W) : k;
    // (dyS of row j + 1 is read only for a face on the grid, j + 1 <= H - 2)
    const double divL = d == 0 ? row_uniform(P.geo.dx, j) : row_uniform(P.geo.dyN, j);
    const double divR = d == 0 ? row_uniform(P.geo.dx, j) : row_uniform(P.geo.dyS, j + 1 < P.H ? j + 1 : j);
    const double M = w.s[d == 0 ? S_FXM : S_FYM][k];
    double lM = lf[4 * d];
    const double loutL = loutK, loutR = d == 0 ? loutE : loutN;
    const FaceSide L = face_side(w, b, k, d, 1.0), R = face_side(w, b, kR, d, -1.0);
    const double lNL = lf[4 * d + 1], lNR = lf[4 * d + 2], lT = lf[4 * d + 3];
    if (!on) continue;
    if (M > 0) lM += loutL / divL;
    if (M < 0) lM -= loutR / divR;
    // the solver sees (normal, tangential) = (u, v) on x-faces, (v, u) on y-faces
    const FaceIn f = face_rev(L.eta, L.b, d == 0 ? L.u : L.v, d == 0 ? L.v : L.u, R.eta, R.b, d == 0 ? R.u : R.v,
                              d == 0 ? R.v : R.u, HC, lM, lNL, lNR, lT);
    double *q = r + 6 * d;
    q[0] = f.etaL; q[1] = d == 0 ? f.uL : f.vL; q[2] = d == 0 ? f.vL : f.uL;
    q[3] = f.etaR; q[4] = d == 0 ? f.uR : f.vR; q[5] = d == 0 ? f.vR : f.uR;
  }
}
// pass 5 reading pass 4's slots
__host__ __device__ inline void adj_faces_rev_at(const Params &P, const double *b, const Work &w, int i, int j) {
  const int k = j * P.W + i;
  const double *lout = w.s[S_LOUT];
  double lf[8], r[12];
  for (int q = 0; q < 8; ++q) lf[q] = w.s[S_LXM + q][k];
  const int kE = xface_on(P, i, j) ? k + 1 : k, kN = yface_on(P, i, j) ? k + P.W : k;
  adj_faces_rev_eval(P, b, w, i, j, lf, lout[k], lout[kE], lout[kN], r);
  for (int q = 0; q < 12; ++q) w.s[S_LXETAL + q][k] = r[q];
}

// pass 6: per cell, gather the face-state adjoints into eta, U, V and the
// slopes; the bed (and spherical geometric) term^T and Coriolis^T
// (swe.mjs:570-628)
struct CellRev { double leta, lU, lV, lsE[2], lsU[2], lsV[2], lhuC, lhvC; };
__host__ __device__ inline CellRev cells_rev(const Params &P, const StateP &S, const double *b, const StateP &lR,
                                             const Work &w, const double *tcs, int i, int j) {
  const int W = P.W, k = j * W + i;
  double leta = 0, lU = 0, lV = 0, lsE[2] = {0, 0}, lsU[2] = {0, 0}, lsV[2] = {0, 0};
  for (int d = 0; d < 2; ++d) {
    const int base = d == 0 ? S_LXETAL : S_LYETAL;
    const bool has_m = d == 0 ? i >= 1 : j >= 1;
    // loaded before the branch (cell k stands in off the grid)
    const int km = has_m ? (d == 0 ? k - 1 : k - W) : k;
    const double pe = w.s[base][k], pu = w.s[base + 1][k], pv = w.s[base + 2][k];
    const double me = w.s[base + 3][km], mu = w.s[base + 4][km], mv = w.s[base + 5][km];
    // + side of this cell: the face it owns; - side: the face its lower neighbour owns
    leta += pe; lU += pu; lV += pv;
    lsE[d] += 0.5 * pe; lsU[d] += 0.5 * pu; lsV[d] += 0.5 * pv;
    if (has_m) {
      leta += me; lU += mu; lV += mv;
      lsE[d] -= 0.5 * me; lsU[d] -= 0.5 * mu; lsV[d] -= 0.5 * mv;
    }
  }
  double lhuC = 0.0, lhvC = 0.0;
  const double lrhu = lR.hu[k], lrhv = lR.hv[k];
  const double eta = w.s[S_ETA][k], bk = b[k];
  if (i >= 2 && i <= W - 3 && j >= 2 && j <= P.H - 3) {
    if (P.order >= 2 || P.sphere) {
      for (int d = 0; d < 2; ++d) {
        const double sE = w.s[S_SEX + d][k], sB = w.s[S_SBX + d][k];
        const double aP = eta + 0.5 * sE - (bk + 0.5 * sB), aM = eta - 0.5 * sE - (bk - 0.5 * sB);
        const double div = d == 0 ? row_uniform(P.geo.dx, j) : row_uniform(P.geo.bedPhi, j);
        const double lt = d == 0 ? lrhu : lrhv;
        // term = (-G/2) (hP + hM) sB / div; on a sphere rhv also gets gc G (hP^2 + hM^2)
        double lhP = lt * P.neg_half_g * sB / div, lhM = lhP;
        if (d == 1 && P.sphere) {
          lhP += lrhv * row_uniform(P.geo.geo, j) * P.g * 2.0 * jmax0(aP);
          lhM += lrhv * row_uniform(P.geo.geo, j) * P.g * 2.0 * jmax0(aM);
        }
        if (aP > 0) { leta += lhP; lsE[d] += 0.5 * lhP; }
        if (aM > 0) { leta += lhM; lsE[d] -= 0.5 * lhM; }
      }
    }
    if (P.sphere) {
      const double tR = row_uniform(P.geo.tanPhi, j) / P.R;
      const double fe = row_uniform(P.geo.fRow, j) + w.s[S_U][k] * tR;
      lhvC = fe * lrhu; lhuC = -fe * lrhv;
      lU += tR * (S.hv[k] * lrhu - S.hu[k] * lrhv);
    } else if (P.coriolis_on) {
      lhvC = P.coriolis * lrhu; lhuC = -P.coriolis * lrhv;
    }
    // equilibrium tide, linear in h: rhu += h X sin(theta), rhv += h Y cos(theta);
    // tcs holds (cos, sin)(omega t_s + phase) of the stage; the row factors and
    // tcs are the same on every lane of a device caller's wavefront (row_uniform)
    for (int c = 0; c < P.tide_n; ++c) {
      const double *T = P.tide + (size_t)c * (2 * P.H + 2 * W);
      const double cl = T[2 * P.H + i], sl = T[2 * P.H + W + i];
      const double tc = row_uniform(tcs, 2 * c), ts = row_uniform(tcs, 2 * c + 1);
      leta += lrhu * row_uniform(T, j) * (ts * cl + tc * sl) + lrhv * row_uniform(T, P.H + j) * (tc * cl - ts * sl);
    }
  }
  CellRev r;
  r.leta = leta; r.lU = lU; r.lV = lV;
  for (int d = 0; d < 2; ++d) { r.lsE[d] = lsE[d]; r.lsU[d] = lsU[d]; r.lsV[d] = lsV[d]; }
  r.lhuC = lhuC; r.lhvC = lhvC;
  return r;
}

__host__ __device__ inline void adj_cells_rev_at(const Params &P, const StateP &S, const double *b, const StateP &lR,
                                                 const Work &w, const double *tcs, int i, int j) {
  const int k = j * P.W + i;
  const CellRev c = cells_rev(P, S, b, lR, w, tcs, i, j);
  w.s[S_LETA][k] = c.leta; w.s[S_LU][k] = c.lU; w.s[S_LV][k] = c.lV;
  w.s[S_LSEX][k] = c.lsE[0]; w.s[S_LSEY][k] = c.lsE[1];
  w.s[S_LSUX][k] = c.lsU[0]; w.s[S_LSUY][k] = c.lsU[1];
  w.s[S_LSVX][k] = c.lsV[0]; w.s[S_LSVY][k] = c.lsV[1];
  w.s[S_LHUC][k] = c.lhuC; w.s[S_LHVC][k] = c.lhvC;
}

// The slope adjoints (eta, U, V) along axis d of cell (i, j) alone, in the same
// operations and order as cells_rev, so bit-identical to its lsE[d], lsU[d],
// lsV[d]: the device's tiled gather + limiter^T pass evaluates them for the
// cells on its tiles' edges instead of reading them back from memory.
struct SlopeAdj { double e, u, v; };
__host__ __device__ inline SlopeAdj slope_adj(const Params &P, const double *b, const StateP &lR, const Work &w,
                                              int i, int j, int d) {
  const int W = P.W, k = j * W + i;
  const int base = d == 0 ? S_LXETAL : S_LYETAL;
  const bool has_m = d == 0 ? i >= 1 : j >= 1;
  const int km = has_m ? (d == 0 ? k - 1 : k - W) : k;
  const double pe = w.s[base][k], pu = w.s[base + 1][k], pv = w.s[base + 2][k];
  const double me = w.s[base + 3][km], mu = w.s[base + 4][km], mv = w.s[base + 5][km];
  SlopeAdj r = {0.0, 0.0, 0.0};
  r.e += 0.5 * pe; r.u += 0.5 * pu; r.v += 0.5 * pv;
  if (has_m) { r.e -= 0.5 * me; r.u -= 0.5 * mu; r.v -= 0.5 * mv; }
  if (i >= 2 && i <= W - 3 && j >= 2 && j <= P.H - 3 && (P.order >= 2 || P.sphere)) {
    const double eta = w.s[S_ETA][k], bk = b[k];
    const double sE = w.s[S_SEX + d][k], sB = w.s[S_SBX + d][k];
    const double aP = eta + 0.5 * sE - (bk + 0.5 * sB), aM = eta - 0.5 * sE - (bk - 0.5 * sB);
    const double div = d == 0 ? P.geo.dx[j] : P.geo.bedPhi[j];
    const double lrhv = lR.hv[k];
    const double lt = d == 0 ? lR.hu[k] : lrhv;
    double lhP = lt * P.neg_half_g * sB / div, lhM = lhP;
    if (d == 1 && P.sphere) {
      lhP += lrhv * P.geo.geo[j] * P.g * 2.0 * jmax0(aP);
      lhM += lrhv * P.geo.geo[j] * P.g * 2.0 * jmax0(aM);
    }
    if (aP > 0) r.e += 0.5 * lhP;
    if (aM > 0) r.e -= 0.5 * lhM;
  }
  return r;
}

// pass 7: limiter^T gathered from the slopes of the cell and its two
// neighbours along each axis, then vel^T: adds the adjoint of the stage input
// state at (i, j) into lS.  slope_of(d, o) gives the slope adjoints of cell
// k + o st along d (o = -1, 0, 1; asked only for cells whose slope exists).
template <class F>
__host__ __device__ inline void slopes_rev_core(const Params &P, const StateP &S, const Work &w, const StateW &lS,
                                                int i, int j, double leta, double lU, double lV, double lhth,
                                                double lhuC, double lhvC, const F &slope_of) {
  const int W = P.W, k = j * W + i;
  if (P.order >= 2) {
    const double *eta = w.s[S_ETA], *U = w.s[S_U], *V = w.s[S_V];
    for (int d = 0; d < 2; ++d) {
      const int st = d == 0 ? 1 : W;
      const int pos = d == 0 ? i : j, len = d == 0 ? W : P.H;
      // every value the three slopes below can use, loaded before any branch (a
      // cell off the grid reads cell k instead and is never used): the five cells
      // k + m st, m = -2..2
      double e5[5], u5[5], v5[5], h5[5];
      for (int m = -2; m <= 2; ++m) {
        const int kc = (pos + m < 0 || pos + m >= len) ? k : k + m * st;
        e5[m + 2] = eta[kc]; u5[m + 2] = U[kc]; v5[m + 2] = V[kc]; h5[m + 2] = S.h[kc];
      }
      for (int o = -1; o <= 1; ++o) {             // slope of cell c = k + o st
        const int ci = d == 0 ? i + o : i, cj = d == 0 ? j : j + o;
        if (ci < 1 || ci > W - 2 || cj < 1 || cj > P.H - 2) continue;
        const int m = o + 2;                        // c, c - st, c + st are m, m - 1, m + 1
        if (h5[m] <= P.min_depth || h5[m + 1] <= P.min_depth || h5[m - 1] <= P.min_depth) continue;
        // s_c = lim(q_c - q_{c-st}, q_{c+st} - q_c)
        const LimD dE = limit_deriv(P.limiter, e5[m] - e5[m - 1], e5[m + 1] - e5[m]);
        const LimD dU = limit_deriv(P.limiter, u5[m] - u5[m - 1], u5[m + 1] - u5[m]);
        const LimD dV = limit_deriv(P.limiter, v5[m] - v5[m - 1], v5[m + 1] - v5[m]);
        const double fe = o == 0 ? dE.da - dE.db : (o == 1 ? -dE.da : dE.db);
        const double fu = o == 0 ? dU.da - dU.db : (o == 1 ? -dU.da : dU.db);
        const double fv = o == 0 ? dV.da - dV.db : (o == 1 ? -dV.da : dV.db);
        const SlopeAdj ls = slope_of(d, o);
        leta += ls.e * fe;
        lU += ls.u * fu;
        lV += ls.v * fv;
      }
    }
  }
  const double h = S.h[k];
  const VelD du = vel_deriv(S.hu[k], h, P.eps4), dv = vel_deriv(S.hv[k], h, P.eps4);
  lS.h[k] += leta + lU * du.dh + lV * dv.dh + lhth;
  lS.hu[k] += lU * du.dhq + lhuC;
  lS.hv[k] += lV * dv.dhq + lhvC;
}

// pass 7 reading pass 6's slots (host)
__host__ __device__ inline void adj_slopes_rev_at(const Params &P, const StateP &S, const Work &w, const StateW &lS,
                                                  int i, int j) {
  const int k = j * P.W + i;
  const int W = P.W;
  slopes_rev_core(P, S, w, lS, i, j, w.s[S_LETA][k], w.s[S_LU][k], w.s[S_LV][k], w.s[S_LHTH][k], w.s[S_LHUC][k],
                  w.s[S_LHVC][k], [&](int d, int o) {
                    const int c = k + o * (d == 0 ? 1 : W);
                    SlopeAdj r = {w.s[S_LSEX + d][c], w.s[S_LSUX + d][c], w.s[S_LSVX + d][c]};
                    return r;
                  });
}

// ---- boundary conditions^T (swe.mjs:289-303, :1026-1045, :1119-1124) ---------
// Source of ghost (i, j) (interior coordinates) for the linear kinds.
__host__ __device__ inline void adj_bc_src(int kind, int side, int i, int j, int nx, int ny, int &si, int &sj) {
  if (kind == BC_PERIODIC) {
    si = i; sj = j;
    if (side < 2) si = ((i % nx) + nx) % nx;
    else sj = ((j % ny) + ny) % ny;
    return;
  }
  si = i; sj = j;
  if (kind == BC_REFLECT) {
    if (side == 0) si = -1 - i; else if (side == 1) si = 2 * nx - 1 - i;
    else if (side == 2) sj = -1 - j; else sj = 2 * ny - 1 - j;
  }
  si = si < 0 ? 0 : (si > nx - 1 ? nx - 1 : si);
  sj = sj < 0 ? 0 : (sj > ny - 1 ? ny - 1 : sj);
}

__host__ __device__ inline void adj_add(double *p, double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  atomicAdd(p, v);
#else
  *p += v;
#endif
}

// Flather's ghost (swe.mjs:1084-1116) as a function of its wet source cell:
// with r = sqrt(g / hInt), uInt = vel(q, hInt) of the normal momentum q,
//   Rin = u0 + sgn e0 r,  Rout = uInt - sgn (b + hInt) r,
//   un = (Rin + Rout) / 2,  en = sgn (Rin - Rout) / (2 r),
//   ghost h = max(0, en - b), normal momentum = ghost h * un,
// tangential momentum copied.  Its partial derivatives with respect to the
// source's hInt and q (the clip at 0 takes the subgradient 0).
struct FlatherD { double dh_dh, dh_dq, dq_dh, dq_dq; };
__host__ __device__ inline FlatherD flather_deriv(const Params &P, int side, double hInt, double q, double bed,
                                                  double e0, double u0) {
  const double sgn = (side == 0 || side == 2) ? 1.0 : -1.0;
  const double r = __builtin_sqrt(P.g / hInt);
  const double dr = -0.5 * r / hInt;                  // d r / d hInt
  const double uInt = vel1(q, hInt, P.eps4);
  const VelD du = vel_deriv(q, hInt, P.eps4);
  const double etaInt = bed + hInt;
  const double Rin = u0 + sgn * e0 * r, Rout = uInt - sgn * etaInt * r;
  const double un = 0.5 * (Rin + Rout);
  const double en = sgn * (Rin - Rout) / (2 * r);
  const double hg = jmax0(en - bed);
  const double dRin_dh = sgn * e0 * dr;
  const double dRout_dh = du.dh - sgn * (r + etaInt * dr), dRout_dq = du.dhq;
  const double dun_dh = 0.5 * (dRin_dh + dRout_dh), dun_dq = 0.5 * dRout_dq;
  const double den_dh = sgn * (dRin_dh - dRout_dh) / (2 * r) - en * dr / r;
  const double den_dq = -sgn * dRout_dq / (2 * r);
  const bool wet = en - bed > 0;
  FlatherD d;
  d.dh_dh = wet ? den_dh : 0.0;
  d.dh_dq = wet ? den_dq : 0.0;
  d.dq_dh = d.dh_dh * un + hg * dun_dh;
  d.dq_dq = d.dh_dq * un + hg * dun_dq;
  return d;
}

// Ghost t of one pass: its cell g, its source cell s and the side it is on.
__host__ __device__ inline void adj_bc_cells(const Params &P, int pass, int t, int &g, int &s, int &side) {
  int gi, gj;
  if (pass == 0) {
    const int rr = t / P.W;
    gi = t % P.W; gj = rr < 2 ? rr : P.H - 4 + rr; side = gj < 2 ? 2 : 3;
  } else {
    const int cc = t / P.ny;
    gj = 2 + t % P.ny; gi = cc < 2 ? cc : P.W - 4 + cc; side = gi < 2 ? 0 : 1;
  }
  int si, sj;
  adj_bc_src(P.bc[side], side, gi - 2, gj - 2, P.nx, P.ny, si, sj);
  g = gj * P.W + gi;
  s = (sj + 2) * P.W + (si + 2);
}

// What a ghost holding adjoints (lh, lhu, lhv) on `side` folds into its source
// cell: dh into l.h, dhu into l.hu, dhv into l.hv (any = false: nothing, a
// Flather ghost with a dry source).  For a Flather side the source's state
// (h, hu, hv and bed) is passed in.
struct GhostFold { bool any; double dh, dhu, dhv; };
__host__ __device__ inline GhostFold adj_bc_fold(const Params &P, int side, double lh, double lhu, double lhv,
                                                 double hInt, double qU, double qV, double bed, const double *ext) {
  const int kind = P.bc[side];
  GhostFold f = {true, 0.0, 0.0, 0.0};
  if (kind == BC_FLATHER) {
    // a dry source: the ghost is max(0, eta_ext - b), zero momentum, no state in it
    if (hInt <= P.min_depth) { f.any = false; return f; }
    const bool we = side < 2;
    const FlatherD d = flather_deriv(P, side, hInt, we ? qU : qV, bed, ext[2 * side], ext[2 * side + 1]);
    const double lqn = we ? lhu : lhv, lqt = we ? lhv : lhu;
    f.dh = lh * d.dh_dh + lqn * d.dq_dh;
    const double ln = lh * d.dh_dq + lqn * d.dq_dq;
    f.dhu = we ? ln : lqt;
    f.dhv = we ? lqt : ln;
    return f;
  }
  const bool flipU = kind == BC_REFLECT && side < 2, flipV = kind == BC_REFLECT && side >= 2;
  f.dh = lh;
  f.dhu = flipU ? -lhu : lhu;
  f.dhv = flipV ? -lhv : lhv;
  return f;
}
// the linear kinds' fold (reflect, outflow, periodic): the ghost's adjoints, the
// normal component negated at a wall
__host__ __device__ inline GhostFold adj_bc_fold_linear(const Params &P, int side, double lh, double lhu, double lhv) {
  const int kind = P.bc[side];
  const bool flipU = kind == BC_REFLECT && side < 2, flipV = kind == BC_REFLECT && side >= 2;
  const GhostFold f = {true, lh, flipU ? -lhu : lhu, flipV ? -lhv : lhv};
  return f;
}

// Fold ghost t of one pass into its source and clear it.  applyBC writes the
// W/E columns (every row) and then the S/N rows (every column, corners
// included, a periodic source possibly a W/E ghost); its transpose runs the S/N
// pass first (t < 4 W), then the W/E pass over the interior rows (t < 4 ny):
// the corner rows' W/E ghosts are overwritten by the S/N pass, so they carry
// no adjoint of their own.  A Flather ghost is a nonlinear function of its
// (interior) source cell: S and b are the state and bed applyBC read, ext the
// step's far-field values (eta_ext, u_ext per side); the linear kinds need none.
__host__ __device__ inline void adj_bc_ghost(const Params &P, const StateW &l, int pass, int t, const StateP &S,
                                             const double *b, const double *ext) {
  int g, s, side;
  adj_bc_cells(P, pass, t, g, s, side);
  const double lh = l.h[g], lhu = l.hu[g], lhv = l.hv[g];
  l.h[g] = 0.0; l.hu[g] = 0.0; l.hv[g] = 0.0;
  const bool fl = P.bc[side] == BC_FLATHER;
  const GhostFold f = adj_bc_fold(P, side, lh, lhu, lhv, fl ? S.h[s] : 0.0, fl ? S.hu[s] : 0.0, fl ? S.hv[s] : 0.0,
                                  fl ? b[s] : 0.0, ext);
  if (!f.any) return;
  adj_add(&l.h[s], f.dh);
  adj_add(&l.hu[s], f.dhu);
  adj_add(&l.hv[s], f.dhv);
}

// BC^T in gather form, for the linear kinds (no Flather side) with south and
// north sides that are not periodic and nx, ny >= 4: every source is then an
// interior cell, the sources of one pass's
// work items are disjoint, and each work item adds its ghosts' folds into its
// sources in the order of the loop over t above, so the result is that loop's
// to the bit, with no atomics.  Pass 0, item q < 2 (W - 4): source column
// sc = 2 + q % (W - 4) on the south (q < W - 4) or north side, whose ghosts are
// that column's two ghost cells and, for the first and last interior columns,
// the corner columns that clamp into them.  Pass 1, item q < 2 ny: interior row
// 2 + q % ny on the west or east side, whose ghosts are that row's two.  An item
// reads all its ghosts and sources before it writes any of them.
template <int N>
__host__ __device__ inline void adj_bc_gather_item(const Params &P, const StateW &l, int pass, const int (&ts)[N]) {
  int g[N], s[N], side[N];
  double v[N][3];
#pragma unroll
  for (int m = 0; m < N; ++m) {
    adj_bc_cells(P, pass, ts[m], g[m], s[m], side[m]);
    v[m][0] = l.h[g[m]]; v[m][1] = l.hu[g[m]]; v[m][2] = l.hv[g[m]];
  }
  // the item's (at most two) distinct sources, in order of first appearance
  const int s0 = s[0];
  int s1 = -1;
#pragma unroll
  for (int m = 1; m < N; ++m)
    if (s1 < 0 && s[m] != s0) s1 = s[m];
  double a0h = l.h[s0], a0u = l.hu[s0], a0v = l.hv[s0];
  double a1h = 0.0, a1u = 0.0, a1v = 0.0;
  if (s1 >= 0) { a1h = l.h[s1]; a1u = l.hu[s1]; a1v = l.hv[s1]; }
#pragma unroll
  for (int m = 0; m < N; ++m) { l.h[g[m]] = 0.0; l.hu[g[m]] = 0.0; l.hv[g[m]] = 0.0; }
#pragma unroll
  for (int m = 0; m < N; ++m) {
    const GhostFold f = adj_bc_fold_linear(P, side[m], v[m][0], v[m][1], v[m][2]);
    if (s[m] == s0) { a0h += f.dh; a0u += f.dhu; a0v += f.dhv; }
    else { a1h += f.dh; a1u += f.dhu; a1v += f.dhv; }
  }
  l.h[s0] = a0h; l.hu[s0] = a0u; l.hv[s0] = a0v;
  if (s1 >= 0) { l.h[s1] = a1h; l.hu[s1] = a1u; l.hv[s1] = a1v; }
}
__host__ __device__ inline void adj_bc_gather(const Params &P, const StateW &l, int pass, int q) {
  if (pass == 0) {
    const int nsc = P.W - 4, r0 = 2 * (q / nsc), sc = 2 + q % nsc;
    if (sc == 2) {                   // the first interior column: corner columns 0 and 1 clamp into it
      const int ts[6] = {r0 * P.W, r0 * P.W + 1, r0 * P.W + 2, (r0 + 1) * P.W, (r0 + 1) * P.W + 1, (r0 + 1) * P.W + 2};
      adj_bc_gather_item<6>(P, l, 0, ts);
    } else if (sc == P.W - 3) {      // the last: corner columns W - 2 and W - 1
      const int c = P.W - 3;
      const int ts[6] = {r0 * P.W + c, r0 * P.W + c + 1, r0 * P.W + c + 2,
                         (r0 + 1) * P.W + c, (r0 + 1) * P.W + c + 1, (r0 + 1) * P.W + c + 2};
      adj_bc_gather_item<6>(P, l, 0, ts);
    } else {
      const int ts[2] = {r0 * P.W + sc, (r0 + 1) * P.W + sc};
      adj_bc_gather_item<2>(P, l, 0, ts);
    }
  } else {
    const int c0 = 2 * (q / P.ny), r = q % P.ny;
    const int ts[2] = {c0 * P.ny + r, (c0 + 1) * P.ny + r};
    adj_bc_gather_item<2>(P, l, 1, ts);
  }
}
__host__ __device__ inline bool adj_bc_gather_ok(const Params &P) {
  for (int k = 0; k < 4; ++k)
    if (P.bc[k] == BC_FLATHER) return false;
  return P.bc[2] != BC_PERIODIC && P.bc[3] != BC_PERIODIC && P.nx >= 4 && P.ny >= 4;
}
// BC^T folded on read, for the configurations of adj_bc_gather: adds to (h, hu,
// hv), the adjoint of interior cell (i, j) held in l, the folds of the ghosts of
// l whose source it is, in the order of the ghost-by-ghost loop (the
// south/north pass's ghosts by row, then column; then the west/east pass's by
// column), so the sum is adj_bc_gather's to the bit.  The ghosts are read, not
// cleared.  Only cells within two of an edge of the interior have any.
__host__ __device__ inline void adj_bc_fold_into(const Params &P, const StateP &l, int i, int j, double &h,
                                                 double &hu, double &hv) {
  const int W = P.W, H = P.H, k = j * W + i;
  if (j <= 3 || j >= H - 4) {                  // south/north pass: ghost rows 0, 1 or H - 2, H - 1
    const int r0 = j <= 3 ? 0 : 2;
    // the ghost columns that can have (i, j) as source, in ascending order: its
    // own and, at the first and last interior column, the corner columns
    const int c0 = i == 2 ? 0 : i, nc = (i == 2 || i == W - 3) ? 3 : 1;
    for (int rr = r0; rr < r0 + 2; ++rr)
      for (int c = 0; c < nc; ++c) {
        int g, s, side;
        adj_bc_cells(P, 0, rr * W + c0 + c, g, s, side);
        if (s != k) continue;
        const GhostFold f = adj_bc_fold_linear(P, side, l.h[g], l.hu[g], l.hv[g]);
        h += f.dh; hu += f.dhu; hv += f.dhv;
      }
  }
  if (i <= 3 || i >= W - 4) {                  // west/east pass: ghost columns 0, 1, W - 2, W - 1 of row j
    for (int cc = 0; cc < 4; ++cc) {
      int g, s, side;
      adj_bc_cells(P, 1, cc * P.ny + (j - 2), g, s, side);
      if (s != k) continue;
      const GhostFold f = adj_bc_fold_linear(P, side, l.h[g], l.hu[g], l.hv[g]);
      h += f.dh; hu += f.dhu; hv += f.dhv;
    }
  }
}

// ---- dryClean^T, friction^T and sponge^T (swe.mjs:659-704, :734-814, :1188-1197) ----
// Reverse of a stage's pointwise part at cell k, given the adjoint lOut of its
// output (interior) and the stage's dryCleaned (pre-friction) state `dc`:
// sponge^T (stage 2: the relaxation h += (target - h) f, q -= q f scales every
// field by 1 - f, f = min(1, strength dt 60)), friction^T (stage 2, accumulating
// the Manning adjoint into lN), dryClean^T, then the stage combination^T:
//   stage 2: lA = 0.5 lC0, lB = 0.5 lC0, lR = 0.5 dt lC0
//   stage 1: lA += lB0,    lR = dt lB0
// dryClean floored the cells whose depth it left at exactly 0 (where the
// combination gave 0 exactly the subgradient 0 is taken), and zeroed the
// momentum of the cells at or below minDepth.  Ghost cells carry no adjoint.
__host__ __device__ inline void adj_epilogue_rev_at(const Params &P, const StateP &dc, const StateP &lOut,
                                                    const double *nfield, const double *sponge, double dt,
                                                    double *lN, int stage, const StateW &lA, const StateW &lB,
                                                    const StateW &lR, int i, int j, bool fold = false) {
  const int k = j * P.W + i;
  double lh = 0.0, lhu = 0.0, lhv = 0.0;
  if (i >= 2 && i <= P.W - 3 && j >= 2 && j <= P.H - 3) {
    const double h = dc.h[k], hu = dc.hu[k], hv = dc.hv[k];
    const bool floored = h == 0.0;
    const bool dry = h <= P.min_depth;
    lh = lOut.h[k]; lhu = lOut.hu[k]; lhv = lOut.hv[k];
    // fold: lOut still holds its ghosts' adjoints (BC^T of the stage output,
    // adj_bc_fold_into)
    if (fold) adj_bc_fold_into(P, lOut, i, j, lh, lhu, lhv);
    if (stage == 2 && P.sponge && sponge) {
      const double str = sponge[k];
      if (str > 0) {
        const double keep = 1.0 - jmin(1.0, str * dt * 60);
        lh *= keep; lhu *= keep; lhv *= keep;
      }
    }
    if (stage == 2 && P.friction && !dry) {
      const double n = P.friction == 2 ? nfield[k] : P.manning;
      if (n != 0.0 && n == n) {
        const Vel2 uv = vel2(hu, hv, h, P.eps4);
        const double speed = js_hypot(uv.u, uv.v);
        if (!(speed < 1e-14)) {
          // out = q / (1 + c n^2 speed / h^(4/3)) for q = hu, hv
          const double qf = dt * P.g * n * n * speed / js_pow_pos(h, 4.0 / 3.0);
          const double den = 1 + qf;
          const double lden = -(lhu * hu + lhv * hv) / (den * den);
          const double lspeed = lden * qf / speed;
          if (lN) lN[k] += lden * 2.0 * qf / n;
          const double lu = lspeed * uv.u / speed, lv = lspeed * uv.v / speed;
          const VelD du = vel_deriv(hu, h, P.eps4), dv = vel_deriv(hv, h, P.eps4);
          lh += lden * qf * (-4.0 / 3.0) / h + lu * du.dh + lv * dv.dh;
          lhu = lhu / den + lu * du.dhq;
          lhv = lhv / den + lv * dv.dhq;
        }
      }
    }
    if (floored) lh = 0.0;
    if (dry) { lhu = 0.0; lhv = 0.0; }
  }
  if (stage == 2) {
    lA.h[k] = 0.5 * lh; lA.hu[k] = 0.5 * lhu; lA.hv[k] = 0.5 * lhv;
    lB.h[k] = 0.5 * lh; lB.hu[k] = 0.5 * lhu; lB.hv[k] = 0.5 * lhv;
    lR.h[k] = 0.5 * dt * lh; lR.hu[k] = 0.5 * dt * lhu; lR.hv[k] = 0.5 * dt * lhv;
  } else {
    lA.h[k] += lh; lA.hu[k] += lhu; lA.hv[k] += lhv;
    lR.h[k] = dt * lh; lR.hu[k] = dt * lhu; lR.hv[k] = dt * lhv;
  }
}

}  // namespace swe
