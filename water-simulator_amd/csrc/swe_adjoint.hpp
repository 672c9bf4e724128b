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
  const double dD = (4.0 * h * h * h * (self ? 2.0 : 1.0)) / (2.0 * D);
  VelD r;
  r.dhq = SQRT2 * h / D;
  r.dh = SQRT2 * hq / D - SQRT2 * h * hq * dD / (D * D);
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

// Reverse of hllc(hL, huL, hvL, hR, huR, hvR) (swe.mjs:974-1011).
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
  const double cL = __builtin_sqrt(G * jmax(hL, 0.0)), cR = __builtin_sqrt(G * jmax(hR, 0.0));
  double sL, sR, a = 0, hStar = 0, qL = 1, qR = 1;
  if (!wetL) { sL = uR - 2 * cR; sR = uR + cR; }
  else if (!wetR) { sL = uL - cL; sR = uL + 2 * cL; }
  else {
    a = (cL + cR) * 0.5 + (uL - uR) * 0.25;
    hStar = a * a / G;
    qL = hStar > hL ? __builtin_sqrt(0.5 * (hStar + hL) * hStar / (hL * hL)) : 1.0;
    qR = hStar > hR ? __builtin_sqrt(0.5 * (hStar + hR) * hStar / (hR * hR)) : 1.0;
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
      double lhStar = 0.0;
      if (hStar > hL) {
        const double X = 0.5 * (hStar + hL) * hStar / (hL * hL);
        const double lX = lqL / (2.0 * qL);
        lhStar += lX * 0.5 * (2.0 * hStar + hL) / (hL * hL);
        g.hL += lX * (0.5 * hStar / (hL * hL) - 2.0 * X / hL);
      }
      if (hStar > hR) {
        const double X = 0.5 * (hStar + hR) * hStar / (hR * hR);
        const double lX = lqR / (2.0 * qR);
        lhStar += lX * 0.5 * (2.0 * hStar + hR) / (hR * hR);
        g.hR += lX * (0.5 * hStar / (hR * hR) - 2.0 * X / hR);
      }
      const double la = lhStar * 2.0 * a / G;
      lcL += 0.5 * la; lcR += 0.5 * la; luL += 0.25 * la; luR -= 0.25 * la;
    }
  }
  if (wetL) {
    g.hL += lcL * G / (2.0 * cL);
    g.huL += luL / hL; g.hvL += lvL / hL;
    g.hL -= (luL * uL + lvL * vL) / hL;
  }
  if (wetR) {
    g.hR += lcR * G / (2.0 * cR);
    g.huR += luR / hR; g.hvR += lvR / hR;
    g.hR -= (luR * uR + lvR * vR) / hR;
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
  // adjoints: face flux terms and positivity factor per face, theta^T per cell
  S_LXM, S_LXNL, S_LXNR, S_LXT, S_LYM, S_LYNL, S_LYNR, S_LYT, S_LXTH, S_LYTH, S_LOUT, S_LHTH,
  // adjoints of the reconstructed face states (grid frame: eta, u, v; L and R)
  S_LXETAL, S_LXUL, S_LXVL, S_LXETAR, S_LXUR, S_LXVR,
  S_LYETAL, S_LYUL, S_LYVL, S_LYETAR, S_LYUR, S_LYVR,
  // adjoints per cell: eta, U, V, slopes, Coriolis into hu / hv
  S_LETA, S_LU, S_LV, S_LSEX, S_LSEY, S_LSUX, S_LSUY, S_LSVX, S_LSVY, S_LHUC, S_LHVC,
  S_COUNT
};
struct Work { double *s[S_COUNT]; };

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
  if (P.order >= 2 && i >= 1 && i <= P.W - 2 && j >= 1 && j <= P.H - 2) {
    for (int d = 0; d < 2; ++d) {
      const int kp = d == 0 ? k + 1 : k + P.W, km = d == 0 ? k - 1 : k - P.W;
      if (h <= P.min_depth || S.h[kp] <= P.min_depth || S.h[km] <= P.min_depth) continue;
      const double eC = b[k] + h, eP = b[kp] + S.h[kp], eM = b[km] + S.h[km];
      const Vel2 up = vel2(S.hu[kp], S.hv[kp], S.h[kp], P.eps4), um = vel2(S.hu[km], S.hv[km], S.h[km], P.eps4);
      sl[0 + d] = limit(P.limiter, eC - eM, eP - eC);
      sl[6 + d] = limit(P.limiter, b[k] - b[km], b[kp] - b[k]);
      sl[2 + d] = limit(P.limiter, uv.u - um.u, up.u - uv.u);
      sl[4 + d] = limit(P.limiter, uv.v - um.v, up.v - uv.v);
    }
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
  const HllcConst HC = {P.g, P.half_g, P.rg, P.min_depth};
  double f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int d = 0; d < 2; ++d) {
    if (!(d == 0 ? xface_on(P, i, j) : yface_on(P, i, j))) continue;
    const FaceSide L = face_side(w, b, k, d, 1.0), R = face_side(w, b, d == 0 ? k + 1 : k + P.W, d, -1.0);
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
  if (dt > 0 && i >= 1 && i <= P.W - 2 && j >= 1 && j <= P.H - 2) {
    const double *fxM = w.s[S_FXM], *fyM = w.s[S_FYM];
    if (fxM[k] > 0) out += fxM[k] / P.geo.dx[j];
    if (fxM[k - 1] < 0) out -= fxM[k - 1] / P.geo.dx[j];
    if (fyM[k] > 0) out += fyM[k] / P.geo.dyN[j];
    if (fyM[k - P.W] < 0) out -= fyM[k - P.W] / P.geo.dyS[j];
    const double avail = S.h[k];
    if (out * dt > avail) th = avail > 0 ? avail / (out * dt) : 0.0;
  }
  w.s[S_TH][k] = th;
  w.s[S_OUT][k] = out;
}

// pass 4: accumulation^T per face (swe.mjs:530-555): the adjoints of its flux
// terms and of the positivity factor that scaled it
__host__ __device__ inline void adj_accum_at(const Params &P, const StateP &lR, const Work &w, int i, int j) {
  const int k = j * P.W + i;
  double lx[5] = {0, 0, 0, 0, 0}, ly[5] = {0, 0, 0, 0, 0};
  const double *th = w.s[S_TH];
  if (xface_on(P, i, j)) {
    const int kL = k, kR = k + 1;
    const double M = w.s[S_FXM][k], NL = w.s[S_FXNL][k], NR = w.s[S_FXNR][k], T = w.s[S_FXT][k];
    const double t = M >= 0 ? th[kL] : th[kR];
    const double inv = t / P.geo.dx[j];
    const double dh = lR.h[kR] - lR.h[kL], dv = lR.hv[kR] - lR.hv[kL];
    lx[0] = inv * dh; lx[1] = -inv * lR.hu[kL]; lx[2] = inv * lR.hu[kR]; lx[3] = inv * dv;
    lx[4] = (M * dh + NR * lR.hu[kR] - NL * lR.hu[kL] + T * dv) / P.geo.dx[j];
  }
  if (yface_on(P, i, j)) {
    const int kL = k, kR = k + P.W;
    const double M = w.s[S_FYM][k], NL = w.s[S_FYNL][k], NR = w.s[S_FYNR][k], T = w.s[S_FYT][k];
    const double t = M >= 0 ? th[kL] : th[kR];
    const double iL = t / P.geo.dyN[j], iR = t / P.geo.dyS[j + 1];
    ly[0] = lR.h[kR] * iR - lR.h[kL] * iL;
    ly[1] = -lR.hv[kL] * iL;
    ly[2] = lR.hv[kR] * iR;
    ly[3] = lR.hu[kR] * iR - lR.hu[kL] * iL;
    ly[4] = (M * lR.h[kR] + NR * lR.hv[kR] + T * lR.hu[kR]) / P.geo.dyS[j + 1] -
            (M * lR.h[kL] + NL * lR.hv[kL] + T * lR.hu[kL]) / P.geo.dyN[j];
  }
  w.s[S_LXM][k] = lx[0]; w.s[S_LXNL][k] = lx[1]; w.s[S_LXNR][k] = lx[2]; w.s[S_LXT][k] = lx[3]; w.s[S_LXTH][k] = lx[4];
  w.s[S_LYM][k] = ly[0]; w.s[S_LYNL][k] = ly[1]; w.s[S_LYNR][k] = ly[2]; w.s[S_LYT][k] = ly[3]; w.s[S_LYTH][k] = ly[4];
}

// pass 5: positivity^T per cell: the factor's adjoint gathered from the faces
// it scaled (as their upwind cell), then through theta = h / (out dt)
__host__ __device__ inline void adj_theta_rev_at(const Params &P, const StateP &S, double dt, const Work &w,
                                                 int i, int j) {
  const int k = j * P.W + i;
  double lth = 0.0;
  const double *fxM = w.s[S_FXM], *fyM = w.s[S_FYM];
  if (fxM[k] >= 0) lth += w.s[S_LXTH][k];                      // east face
  if (i >= 1 && fxM[k - 1] < 0) lth += w.s[S_LXTH][k - 1];     // west face
  if (fyM[k] >= 0) lth += w.s[S_LYTH][k];                      // north face
  if (j >= 1 && fyM[k - P.W] < 0) lth += w.s[S_LYTH][k - P.W]; // south face
  double lh = 0.0, lout = 0.0;
  const double th = w.s[S_TH][k], out = w.s[S_OUT][k], h = S.h[k];
  if (th < 1.0 && h > 0 && out * dt > h) {
    lh = lth / (out * dt);
    lout = -lth * th / out;
  }
  w.s[S_LHTH][k] = lh;
  w.s[S_LOUT][k] = lout;
}

// pass 6: face^T: the adjoints of both sides' reconstructed states of the
// east x-face and north y-face of cell (i, j)
__host__ __device__ inline void adj_faces_rev_at(const Params &P, const double *b, const Work &w, int i, int j) {
  const int k = j * P.W + i;
  const HllcConst HC = {P.g, P.half_g, P.rg, P.min_depth};
  const double *lout = w.s[S_LOUT];
  double r[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int d = 0; d < 2; ++d) {
    if (!(d == 0 ? xface_on(P, i, j) : yface_on(P, i, j))) continue;
    const int kR = d == 0 ? k + 1 : k + P.W;
    const double divL = d == 0 ? P.geo.dx[j] : P.geo.dyN[j], divR = d == 0 ? P.geo.dx[j] : P.geo.dyS[j + 1];
    const double M = w.s[d == 0 ? S_FXM : S_FYM][k];
    double lM = w.s[d == 0 ? S_LXM : S_LYM][k];
    if (M > 0) lM += lout[k] / divL;
    if (M < 0) lM -= lout[kR] / divR;
    const FaceSide L = face_side(w, b, k, d, 1.0), R = face_side(w, b, kR, d, -1.0);
    const int o = d == 0 ? S_LXNL : S_LYNL;
    // the solver sees (normal, tangential) = (u, v) on x-faces, (v, u) on y-faces
    const FaceIn f = face_rev(L.eta, L.b, d == 0 ? L.u : L.v, d == 0 ? L.v : L.u, R.eta, R.b, d == 0 ? R.u : R.v,
                              d == 0 ? R.v : R.u, HC, lM, w.s[o][k], w.s[o + 1][k], w.s[o + 2][k]);
    double *q = r + 6 * d;
    q[0] = f.etaL; q[1] = d == 0 ? f.uL : f.vL; q[2] = d == 0 ? f.vL : f.uL;
    q[3] = f.etaR; q[4] = d == 0 ? f.uR : f.vR; q[5] = d == 0 ? f.vR : f.uR;
  }
  for (int q = 0; q < 12; ++q) w.s[S_LXETAL + q][k] = r[q];
}

// pass 7: per cell, gather the face-state adjoints into eta, U, V and the
// slopes; the bed (and spherical geometric) term^T and Coriolis^T
// (swe.mjs:570-628)
__host__ __device__ inline void adj_cells_rev_at(const Params &P, const StateP &S, const double *b, const StateP &lR,
                                                 const Work &w, const double *tcs, int i, int j) {
  const int W = P.W, k = j * W + i;
  double leta = 0, lU = 0, lV = 0, lsE[2] = {0, 0}, lsU[2] = {0, 0}, lsV[2] = {0, 0};
  for (int d = 0; d < 2; ++d) {
    const int base = d == 0 ? S_LXETAL : S_LYETAL;
    const int km = d == 0 ? k - 1 : k - W;
    const bool has_m = d == 0 ? i >= 1 : j >= 1;
    // + side of this cell: the face it owns; - side: the face its lower neighbour owns
    leta += w.s[base][k]; lU += w.s[base + 1][k]; lV += w.s[base + 2][k];
    lsE[d] += 0.5 * w.s[base][k]; lsU[d] += 0.5 * w.s[base + 1][k]; lsV[d] += 0.5 * w.s[base + 2][k];
    if (has_m) {
      leta += w.s[base + 3][km]; lU += w.s[base + 4][km]; lV += w.s[base + 5][km];
      lsE[d] -= 0.5 * w.s[base + 3][km]; lsU[d] -= 0.5 * w.s[base + 4][km]; lsV[d] -= 0.5 * w.s[base + 5][km];
    }
  }
  double lhuC = 0.0, lhvC = 0.0;
  if (i >= 2 && i <= W - 3 && j >= 2 && j <= P.H - 3) {
    const double lrhu = lR.hu[k], lrhv = lR.hv[k];
    if (P.order >= 2 || P.sphere) {
      const double eta = w.s[S_ETA][k], bk = b[k];
      for (int d = 0; d < 2; ++d) {
        const double sE = w.s[S_SEX + d][k], sB = w.s[S_SBX + d][k];
        const double aP = eta + 0.5 * sE - (bk + 0.5 * sB), aM = eta - 0.5 * sE - (bk - 0.5 * sB);
        const double div = d == 0 ? P.geo.dx[j] : P.geo.bedPhi[j];
        const double lt = d == 0 ? lrhu : lrhv;
        // term = (-G/2) (hP + hM) sB / div; on a sphere rhv also gets gc G (hP^2 + hM^2)
        double lhP = lt * P.neg_half_g * sB / div, lhM = lhP;
        if (d == 1 && P.sphere) {
          lhP += lrhv * P.geo.geo[j] * P.g * 2.0 * jmax0(aP);
          lhM += lrhv * P.geo.geo[j] * P.g * 2.0 * jmax0(aM);
        }
        if (aP > 0) { leta += lhP; lsE[d] += 0.5 * lhP; }
        if (aM > 0) { leta += lhM; lsE[d] -= 0.5 * lhM; }
      }
    }
    if (P.sphere) {
      const double tR = P.geo.tanPhi[j] / P.R;
      const double fe = P.geo.fRow[j] + w.s[S_U][k] * tR;
      lhvC = fe * lrhu; lhuC = -fe * lrhv;
      lU += tR * (S.hv[k] * lrhu - S.hu[k] * lrhv);
    } else if (P.coriolis_on) {
      lhvC = P.coriolis * lrhu; lhuC = -P.coriolis * lrhv;
    }
    // equilibrium tide, linear in h: rhu += h X sin(theta), rhv += h Y cos(theta);
    // tcs holds (cos, sin)(omega t_s + phase) of the stage
    for (int c = 0; c < P.tide_n; ++c) {
      const double *T = P.tide + (size_t)c * (2 * P.H + 2 * W);
      const double cl = T[2 * P.H + i], sl = T[2 * P.H + W + i];
      const double tc = tcs[2 * c], ts = tcs[2 * c + 1];
      leta += lrhu * T[j] * (ts * cl + tc * sl) + lrhv * T[P.H + j] * (tc * cl - ts * sl);
    }
  }
  w.s[S_LETA][k] = leta; w.s[S_LU][k] = lU; w.s[S_LV][k] = lV;
  w.s[S_LSEX][k] = lsE[0]; w.s[S_LSEY][k] = lsE[1];
  w.s[S_LSUX][k] = lsU[0]; w.s[S_LSUY][k] = lsU[1];
  w.s[S_LSVX][k] = lsV[0]; w.s[S_LSVY][k] = lsV[1];
  w.s[S_LHUC][k] = lhuC; w.s[S_LHVC][k] = lhvC;
}

// pass 8: limiter^T gathered from the slopes of the cell and its two
// neighbours along each axis, then vel^T: adds the adjoint of the stage input
// state at (i, j) into lS
__host__ __device__ inline void adj_slopes_rev_at(const Params &P, const StateP &S, const Work &w, const StateW &lS,
                                                  int i, int j) {
  const int W = P.W, k = j * W + i;
  double leta = w.s[S_LETA][k], lU = w.s[S_LU][k], lV = w.s[S_LV][k];
  if (P.order >= 2) {
    const double *eta = w.s[S_ETA], *U = w.s[S_U], *V = w.s[S_V];
    for (int d = 0; d < 2; ++d) {
      const int st = d == 0 ? 1 : W;
      for (int o = -1; o <= 1; ++o) {             // slope of cell c = k + o st
        const int c = k + o * st;
        const int ci = d == 0 ? i + o : i, cj = d == 0 ? j : j + o;
        if (ci < 1 || ci > W - 2 || cj < 1 || cj > P.H - 2) continue;
        if (S.h[c] <= P.min_depth || S.h[c + st] <= P.min_depth || S.h[c - st] <= P.min_depth) continue;
        // s_c = lim(q_c - q_{c-st}, q_{c+st} - q_c)
        const LimD dE = limit_deriv(P.limiter, eta[c] - eta[c - st], eta[c + st] - eta[c]);
        const LimD dU = limit_deriv(P.limiter, U[c] - U[c - st], U[c + st] - U[c]);
        const LimD dV = limit_deriv(P.limiter, V[c] - V[c - st], V[c + st] - V[c]);
        const double fe = o == 0 ? dE.da - dE.db : (o == 1 ? -dE.da : dE.db);
        const double fu = o == 0 ? dU.da - dU.db : (o == 1 ? -dU.da : dU.db);
        const double fv = o == 0 ? dV.da - dV.db : (o == 1 ? -dV.da : dV.db);
        leta += w.s[S_LSEX + d][c] * fe;
        lU += w.s[S_LSUX + d][c] * fu;
        lV += w.s[S_LSVX + d][c] * fv;
      }
    }
  }
  const double h = S.h[k];
  const VelD du = vel_deriv(S.hu[k], h, P.eps4), dv = vel_deriv(S.hv[k], h, P.eps4);
  lS.h[k] += leta + lU * du.dh + lV * dv.dh + w.s[S_LHTH][k];
  lS.hu[k] += lU * du.dhq + w.s[S_LHUC][k];
  lS.hv[k] += lV * dv.dhq + w.s[S_LHVC][k];
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

// Fold ghost t of one pass into its source and clear it.  applyBC writes the
// W/E columns (every row) and then the S/N rows (every column, corners
// included, a periodic source possibly a W/E ghost); its transpose runs the S/N
// pass first (t < 4 W), then the W/E pass over the interior rows (t < 4 ny):
// the corner rows' W/E ghosts are overwritten by the S/N pass, so they carry
// no adjoint of their own.
__host__ __device__ inline void adj_bc_ghost(const Params &P, const StateW &l, int pass, int t) {
  int gi, gj, side;
  if (pass == 0) {
    const int rr = t / P.W;
    gi = t % P.W; gj = rr < 2 ? rr : P.H - 4 + rr; side = gj < 2 ? 2 : 3;
  } else {
    const int cc = t / P.ny;
    gj = 2 + t % P.ny; gi = cc < 2 ? cc : P.W - 4 + cc; side = gi < 2 ? 0 : 1;
  }
  const int kind = P.bc[side];
  int si, sj;
  adj_bc_src(kind, side, gi - 2, gj - 2, P.nx, P.ny, si, sj);
  const int g = gj * P.W + gi, s = (sj + 2) * P.W + (si + 2);
  const bool flipU = kind == BC_REFLECT && side < 2, flipV = kind == BC_REFLECT && side >= 2;
  const double lh = l.h[g], lhu = l.hu[g], lhv = l.hv[g];
  l.h[g] = 0.0; l.hu[g] = 0.0; l.hv[g] = 0.0;
  adj_add(&l.h[s], lh);
  adj_add(&l.hu[s], flipU ? -lhu : lhu);
  adj_add(&l.hv[s], flipV ? -lhv : lhv);
}

// ---- dryClean^T and friction^T (swe.mjs:659-704, :734-814) -------------------------
// Reverse of a stage's pointwise part at cell k, given the adjoint lOut of its
// output (interior) and the stage's dryCleaned (pre-friction) state `dc`:
// friction^T (stage 2, accumulating the Manning adjoint into lN), dryClean^T,
// then the stage combination^T:
//   stage 2: lA = 0.5 lC0, lB = 0.5 lC0, lR = 0.5 dt lC0
//   stage 1: lA += lB0,    lR = dt lB0
// dryClean floored the cells whose depth it left at exactly 0 (where the
// combination gave 0 exactly the subgradient 0 is taken), and zeroed the
// momentum of the cells at or below minDepth.  Ghost cells carry no adjoint.
__host__ __device__ inline void adj_epilogue_rev_at(const Params &P, const StateP &dc, const StateP &lOut,
                                                    const double *nfield, double dt, double *lN, int stage,
                                                    const StateW &lA, const StateW &lB, const StateW &lR,
                                                    int i, int j) {
  const int k = j * P.W + i;
  double lh = 0.0, lhu = 0.0, lhv = 0.0;
  if (i >= 2 && i <= P.W - 3 && j >= 2 && j <= P.H - 3) {
    const double h = dc.h[k], hu = dc.hu[k], hv = dc.hv[k];
    const bool floored = h == 0.0;
    const bool dry = h <= P.min_depth;
    lh = lOut.h[k]; lhu = lOut.hu[k]; lhv = lOut.hv[k];
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
