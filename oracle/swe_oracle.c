/*
 * CPU oracle -- test infrastructure only (tests/, smoke(), bench.py cpu_baseline).
 * Sequential restatement of the reference solver src/swe.mjs; see swe_oracle.h.
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math (oracle/Makefile).
 *
 * Parity status per routine (what pins it):
 *   orc_vel, orc_limit, orc_hllc, orc_apply_bc, orc_residual, orc_max_dt,
 *   orc_dry_clean, orc_step, orc_sponge  -- bit-exact against golden vectors
 *       written by the reference (tests/golden/make_golden.mjs) plus closed-form
 *       checks (tests/test_oracle_*.py).
 *   orc_apply_friction -- golden vectors to 1e-13 relative: Math.pow is not
 *       bit-reproduced (see orc_pow); Math.hypot is.
 *   orc_volume/energy/max_speed/max_surface_deviation/finite -- golden values.
 */
#include "swe_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orc_sim {
  orc_params p;
  int W, H, ng, N;
  double t; int64_t steps; double mass_floored;
  int bc[4];
  double ext[8];
  double *f[64];          /* fields by id (N or 2N), NULL where unused */
  double *row[ORC_NROWS];
};

/* ---- JS Math semantics ------------------------------------------------- */
/* Math.max(0, x): NaN propagates, -0 becomes +0. */
static inline double jmax0(double x) { return (x > 0.0 || x != x) ? x : 0.0; }
/* Math.max(a, b) for two doubles */
static inline double jmax(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a > b) return a;
  if (b > a) return b;
  if (a == 0.0) return (signbit(a) && signbit(b)) ? -0.0 : 0.0;
  return a;
}
static inline double jmin(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a < b) return a;
  if (b < a) return b;
  if (a == 0.0) return (signbit(a) || signbit(b)) ? -0.0 : 0.0;
  return a;
}
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

/* ---- Math.hypot (2 args): V8 normalise-by-max + Kahan sum --------------- */
double orc_hypot(double a, double b) {
  if (a != a || b != b) {
    if (isinf(a) || isinf(b)) return INFINITY;
    return NAN;
  }
  double x = fabs(a), y = fabs(b);
  double m = 0.0;
  if (x > m) m = x;
  if (y > m) m = y;
  if (m == INFINITY) return INFINITY;
  if (m == 0.0) return 0.0;
  double sum = 0.0, comp = 0.0;
  double v[2] = {x, y};
  for (int i = 0; i < 2; i++) {
    double n = v[i] / m;
    double summand = n * n - comp;
    double pre = sum + summand;
    comp = (pre - sum) - summand;
    sum = pre;
  }
  return sqrt(sum) * m;
}

/* ---- Math.pow restated from fdlibm e_pow.c, positive finite x ---------- */
static inline uint32_t hiw(double x) { uint64_t u; memcpy(&u, &x, 8); return (uint32_t)(u >> 32); }
static inline uint32_t low(double x) { uint64_t u; memcpy(&u, &x, 8); return (uint32_t)u; }
static inline double mkd(uint32_t hi, uint32_t lo) { uint64_t u = ((uint64_t)hi << 32) | lo; double d; memcpy(&d, &u, 8); return d; }
static inline double zlo(double x) { return mkd(hiw(x), 0); }

double orc_pow(double x, double y) {
  static const double bp[2] = {1.0, 1.5}, dp_h[2] = {0.0, 5.84962487220764160156e-01},
      dp_l[2] = {0.0, 1.35003920212974897128e-08}, two53 = 9007199254740992.0,
      L1 = 5.99999999999994648725e-01, L2 = 4.28571428578550184252e-01,
      L3 = 3.33333329818377432918e-01, L4 = 2.72728123808534006489e-01,
      L5 = 2.30660745775561754067e-01, L6 = 2.06975017800338417784e-01,
      P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
      P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
      P5 = 4.13813679705723846039e-08, lg2 = 6.93147180559945286227e-01,
      lg2_h = 6.93147182464599609375e-01, lg2_l = -1.90465429995776804525e-09,
      cp = 9.61796693925975554329e-01, cp_h = 9.61796700954437255859e-01,
      cp_l = -7.02846165095275826516e-09;
  if (y == 0.0) return 1.0;
  if (x != x || y != y) return x + y;
  if (y == 1.0) return x;
  if (y == 2.0) return x * x;
  if (y == 0.5 && x >= 0.0) return sqrt(x);
  if (!(x > 0.0) || isinf(x) || isinf(y)) return pow(x, y); /* never reached by the solver */
  double z, ax, z_h, z_l, p_h, p_l, y1, t1, t2, r, t, u, v, w;
  double ss, s2, s_h, s_l, t_h, t_l;
  int32_t i, j, k, n;
  int32_t ix = (int32_t)(hiw(x) & 0x7fffffffu);
  ax = fabs(x);
  n = 0;
  if (ix < 0x00100000) { ax *= two53; n -= 53; ix = (int32_t)hiw(ax); }
  n += (ix >> 20) - 0x3ff;
  j = ix & 0x000fffff;
  ix = j | 0x3ff00000;
  if (j <= 0x3988E) k = 0;
  else if (j < 0xBB67A) k = 1;
  else { k = 0; n += 1; ix -= 0x00100000; }
  ax = mkd((uint32_t)ix, low(ax));
  u = ax - bp[k];
  v = 1.0 / (ax + bp[k]);
  ss = u * v;
  s_h = zlo(ss);
  t_h = mkd((uint32_t)(((ix >> 1) | 0x20000000) + 0x00080000 + (k << 18)), 0);
  t_l = ax - (t_h - bp[k]);
  s_l = v * ((u - s_h * t_h) - s_h * t_l);
  s2 = ss * ss;
  r = s2 * s2 * (L1 + s2 * (L2 + s2 * (L3 + s2 * (L4 + s2 * (L5 + s2 * L6)))));
  r += s_l * (s_h + ss);
  s2 = s_h * s_h;
  t_h = zlo(3.0 + s2 + r);
  t_l = r - ((t_h - 3.0) - s2);
  u = s_h * t_h;
  v = s_l * t_h + t_l * ss;
  p_h = zlo(u + v);
  p_l = v - (p_h - u);
  z_h = cp_h * p_h;
  z_l = cp_l * p_h + p_l * cp + dp_l[k];
  t = (double)n;
  t1 = zlo((((z_h + z_l) + dp_h[k]) + t));
  t2 = z_l - (((t1 - t) - dp_h[k]) - z_h);
  y1 = zlo(y);
  p_l = (y - y1) * t1 + y * t2;
  p_h = y1 * t1;
  z = p_l + p_h;
  j = (int32_t)hiw(z);
  if (j >= 0x40900000 || (j & 0x7fffffff) >= 0x4090cc00) return pow(x, y); /* over/underflow */
  i = j & 0x7fffffff;
  k = (i >> 20) - 0x3ff;
  n = 0;
  if (i > 0x3fe00000) {
    n = j + (0x00100000 >> (k + 1));
    k = ((n & 0x7fffffff) >> 20) - 0x3ff;
    t = mkd((uint32_t)(n & ~(0x000fffff >> k)), 0);
    n = ((n & 0x000fffff) | 0x00100000) >> (20 - k);
    if (j < 0) n = -n;
    p_h -= t;
  }
  t = zlo(p_l + p_h);
  u = t * lg2_h;
  v = (p_l - (t - p_h)) * lg2 + t * lg2_l;
  z = u + v;
  w = v - (z - u);
  t = z * z;
  t1 = z - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
  r = (z * t1) / (t1 - 2.0) - (w + z * w);
  z = 1.0 - (r - z);
  j = (int32_t)hiw(z);
  j += (int32_t)((uint32_t)n << 20);
  if ((j >> 20) <= 0) z = scalbn(z, n);
  else z = mkd((uint32_t)j, low(z));
  return z;
}

/* ---- vel(): swe.mjs:282-285 --------------------------------------------- */
double orc_vel(double hq, double h, double eps4) {
  const double h4 = h * h * h * h;
  return (M_SQRT2 * h * hq) / sqrt(h4 + jmax(h4, eps4));
}

/* ---- limiters: swe.mjs:88-141 ------------------------------------------ */
double orc_limit(int lim, double a, double b) {
  if (lim == ORC_LIM_MINMOD) {                 /* swe.mjs:88-92 */
    if (a > 0 && b > 0) return a < b ? a : b;
    if (a < 0 && b < 0) return a > b ? a : b;
    return 0;
  }
  if (lim == ORC_LIM_MC) {                     /* swe.mjs:131-136 */
    if (a * b <= 0) return 0;
    const double c = 0.5 * (a + b);
    const double m = jmin(jmin(fabs(c), 2 * fabs(a)), 2 * fabs(b));
    return a > 0 ? m : -m;
  }
  return 0.5 * (a + b);                        /* central, swe.mjs:139 */
}

/* ---- flux1D / hllc: swe.mjs:974-1018 ----------------------------------- */
static inline void flux1d(double out[3], double h, double hu, double hv, double u, double g) {
  out[0] = hu;
  out[1] = hu * u + 0.5 * g * h * h;
  out[2] = hv * u;
}
void orc_hllc(double out[3], double hL, double huL, double hvL,
              double hR, double huR, double hvR, double dry, double g) {
  const int wetL = hL > dry, wetR = hR > dry;
  if (!wetL && !wetR) { out[0] = out[1] = out[2] = 0; return; }
  const double uL = wetL ? huL / hL : 0, vL = wetL ? hvL / hL : 0;
  const double uR = wetR ? huR / hR : 0, vR = wetR ? hvR / hR : 0;
  const double cL = sqrt(g * jmax(hL, 0)), cR = sqrt(g * jmax(hR, 0));
  double sL, sR;
  if (!wetL) { sL = uR - 2 * cR; sR = uR + cR; }
  else if (!wetR) { sL = uL - cL; sR = uL + 2 * cL; }
  else {
    const double a = (cL + cR) / 2 + (uL - uR) / 4;
    const double hStar = a * a / g;             /* x ** 2 is x*x in the engine */
    const double qL = hStar > hL ? sqrt(0.5 * (hStar + hL) * hStar / (hL * hL)) : 1;
    const double qR = hStar > hR ? sqrt(0.5 * (hStar + hR) * hStar / (hR * hR)) : 1;
    sL = uL - cL * qL;
    sR = uR + cR * qR;
  }
  if (sL >= 0) { flux1d(out, hL, huL, hvL, uL, g); return; }
  if (sR <= 0) { flux1d(out, hR, huR, hvR, uR, g); return; }
  const double num = sL * hR * (uR - sR) - sR * hL * (uL - sL);
  const double den = hR * (uR - sR) - hL * (uL - sL);
  const double sM = fabs(den) > 1e-30 ? num / den : 0.5 * (sL + sR);
  double FL[3], FR[3];
  flux1d(FL, hL, huL, hvL, uL, g);
  flux1d(FR, hR, huR, hvR, uR, g);
  const double inv = 1 / (sR - sL);
  out[0] = (sR * FL[0] - sL * FR[0] + sL * sR * (hR - hL)) * inv;
  out[1] = (sR * FL[1] - sL * FR[1] + sL * sR * (huR - huL)) * inv;
  out[2] = out[0] * (sM >= 0 ? vL : vR);
}

/* ---- construction ------------------------------------------------------ */
static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

orc_sim *orc_create(const orc_params *p, const double *const rows[ORC_NROWS]) {
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  if (!s) return NULL;
  s->p = *p;
  s->ng = 2;
  s->W = p->nx + 2 * s->ng;
  s->H = p->ny + 2 * s->ng;
  s->N = s->W * s->H;
  const int N = s->N;
  const int ids[] = {ORC_B, ORC_H, ORC_HU, ORC_HV, ORC_MANNING, ORC_RH, ORC_RHU, ORC_RHV,
                     ORC_THETA, ORC_SPONGE, ORC_H0, ORC_HU0, ORC_HV0, ORC_U, ORC_V,
                     ORC_FXM, ORC_FXN, ORC_FXT, ORC_FYM, ORC_FYN, ORC_FYT,
                     ORC_PXL, ORC_PXR, ORC_PYL, ORC_PYR};
  for (size_t i = 0; i < sizeof(ids) / sizeof(ids[0]); i++) s->f[ids[i]] = zalloc(N);
  for (int i = ORC_SE; i <= ORC_SV; i++) s->f[i] = zalloc(2 * (size_t)N);
  for (int r = 0; r < ORC_NROWS; r++) {
    s->row[r] = zalloc(s->H);
    if (rows && rows[r]) memcpy(s->row[r], rows[r], sizeof(double) * s->H);
  }
  for (int k = 0; k < 4; k++) s->bc[k] = p->sphere ? (k < 2 ? ORC_BC_PERIODIC : ORC_BC_REFLECT) : ORC_BC_REFLECT;
  return s;
}
void orc_destroy(orc_sim *s) {
  if (!s) return;
  for (int i = 0; i < 64; i++) free(s->f[i]);
  for (int r = 0; r < ORC_NROWS; r++) free(s->row[r]);
  free(s);
}
double *orc_field(orc_sim *s, int w) { return (w >= 0 && w < 64) ? s->f[w] : NULL; }
double *orc_row(orc_sim *s, int w) { return (w >= 0 && w < ORC_NROWS) ? s->row[w] : NULL; }
int orc_W(const orc_sim *s) { return s->W; }
int orc_H(const orc_sim *s) { return s->H; }
void orc_set_bc(orc_sim *s, int side, int kind) { if (side >= 0 && side < 4) s->bc[side] = kind; }
void orc_set_time(orc_sim *s, double t, int64_t steps) { s->t = t; s->steps = steps; }
double orc_time(const orc_sim *s) { return s->t; }
double orc_mass_floored(const orc_sim *s) { return s->mass_floored; }

#define IDX(s, i, j) (((j) + (s)->ng) * (s)->W + ((i) + (s)->ng))

/* ---- boundary conditions: swe.mjs:289-303, 1026-1126 ------------------- */
static void bc_cell(orc_sim *s, int kind, int i, int j, int side) {
  const int nx = s->p.nx, ny = s->p.ny;
  double *b = s->f[ORC_B], *h = s->f[ORC_H], *hu = s->f[ORC_HU], *hv = s->f[ORC_HV];
  if (kind == ORC_BC_REFLECT) {                       /* swe.mjs:1026-1039 */
    int si = i, sj = j;
    if (side == 0) si = -1 - i; else if (side == 1) si = 2 * nx - 1 - i;
    else if (side == 2) sj = -1 - j; else sj = 2 * ny - 1 - j;
    si = imax(0, imin(nx - 1, si));
    sj = imax(0, imin(ny - 1, sj));
    const int g = IDX(s, i, j), k = IDX(s, si, sj);
    b[g] = b[k]; h[g] = h[k];
    const int flipX = side == 0 || side == 1;
    hu[g] = flipX ? -hu[k] : hu[k];
    hv[g] = flipX ? hv[k] : -hv[k];
  } else if (kind == ORC_BC_OUTFLOW) {                /* swe.mjs:1042-1048 */
    const int si = imax(0, imin(nx - 1, i)), sj = imax(0, imin(ny - 1, j));
    const int g = IDX(s, i, j), k = IDX(s, si, sj);
    b[g] = b[k]; h[g] = h[k]; hu[g] = hu[k]; hv[g] = hv[k];
  } else if (kind == ORC_BC_PERIODIC) {               /* swe.mjs:1119-1126 */
    int si = i, sj = j;
    if (side == 0) si = i + nx; else if (side == 1) si = i - nx;
    else if (side == 2) sj = j + ny; else sj = j - ny;
    const int g = IDX(s, i, j), k = IDX(s, si, sj);
    b[g] = b[k]; h[g] = h[k]; hu[g] = hu[k]; hv[g] = hv[k];
  } else {                                            /* flather, swe.mjs:1084-1116 */
    const int si = imax(0, imin(nx - 1, i)), sj = imax(0, imin(ny - 1, j));
    const int g = IDX(s, i, j), k = IDX(s, si, sj);
    b[g] = b[k];
    const double bed = b[g];
    const double hInt = h[k];
    const double e0 = s->ext[2 * side], u0 = s->ext[2 * side + 1];
    if (hInt <= s->p.min_depth) { h[g] = jmax0(e0 - bed); hu[g] = 0; hv[g] = 0; return; }
    const double etaInt = bed + hInt;
    const double r = sqrt(s->p.g / hInt);
    const double sgn = (side == 0 || side == 2) ? 1 : -1;
    const double uInt = (side == 0 || side == 1) ? orc_vel(hu[k], hInt, s->p.eps4)
                                                 : orc_vel(hv[k], hInt, s->p.eps4);
    const double Rin = u0 + sgn * e0 * r;
    const double Rout = uInt - sgn * etaInt * r;
    const double un = 0.5 * (Rin + Rout);
    const double en = sgn * (Rin - Rout) / (2 * r);
    h[g] = jmax0(en - bed);
    if (side == 0 || side == 1) { hu[g] = h[g] * un; hv[g] = hv[k]; }
    else { hv[g] = h[g] * un; hu[g] = hu[k]; }
  }
}

void orc_apply_bc(orc_sim *s, const double ext[8]) {
  if (ext) memcpy(s->ext, ext, sizeof(s->ext)); else memset(s->ext, 0, sizeof(s->ext));
  const int H = s->H, ng = s->ng, nx = s->p.nx, ny = s->p.ny;
  for (int j = 0; j < H; j++)
    for (int g = 0; g < ng; g++) {
      bc_cell(s, s->bc[0], -1 - g, j - ng, 0);
      bc_cell(s, s->bc[1], nx + g, j - ng, 1);
    }
  for (int i = -ng; i < nx + ng; i++)
    for (int g = 0; g < ng; g++) {
      bc_cell(s, s->bc[2], i, -1 - g, 2);
      bc_cell(s, s->bc[3], i, ny + g, 3);
    }
}

/* ---- residual(dt): swe.mjs:308-629 ------------------------------------- */
void orc_residual(orc_sim *s, double dt) {
  const int W = s->W, H = s->H, ng = s->ng, N = s->N;
  const double minDepth = s->p.min_depth, G = s->p.g, eps4 = s->p.eps4;
  const int sph = s->p.sphere;
  const double *dxRow = s->row[ORC_ROW_DX], *dyRowN = s->row[ORC_ROW_DYN],
               *dyRowS = s->row[ORC_ROW_DYS], *bedPhi = s->row[ORC_ROW_BEDPHI];
  const double *h = s->f[ORC_H], *hu = s->f[ORC_HU], *hv = s->f[ORC_HV], *b = s->f[ORC_B];
  double *rh = s->f[ORC_RH], *rhu = s->f[ORC_RHU], *rhv = s->f[ORC_RHV];
  double *fxM = s->f[ORC_FXM], *fxN = s->f[ORC_FXN], *fxT = s->f[ORC_FXT];
  double *fyM = s->f[ORC_FYM], *fyN = s->f[ORC_FYN], *fyT = s->f[ORC_FYT];
  double *pxL = s->f[ORC_PXL], *pxR = s->f[ORC_PXR], *pyL = s->f[ORC_PYL], *pyR = s->f[ORC_PYR];
  double *U = s->f[ORC_U], *V = s->f[ORC_V];
  double *sE = s->f[ORC_SE], *sB = s->f[ORC_SB], *sU = s->f[ORC_SU], *sV = s->f[ORC_SV];
  double *theta = s->f[ORC_THETA];
  for (int k = 0; k < N; k++) { rh[k] = 0; rhu[k] = 0; rhv[k] = 0; fxM[k] = 0; fyM[k] = 0; }

  for (int k = 0; k < N; k++) { const double hk = h[k]; U[k] = orc_vel(hu[k], hk, eps4); V[k] = orc_vel(hv[k], hk, eps4); }

  if (s->p.order >= 2) {                                         /* swe.mjs:364-411 */
    const int lim = s->p.limiter;
    for (int j = 1; j < H - 1; j++)
      for (int i = 1; i < W - 1; i++) {
        const int k = j * W + i, k2 = 2 * k;
        const double hk = h[k], bk = b[k], eC = bk + hk, uk = U[k], vk = V[k];
        for (int d = 0; d < 2; d++) {
          const int kp = d == 0 ? k + 1 : k + W;
          const int km = d == 0 ? k - 1 : k - W;
          if (hk <= minDepth || h[kp] <= minDepth || h[km] <= minDepth) {
            sE[k2 + d] = 0; sB[k2 + d] = 0; sU[k2 + d] = 0; sV[k2 + d] = 0;
            continue;
          }
          const double eP = b[kp] + h[kp], eM = b[km] + h[km];
          sE[k2 + d] = orc_limit(lim, eC - eM, eP - eC);
          sB[k2 + d] = orc_limit(lim, bk - b[km], b[kp] - bk);
          sU[k2 + d] = orc_limit(lim, uk - U[km], U[kp] - uk);
          sV[k2 + d] = orc_limit(lim, vk - V[km], V[kp] - vk);
        }
      }
  } else {
    memset(sE, 0, sizeof(double) * 2 * N); memset(sU, 0, sizeof(double) * 2 * N);
    memset(sV, 0, sizeof(double) * 2 * N); memset(sB, 0, sizeof(double) * 2 * N);
  }

  double flux[3];
  for (int j = ng; j < H - ng; j++)                              /* x faces :413-445 */
    for (int i = ng - 1; i < W - ng; i++) {
      const int kL = j * W + i, kR = kL + 1, kL2 = 2 * kL, kR2 = 2 * kR;
      const double etaL = b[kL] + h[kL] + 0.5 * sE[kL2];
      const double etaR = b[kR] + h[kR] - 0.5 * sE[kR2];
      const double bL = b[kL] + 0.5 * sB[kL2];
      const double bR = b[kR] - 0.5 * sB[kR2];
      const double hLf = jmax0(etaL - bL);
      const double hRf = jmax0(etaR - bR);
      const double uL = U[kL] + 0.5 * sU[kL2];
      const double uR = U[kR] - 0.5 * sU[kR2];
      const double vL = V[kL] + 0.5 * sV[kL2];
      const double vR = V[kR] - 0.5 * sV[kR2];
      const double bStar = jmax(bL, bR);
      const double hsL = jmax0(etaL - bStar);
      const double hsR = jmax0(etaR - bStar);
      orc_hllc(flux, hsL, hsL * uL, hsL * vL, hsR, hsR * uR, hsR * vR, minDepth, G);
      const double pL = 0.5 * G * (hLf * hLf - hsL * hsL);
      const double pR = 0.5 * G * (hRf * hRf - hsR * hsR);
      fxM[kL] = flux[0]; fxN[kL] = flux[1]; fxT[kL] = flux[2];
      pxL[kL] = pL; pxR[kL] = pR;
    }
  for (int j = ng - 1; j < H - ng; j++)                          /* y faces :448-475 */
    for (int i = ng; i < W - ng; i++) {
      const int kL = j * W + i, kR = kL + W, kL2 = 2 * kL + 1, kR2 = 2 * kR + 1;
      const double etaL = b[kL] + h[kL] + 0.5 * sE[kL2];
      const double etaR = b[kR] + h[kR] - 0.5 * sE[kR2];
      const double bL = b[kL] + 0.5 * sB[kL2];
      const double bR = b[kR] - 0.5 * sB[kR2];
      const double hLf = jmax0(etaL - bL);
      const double hRf = jmax0(etaR - bR);
      const double uL = U[kL] + 0.5 * sU[kL2];
      const double uR = U[kR] - 0.5 * sU[kR2];
      const double vL = V[kL] + 0.5 * sV[kL2];
      const double vR = V[kR] - 0.5 * sV[kR2];
      const double bStar = jmax(bL, bR);
      const double hsL = jmax0(etaL - bStar);
      const double hsR = jmax0(etaR - bStar);
      orc_hllc(flux, hsL, hsL * vL, hsL * uL, hsR, hsR * vR, hsR * uR, minDepth, G);
      const double pL = 0.5 * G * (hLf * hLf - hsL * hsL);
      const double pR = 0.5 * G * (hRf * hRf - hsR * hsR);
      fyM[kL] = flux[0]; fyN[kL] = flux[1]; fyT[kL] = flux[2];
      pyL[kL] = pL; pyR[kL] = pR;
    }

  for (int k = 0; k < N; k++) theta[k] = 1;                      /* positivity :507-529 */
  if (dt > 0) {
    for (int j = ng - 1; j < H - ng + 1; j++)
      for (int i = ng - 1; i < W - ng + 1; i++) {
        const int k = j * W + i;
        double out = 0;
        if (fxM[k] > 0) out += fxM[k] / dxRow[j];
        if (fxM[k - 1] < 0) out -= fxM[k - 1] / dxRow[j];
        if (fyM[k] > 0) out += fyM[k] / dyRowN[j];
        if (fyM[k - W] < 0) out -= fyM[k - W] / dyRowS[j];
        const double avail = h[k];
        if (out * dt > avail) theta[k] = avail > 0 ? avail / (out * dt) : 0;
      }
  }
  for (int j = ng; j < H - ng; j++)                              /* accumulate :530-555 */
    for (int i = ng - 1; i < W - ng; i++) {
      const int kL = j * W + i, kR = kL + 1;
      const double th = fxM[kL] >= 0 ? theta[kL] : theta[kR];
      const double inv = th / dxRow[j];
      rh[kL] -= fxM[kL] * inv; rhu[kL] -= (fxN[kL] + pxL[kL]) * inv; rhv[kL] -= fxT[kL] * inv;
      rh[kR] += fxM[kL] * inv; rhu[kR] += (fxN[kL] + pxR[kL]) * inv; rhv[kR] += fxT[kL] * inv;
    }
  for (int j = ng - 1; j < H - ng; j++)
    for (int i = ng; i < W - ng; i++) {
      const int kL = j * W + i, kR = kL + W;
      const double th = fyM[kL] >= 0 ? theta[kL] : theta[kR];
      const double invL = th / dyRowN[j];
      const double invR = th / dyRowS[j + 1];
      rh[kL] -= fyM[kL] * invL; rhv[kL] -= (fyN[kL] + pyL[kL]) * invL; rhu[kL] -= fyT[kL] * invL;
      rh[kR] += fyM[kL] * invR; rhv[kR] += (fyN[kL] + pyR[kL]) * invR; rhu[kR] += fyT[kL] * invR;
    }

  if (s->p.order >= 2 || sph) {                                  /* bed + geometric :570-598 */
    const double *geoCoef = s->row[ORC_ROW_GEO];
    for (int j = ng; j < H - ng; j++) {
      const double bedDiv1 = bedPhi[j], bedDiv0 = dxRow[j], gc = geoCoef[j];
      for (int i = ng; i < W - ng; i++) {
        const int k = j * W + i, k2 = 2 * k;
        const double bk = b[k];
        const double eC = bk + h[k];
        for (int d = 0; d < 2; d++) {
          const double hP = jmax0(eC + 0.5 * sE[k2 + d] - (bk + 0.5 * sB[k2 + d]));
          const double hM = jmax0(eC - 0.5 * sE[k2 + d] - (bk - 0.5 * sB[k2 + d]));
          const double db = sB[k2 + d];
          const double term = -G * 0.5 * (hP + hM) * db / (d == 0 ? bedDiv0 : bedDiv1);
          if (d == 0) rhu[k] += term;
          else {
            rhv[k] += term;
            if (sph) rhv[k] += gc * G * (hP * hP + hM * hM);
          }
        }
      }
    }
  }

  if (sph) {                                                     /* Coriolis :600-628 */
    const double *fRow = s->row[ORC_ROW_F], *tanPhi = s->row[ORC_ROW_TANPHI];
    const double R = s->p.R;
    for (int j = ng; j < H - ng; j++) {
      const double f0 = fRow[j], tR = tanPhi[j] / R;
      for (int i = ng; i < W - ng; i++) {
        const int k = j * W + i;
        const double fe = f0 + U[k] * tR;
        rhu[k] += fe * hv[k];
        rhv[k] -= fe * hu[k];
      }
    }
  } else if (s->p.coriolis != 0) {
    const double f = s->p.coriolis;
    for (int j = ng; j < H - ng; j++)
      for (int i = ng; i < W - ng; i++) {
        const int k = j * W + i;
        rhu[k] += f * hv[k];
        rhv[k] -= f * hu[k];
      }
  }
}

/* ---- maxDt(): swe.mjs:632-647 ------------------------------------------ */
double orc_max_dt(orc_sim *s) {
  const int W = s->W, H = s->H, ng = s->ng;
  const double *dxRow = s->row[ORC_ROW_DX], *dyCFL = s->row[ORC_ROW_DYCFL];
  const double *h = s->f[ORC_H], *hu = s->f[ORC_HU], *hv = s->f[ORC_HV];
  double inv = 0;
  for (int j = ng; j < H - ng; j++)
    for (int i = ng; i < W - ng; i++) {
      const int k = j * W + i;
      const double hk = h[k];
      if (hk <= s->p.min_depth) continue;
      const double c = sqrt(s->p.g * hk);
      const double u = fabs(orc_vel(hu[k], hk, s->p.eps4)), v = fabs(orc_vel(hv[k], hk, s->p.eps4));
      const double sp = (u + c) / dxRow[j] + (v + c) / dyCFL[j];
      if (sp > inv) inv = sp;
    }
  return inv > 0 ? s->p.cfl / inv : INFINITY;
}

/* ---- applyFriction(dt): swe.mjs:659-704 -------------------------------- */
void orc_apply_friction(orc_sim *s, double dt) {
  const int W = s->W, H = s->H, ng = s->ng;
  double *h = s->f[ORC_H], *hu = s->f[ORC_HU], *hv = s->f[ORC_HV];
  const double G = s->p.g, md = s->p.min_depth, e4 = s->p.eps4;
  const double four3 = 4.0 / 3.0;
  if (s->p.use_manning_field) {
    const double *nf = s->f[ORC_MANNING];
    for (int j = ng; j < H - ng; j++)
      for (int i = ng; i < W - ng; i++) {
        const int k = j * W + i;
        const double hk = h[k];
        if (hk <= md) { hu[k] = 0; hv[k] = 0; continue; }
        const double nk = nf[k];
        if (nk == 0 || nk != nk) continue;                         /* !nk */
        const double u = orc_vel(hu[k], hk, e4), v = orc_vel(hv[k], hk, e4);
        const double speed = orc_hypot(u, v);
        if (speed < 1e-14) continue;
        const double denom = 1 + dt * G * nk * nk * speed / orc_pow(hk, four3);
        hu[k] /= denom;
        hv[k] /= denom;
      }
    return;
  }
  const double n = s->p.manning;
  if (n == 0 || n != n) return;
  for (int j = ng; j < H - ng; j++)
    for (int i = ng; i < W - ng; i++) {
      const int k = j * W + i;
      const double hk = h[k];
      if (hk <= md) { hu[k] = 0; hv[k] = 0; continue; }
      const double u = orc_vel(hu[k], hk, e4), v = orc_vel(hv[k], hk, e4);
      const double speed = orc_hypot(u, v);
      if (speed < 1e-14) continue;
      const double denom = 1 + dt * G * n * n * speed / orc_pow(hk, four3);
      hu[k] /= denom;
      hv[k] /= denom;
    }
}

/* ---- dryClean(): swe.mjs:734-745 --------------------------------------- */
double orc_dry_clean(orc_sim *s) {
  const int W = s->W, H = s->H, ng = s->ng;
  double *h = s->f[ORC_H], *hu = s->f[ORC_HU], *hv = s->f[ORC_HV];
  double added = 0;
  for (int j = ng; j < H - ng; j++)
    for (int i = ng; i < W - ng; i++) {
      const int k = j * W + i;
      if (h[k] < 0) { added -= h[k]; h[k] = 0; }
      if (h[k] <= s->p.min_depth) { hu[k] = 0; hv[k] = 0; }
    }
  return added;
}

/* ---- step(dt): swe.mjs:754-814 (forcing hook excluded) ----------------- */
double orc_step(orc_sim *s, double dt_in, const double ext[8]) {
  orc_apply_bc(s, ext);
  const double dt = (dt_in != dt_in) ? orc_max_dt(s) : dt_in;
  if (!(dt > 0) || isinf(dt)) return 0;
  const int N = s->N;
  double *h = s->f[ORC_H], *hu = s->f[ORC_HU], *hv = s->f[ORC_HV];
  double *h0 = s->f[ORC_H0], *hu0 = s->f[ORC_HU0], *hv0 = s->f[ORC_HV0];
  const double *rh = s->f[ORC_RH], *rhu = s->f[ORC_RHU], *rhv = s->f[ORC_RHV];
  memcpy(h0, h, sizeof(double) * N); memcpy(hu0, hu, sizeof(double) * N); memcpy(hv0, hv, sizeof(double) * N);
  orc_residual(s, dt);
  for (int k = 0; k < N; k++) { h[k] += dt * rh[k]; hu[k] += dt * rhu[k]; hv[k] += dt * rhv[k]; }
  s->mass_floored = orc_dry_clean(s);
  orc_apply_bc(s, ext);
  orc_residual(s, dt);
  for (int k = 0; k < N; k++) {
    h[k] = 0.5 * (h0[k] + h[k] + dt * rh[k]);
    hu[k] = 0.5 * (hu0[k] + hu[k] + dt * rhu[k]);
    hv[k] = 0.5 * (hv0[k] + hv[k] + dt * rhv[k]);
  }
  s->mass_floored += orc_dry_clean(s);
  orc_apply_friction(s, dt);
  s->t += dt;
  s->steps++;
  return dt;
}

int64_t orc_run(orc_sim *s, int64_t nsteps, const double ext[8]) {
  int64_t n = 0;
  for (; n < nsteps; n++) if (orc_step(s, NAN, ext) == 0) break;
  return n;
}

/* ---- makeSponge closure body: swe.mjs:1188-1197 ------------------------ */
void orc_sponge(orc_sim *s, double dt, double ref) {
  const int W = s->W, H = s->H, ng = s->ng;
  const double *str = s->f[ORC_SPONGE], *b = s->f[ORC_B];
  double *h = s->f[ORC_H], *hu = s->f[ORC_HU], *hv = s->f[ORC_HV];
  for (int j = ng; j < H - ng; j++)
    for (int i = ng; i < W - ng; i++) {
      const int k = j * W + i;
      if (!(str[k] > 0)) continue;
      const double f = jmin(1, str[k] * dt * 60);
      const double target = jmax0(ref - b[k]);
      h[k] += (target - h[k]) * f;
      hu[k] -= hu[k] * f;
      hv[k] -= hv[k] * f;
    }
}

/* ---- diagnostics: swe.mjs:819-917 -------------------------------------- */
double orc_volume(orc_sim *s) {
  const int W = s->W, ng = s->ng, nx = s->p.nx, ny = s->p.ny;
  const double *h = s->f[ORC_H];
  if (!s->p.sphere) {
    double v = 0;
    for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) v += h[(j + ng) * W + (i + ng)];
    return v * s->p.dx * s->p.dy;
  }
  const double *area = s->row[ORC_ROW_AREA];
  double v = 0;
  for (int j = 0; j < ny; j++) {
    double row = 0;
    for (int i = 0; i < nx; i++) row += h[(j + ng) * W + (i + ng)];
    v += row * area[j + ng];
  }
  return v;
}
double orc_energy(orc_sim *s) {
  const int W = s->W, ng = s->ng, nx = s->p.nx, ny = s->p.ny;
  const double *h = s->f[ORC_H], *hu = s->f[ORC_HU], *hv = s->f[ORC_HV], *b = s->f[ORC_B];
  const double *area = s->row[ORC_ROW_AREA];
  double e = 0;
  for (int j = 0; j < ny; j++) {
    double row = 0;
    for (int i = 0; i < nx; i++) {
      const int k = (j + ng) * W + (i + ng);
      const double hk = h[k];
      if (hk <= 0) continue;
      const double u = orc_vel(hu[k], hk, s->p.eps4), v = orc_vel(hv[k], hk, s->p.eps4);
      row += 0.5 * hk * (u * u + v * v) + 0.5 * s->p.g * hk * (hk + 2 * b[k]);
    }
    if (s->p.sphere) e += row * area[j + ng]; else e += row;
  }
  return s->p.sphere ? e : e * s->p.dx * s->p.dy;
}
double orc_max_surface_deviation(orc_sim *s, double ref) {
  const int W = s->W, ng = s->ng, nx = s->p.nx, ny = s->p.ny;
  const double *h = s->f[ORC_H], *b = s->f[ORC_B];
  double m = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = (j + ng) * W + (i + ng);
      if (h[k] <= s->p.min_depth) continue;
      const double d = fabs(b[k] + h[k] - ref);
      if (d > m) m = d;
    }
  return m;
}
double orc_max_speed(orc_sim *s) {
  const int W = s->W, ng = s->ng, nx = s->p.nx, ny = s->p.ny;
  const double *h = s->f[ORC_H], *hu = s->f[ORC_HU], *hv = s->f[ORC_HV];
  double m = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = (j + ng) * W + (i + ng);
      const double hk = h[k];
      if (hk <= s->p.min_depth) continue;
      const double sp = orc_hypot(orc_vel(hu[k], hk, s->p.eps4), orc_vel(hv[k], hk, s->p.eps4));
      if (sp > m) m = sp;
    }
  return m;
}
int orc_finite(orc_sim *s) {
  const double *h = s->f[ORC_H], *hu = s->f[ORC_HU], *hv = s->f[ORC_HV];
  for (int k = 0; k < s->N; k++)
    if (!isfinite(h[k]) || !isfinite(hu[k]) || !isfinite(hv[k])) return 0;
  return 1;
}

/* FNV-1a 64-bit over raw little-endian bytes (the fingerprint idea of swe.mjs:349-357) */
uint64_t orc_fnv1a(const void *bytes, size_t n, uint64_t h) {
  const unsigned char *p = (const unsigned char *)bytes;
  if (h == 0) h = 0xcbf29ce484222325ULL;
  for (size_t i = 0; i < n; i++) { h ^= p[i]; h *= 0x100000001b3ULL; }
  return h;
}
