// Device-side arithmetic of the shallow-water step, written to reproduce the
// reference solver (src/swe.mjs) double for double:
//   * every expression keeps JavaScript's left-to-right association;
//   * compiled with -ffp-contract=off, so a*b+c is never fused behind our back;
//   * the only explicit FMAs are inside exact-division helpers whose result is
//     the correctly rounded quotient (Markstein's correction step), i.e. the
//     same double that an IEEE divide returns;
//   * Math.max / Math.min keep their NaN and signed-zero behaviour.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swe {

constexpr double SQRT2 = 1.4142135623730951;   // Math.SQRT2

// ---- JS Math semantics ----------------------------------------------------
__host__ __device__ __forceinline__ uint64_t dbits(double x) { return __builtin_bit_cast(uint64_t, x); }
__host__ __device__ __forceinline__ double dfrom(uint64_t u) { return __builtin_bit_cast(double, u); }
// Math.max(0, x): NaN propagates, -0 becomes +0
__host__ __device__ __forceinline__ double jmax0(double x) { return (x > 0.0 || x != x) ? x : 0.0; }
// Math.max(a, b) / Math.min(a, b), branch-free: NaN (the canonical quiet NaN)
// when either is NaN; of two equal values -- +0 and -0 among them -- the one
// whose sign bit is the AND (max) or the OR (min) of theirs, so max(+0, -0) =
// +0 and min(+0, -0) = -0; otherwise the hardware max / min, which then only
// decides between two distinct ordered values.
__host__ __device__ __forceinline__ double jmax(double a, double b) {
  const double m = a == b ? dfrom(dbits(a) & dbits(b)) : __builtin_fmax(a, b);
  return (a != a || b != b) ? __builtin_nan("") : m;
}
__host__ __device__ __forceinline__ double jmin(double a, double b) {
  const double m = a == b ? dfrom(dbits(a) | dbits(b)) : __builtin_fmin(a, b);
  return (a != a || b != b) ? __builtin_nan("") : m;
}
// Math.min of two values known to be >= +0 or NaN (no -0): the hardware min,
// NaN when either is NaN
__host__ __device__ __forceinline__ double jmin_nn(double a, double b) {
  return (a != a || b != b) ? __builtin_nan("") : __builtin_fmin(a, b);
}

// ---- exact division helpers -----------------------------------------------
// x / d given r = RN(1/d): q = RN(x r); e = x - q d (exact by FMA); RN(q + e r)
// is RN(x/d) for d, x, x/d normal (Markstein).  Where the corrected quotient is
// zero or NaN, q = x * r is x / d itself: a zero quotient keeps the sign of x*r
// (the correction would give +0 for x = -0); an infinite d (a pole face) has
// r = 0 (q a signed zero, or NaN for an infinite or NaN x); an infinite x over
// a finite d gives q = +-inf, and a NaN x or d gives NaN.  One compare on |q1|
// (false for zero and NaN) and a select.  Branch-free.
__host__ __device__ __forceinline__ double div_r(double x, double d, double r) {
  const double q = x * r;
  const double e = __builtin_fma(-q, d, x);
  const double q1 = __builtin_fma(e, r, q);
  return __builtin_fabs(q1) > 0.0 ? q1 : q;
}

// ---- square roots ----------------------------------------------------------
// Math.sqrt (correctly rounded) of an x that is NaN, negative, +-0, +inf or at
// least 2^-767.  Goldschmidt's iteration from the hardware reciprocal square
// root estimate y ~ 1/sqrt(x): g = x y ~ sqrt(x) and h = y/2 ~ 1/(2 sqrt(x)),
// one coupled step r = 1/2 - h g, g += g r, h += h r, then two Newton-Raphson
// corrections g += (x - g^2) h with the residual exact by FMA.  These are the
// steps the compiler's own sqrt takes once it has rescaled an x below 2^-767
// (and rescaled back), so the two agree bit for bit wherever x needs no
// rescaling; +-0 and +inf are returned as they are (the estimate is inf or 0
// there), NaN and negative x give NaN through the estimate.
__host__ __device__ __forceinline__ double sqrt_n(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return __builtin_amdgcn_class(x, 0x260) ? x : g;   // 0x260: -0, +0, +inf
#else
  return __builtin_sqrt(x);
#endif
}
// sqrt(g h) for a depth h above minDepth: sqrt_n when g minDepth >= 2^-767
// (`normal`, uniform over a launch), the full sqrt otherwise
__host__ __device__ __forceinline__ double sqrt_gh(double x, bool normal) {
  return normal ? sqrt_n(x) : __builtin_sqrt(x);
}
constexpr double SQRT_N_MIN = 0x1p-767;

// ---- vel(): swe.mjs:282-285.  Both components share one denominator. --------
struct Vel2 { double u, v; };
// Math.max(h4, eps4) with eps4 > 0 is the hardware max: the two cannot be equal
// zeros, and where h4 is NaN the sum is NaN whichever operand the max returns.
__host__ __device__ __forceinline__ double vel_den(double h, double eps4) {
  const double h4 = h * h * h * h;
  // the argument is at least eps4 = max(minDepth, 1e-12)^4 >= 1e-48 (or NaN, +inf)
  return sqrt_n(h4 + __builtin_fmax(h4, eps4));
}
__host__ __device__ __forceinline__ Vel2 vel2(double hu, double hv, double h, double eps4) {
  const double den = vel_den(h, eps4);
  const double r = 1.0 / den;
  const double s = SQRT2 * h;
  Vel2 o;
  o.u = div_r(s * hu, den, r);
  o.v = div_r(s * hv, den, r);
  return o;
}
__host__ __device__ __forceinline__ double vel1(double hq, double h, double eps4) {
  return (SQRT2 * h * hq) / vel_den(h, eps4);
}

// ---- limiters: swe.mjs:88-141 ---------------------------------------------
enum { LIM_MINMOD = 0, LIM_MC = 1, LIM_CENTRAL = 2 };
__host__ __device__ __forceinline__ double limit(int lim, double a, double b) {
  if (lim == LIM_MC) {
    // min(|c|, 2|a|, 2|b|) as Math.min gives it, NaN included: the three are
    // absolute values (no -0), so the hardware min picks the same double, and
    // one of them is NaN where a result is used (a * b > 0) only if a or b is
    const double c = 0.5 * (a + b);
    const double m0 = __builtin_fmin(__builtin_fmin(__builtin_fabs(c), 2 * __builtin_fabs(a)), 2 * __builtin_fabs(b));
    const double m = (a != a || b != b) ? __builtin_nan("") : m0;
    return a * b <= 0 ? 0.0 : (a > 0 ? m : -m);
  }
  if (lim == LIM_MINMOD) {
    if (a > 0 && b > 0) return a < b ? a : b;
    if (a < 0 && b < 0) return a > b ? a : b;
    return 0.0;
  }
  return 0.5 * (a + b);
}

// ---- HLLC with Einfeldt speeds and dry states: swe.mjs:974-1018 -----------
struct Flux3 { double m, n, t; };
// rg = RN(1/g); normal: g * dry >= 2^-767, so sqrt(g h) of a wet depth h > dry
// needs no rescaling (sqrt_gh)
struct HllcConst { double g, half_g, rg, dry; bool normal; };

// H_NONNEG: the caller guarantees hL, hR >= +0 or NaN (the hydrostatic
// reconstruction's depths are), where Math.max(h, 0) is h itself wherever the
// wave speed it feeds is used (the side is wet: h > dry)
template <bool H_NONNEG = false>
__host__ __device__ __forceinline__ Flux3 hllc(double hL, double huL, double hvL,
                                      double hR, double huR, double hvR, const HllcConst &C) {
  Flux3 out;
  const bool wetL = hL > C.dry, wetR = hR > C.dry;
  if (!wetL && !wetR) { out.m = 0.0; out.n = 0.0; out.t = 0.0; return out; }
  double uL = 0.0, vL = 0.0, uR = 0.0, vR = 0.0;
  if (wetL) { const double r = 1.0 / hL; uL = div_r(huL, hL, r); vL = div_r(hvL, hL, r); }
  if (wetR) { const double r = 1.0 / hR; uR = div_r(huR, hR, r); vR = div_r(hvR, hR, r); }
  // cL is used only where the left side is wet (hL > dry), cR likewise
  const double cL = sqrt_gh(C.g * (H_NONNEG ? hL : jmax(hL, 0.0)), C.normal);
  const double cR = sqrt_gh(C.g * (H_NONNEG ? hR : jmax(hR, 0.0)), C.normal);
  double sL, sR;
  if (!wetL) { sL = uR - 2 * cR; sR = uR + cR; }
  else if (!wetR) { sL = uL - cL; sR = uL + 2 * cL; }
  else {
    const double a = (cL + cR) * 0.5 + (uL - uR) * 0.25;   // (cL+cR)/2 + (uL-uR)/4, exact halving
    const double hStar = div_r(a * a, C.g, C.rg);
    // the arguments exceed 1 where the root is taken (hStar > h), or are NaN / +inf
    const double qL = hStar > hL ? sqrt_n(0.5 * (hStar + hL) * hStar / (hL * hL)) : 1.0;
    const double qR = hStar > hR ? sqrt_n(0.5 * (hStar + hR) * hStar / (hR * hR)) : 1.0;
    sL = uL - cL * qL;
    sR = uR + cR * qR;
  }
  if (sL >= 0) { out.m = huL; out.n = huL * uL + C.half_g * hL * hL; out.t = hvL * uL; return out; }
  if (sR <= 0) { out.m = huR; out.n = huR * uR + C.half_g * hR * hR; out.t = hvR * uR; return out; }
  const double num = sL * hR * (uR - sR) - sR * hL * (uL - sL);
  const double den = hR * (uR - sR) - hL * (uL - sL);
  const double sM = __builtin_fabs(den) > 1e-30 ? num / den : 0.5 * (sL + sR);
  const double FL0 = huL, FL1 = huL * uL + C.half_g * hL * hL;
  const double FR0 = huR, FR1 = huR * uR + C.half_g * hR * hR;
  const double inv = 1.0 / (sR - sL);
  out.m = (sR * FL0 - sL * FR0 + sL * sR * (hR - hL)) * inv;
  out.n = (sR * FL1 - sL * FR1 + sL * sR * (huR - huL)) * inv;
  out.t = out.m * (sM >= 0 ? vL : vR);
  return out;
}

// ---- Math.hypot (2 args), V8's normalise-by-max Kahan form -----------------
// V8 sums (x/m)^2 and (y/m)^2, m the larger magnitude, with Kahan's
// compensation.  For two values the compensation after the first term is
// exactly 0, one quotient is m/m = 1 exactly, and IEEE addition commutes, so
// the sum is 1 + q^2 with q the smaller magnitude over the larger: one
// division where the loop takes two, and the same double.
__host__ __device__ __forceinline__ double js_hypot(double a, double b) {
  if (a != a || b != b) {
    if (__builtin_isinf(a) || __builtin_isinf(b)) return __builtin_inf();
    return __builtin_nan("");
  }
  const double x = __builtin_fabs(a), y = __builtin_fabs(b);
  double m = 0.0;
  if (x > m) m = x;
  if (y > m) m = y;
  if (m == __builtin_inf()) return __builtin_inf();
  if (m == 0.0) return 0.0;
  const double q = (x > y ? y : x) / m;
  return sqrt_n(1.0 + q * q) * m;   // the sum is in [1, 2]
}

// ---- x**y for finite x > 0: fdlibm e_pow general path ----------------------
__host__ __device__ __forceinline__ uint32_t hiw(double x) { return (uint32_t)(__builtin_bit_cast(uint64_t, x) >> 32); }
__host__ __device__ __forceinline__ uint32_t low(double x) { return (uint32_t)__builtin_bit_cast(uint64_t, x); }
__host__ __device__ __forceinline__ double mkd(uint32_t hi, uint32_t lo) {
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__host__ __device__ __forceinline__ double zlo(double x) { return mkd(hiw(x), 0u); }

// js_pow_core is inlined where y is a constant (the friction's 4/3, whose
// split into high and low parts then folds at compile time); js_pow_pos is the
// out-of-line form.  Also compiled for the host: the C ABI exports it
// (swe_js_pow) so host code can form eps4 = max(minDepth,1e-12)**4 the way the
// JS engine does.
__host__ __device__ __forceinline__ double js_pow_core(double x, double y) {
  const double two53 = 9007199254740992.0,
      L1 = 5.99999999999994648725e-01, L2 = 4.28571428578550184252e-01,
      L3 = 3.33333329818377432918e-01, L4 = 2.72728123808534006489e-01,
      L5 = 2.30660745775561754067e-01, L6 = 2.06975017800338417784e-01,
      P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
      P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
      P5 = 4.13813679705723846039e-08, lg2 = 6.93147180559945286227e-01,
      lg2_h = 6.93147182464599609375e-01, lg2_l = -1.90465429995776804525e-09,
      cp = 9.61796693925975554329e-01, cp_h = 9.61796700954437255859e-01,
      cp_l = -7.02846165095275826516e-09;
  double z, ax, z_h, z_l, p_h, p_l, y1, t1, t2, r, t, u, v, w, ss, s2, s_h, s_l, t_h, t_l;
  int32_t i, j, k, n;
  int32_t ix = (int32_t)(hiw(x) & 0x7fffffffu);
  ax = __builtin_fabs(x);
  n = 0;
  if (ix < 0x00100000) { ax *= two53; n -= 53; ix = (int32_t)hiw(ax); }
  n += (ix >> 20) - 0x3ff;
  j = ix & 0x000fffff;
  ix = j | 0x3ff00000;
  if (j <= 0x3988E) k = 0;
  else if (j < 0xBB67A) k = 1;
  else { k = 0; n += 1; ix -= 0x00100000; }
  ax = mkd((uint32_t)ix, low(ax));
  const double bpk = k ? 1.5 : 1.0;
  const double dphk = k ? 5.84962487220764160156e-01 : 0.0;
  const double dplk = k ? 1.35003920212974897128e-08 : 0.0;
  u = ax - bpk;
  v = 1.0 / (ax + bpk);
  ss = u * v;
  s_h = zlo(ss);
  t_h = mkd((uint32_t)(((ix >> 1) | 0x20000000) + 0x00080000 + (k << 18)), 0u);
  t_l = ax - (t_h - bpk);
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
  z_l = cp_l * p_h + p_l * cp + dplk;
  t = (double)n;
  t1 = zlo((((z_h + z_l) + dphk) + t));
  t2 = z_l - (((t1 - t) - dphk) - z_h);
  y1 = zlo(y);
  p_l = (y - y1) * t1 + y * t2;
  p_h = y1 * t1;
  z = p_l + p_h;
  j = (int32_t)hiw(z);
  i = j & 0x7fffffff;
  k = (i >> 20) - 0x3ff;
  n = 0;
  if (i > 0x3fe00000) {
    n = j + (0x00100000 >> (k + 1));
    k = ((n & 0x7fffffff) >> 20) - 0x3ff;
    t = mkd((uint32_t)(n & ~(0x000fffff >> k)), 0u);
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
  if ((j >> 20) <= 0) z = __builtin_ldexp(z, n);
  else z = mkd((uint32_t)j, low(z));
  return z;
}
static __host__ __device__ __noinline__ double js_pow_pos(double x, double y) { return js_pow_core(x, y); }
// x**(4/3) (Manning's depth power), out of line: inlined, its registers would
// add to the stage kernel's peak
static __device__ __noinline__ double js_pow43(double x) { return js_pow_core(x, 4.0 / 3.0); }

}  // namespace swe
