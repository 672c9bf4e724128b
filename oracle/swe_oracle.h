/*
 * CPU oracle for the shallow-water time step: a sequential C restatement of
 * ShallowWater in src/swe.mjs (the reference solver) used ONLY as a checker by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 * Nothing in the product path may link or call it.
 *
 * Every routine follows the reference loop by loop and operation by operation
 * (left-to-right association, divisions kept as divisions, no FMA contraction:
 * build with -ffp-contract=off), so that on identical inputs it reproduces the
 * reference's doubles bit for bit.  The file:line of the reference routine each
 * function restates is given at its definition in swe_oracle.c.
 *
 * Two JS engine builtins are restated, not called:
 *   Math.hypot  -> orc_hypot: V8's normalise-by-max Kahan sum (V8 7.8 / Node
 *                  12.22.9, builtins math.tq MathHypot). Pinned bit-exact
 *                  against Node on 2e5 random pairs (tests/test_oracle_math.py).
 *   Math.pow    -> orc_pow:   fdlibm e_pow (Sun, 1993) general path. Node
 *                  12.22.9's Math.pow agrees with it to 1 ulp but not bit for
 *                  bit (about 4.5% of arguments differ by 1 ulp), so friction
 *                  results are pinned to a tolerance, not to the bit.
 */
#ifndef SWE_ORACLE_H
#define SWE_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* field ids (shared numbering with the HIP library's swe_field) */
enum {
  ORC_B = 0, ORC_H, ORC_HU, ORC_HV, ORC_MANNING, ORC_RH, ORC_RHU, ORC_RHV,
  ORC_THETA, ORC_SPONGE, ORC_NFIELDS_PUBLIC,
  /* internal work arrays, exposed read-only for stage-by-stage debugging */
  ORC_H0 = 16, ORC_HU0, ORC_HV0, ORC_U, ORC_V,
  ORC_FXM, ORC_FXN, ORC_FXT, ORC_FYM, ORC_FYN, ORC_FYT,
  ORC_PXL, ORC_PXR, ORC_PYL, ORC_PYR,
  ORC_SE = 40, ORC_SB, ORC_SU, ORC_SV /* 2N, interleaved [2k+d] */
};
/* geometry row ids (length H each) */
enum {
  ORC_ROW_DX = 0, ORC_ROW_DYN, ORC_ROW_DYS, ORC_ROW_DYCFL, ORC_ROW_BEDPHI,
  ORC_ROW_TANPHI, ORC_ROW_F, ORC_ROW_GEO, ORC_ROW_AREA, ORC_NROWS
};
enum { ORC_BC_REFLECT = 0, ORC_BC_OUTFLOW = 1, ORC_BC_PERIODIC = 2, ORC_BC_FLATHER = 3 };
enum { ORC_LIM_MINMOD = 0, ORC_LIM_MC = 1, ORC_LIM_CENTRAL = 2 };

typedef struct orc_params {
  int32_t nx, ny;        /* interior cells */
  int32_t sphere;        /* 0 Cartesian, 1 latitude-longitude */
  int32_t order;         /* 1 or 2 */
  int32_t limiter;       /* ORC_LIM_* */
  int32_t use_manning_field;
  double g;              /* gravity; the reference hard-codes 9.80665 */
  double min_depth;
  double eps4;           /* max(minDepth,1e-12)**4 exactly as the engine computed it */
  double cfl;
  double coriolis;       /* Cartesian f */
  double manning;        /* scalar n */
  double R;              /* sphere radius (Infinity for Cartesian) */
  double dx, dy;         /* Cartesian cell size (volume()) */
} orc_params;

typedef struct orc_sim orc_sim;

orc_sim *orc_create(const orc_params *p, const double *const rows[ORC_NROWS]);
void     orc_destroy(orc_sim *s);
double  *orc_field(orc_sim *s, int which);     /* N (or 2N for slopes) */
double  *orc_row(orc_sim *s, int which);       /* H */
int      orc_W(const orc_sim *s);
int      orc_H(const orc_sim *s);
void     orc_set_bc(orc_sim *s, int side, int kind);
void     orc_set_time(orc_sim *s, double t, int64_t steps);
double   orc_time(const orc_sim *s);
double   orc_mass_floored(const orc_sim *s);

void   orc_apply_bc(orc_sim *s, const double ext[8]);
void   orc_residual(orc_sim *s, double dt);
double orc_max_dt(orc_sim *s);
double orc_dry_clean(orc_sim *s);
void   orc_apply_friction(orc_sim *s, double dt);
/* dt_in NaN means "self-timed" (the reference's step() with no argument). */
double orc_step(orc_sim *s, double dt_in, const double ext[8]);
void   orc_sponge(orc_sim *s, double dt, double eta_ref);
/* nsteps self-timed steps with a constant ext; returns steps taken */
int64_t orc_run(orc_sim *s, int64_t nsteps, const double ext[8]);

double orc_volume(orc_sim *s);
double orc_max_speed(orc_sim *s);
double orc_max_surface_deviation(orc_sim *s, double ref);
double orc_energy(orc_sim *s);
int    orc_finite(orc_sim *s);

/* scalar pieces, exported for direct unit probes */
void   orc_hllc(double out[3], double hL, double huL, double hvL,
                double hR, double huR, double hvR, double dry, double g);
double orc_vel(double hq, double h, double eps4);
double orc_pow(double x, double y);
double orc_hypot(double a, double b);
double orc_limit(int limiter, double a, double b);
uint64_t orc_fnv1a(const void *bytes, size_t n, uint64_t h);

#ifdef __cplusplus
}
#endif
#endif
