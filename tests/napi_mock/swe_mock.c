/* A stand-in for libswe_amd.so with the same exported C ABI (include/swe_amd.h)
 * and no GPU behind it, so that the N-API addon's argument checks can be tested
 * on a machine without one: node loads the addon with this library first on
 * LD_LIBRARY_PATH.  Every entry point counts its calls and succeeds; the count
 * reaches JS through swe_build_info() ("mock <calls>"), so a test can tell that
 * an array the addon refused never reached the library. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/swe_amd.h"

static long long calls;
static char info[64];

struct swe_ctx { int nx, ny; };
struct swe_adj { int dummy; };

#define COUNT() (++calls)

const char *swe_build_info(void) {
  snprintf(info, sizeof info, "mock %lld", calls);
  return info;
}
const char *swe_last_error(const swe_ctx *c) { (void)c; return "mock"; }
const char *swe_adj_last_error(const swe_adj *a) { (void)a; return "mock"; }
int swe_create(const swe_desc *d, swe_ctx **out) {
  COUNT();
  swe_ctx *c = (swe_ctx *)calloc(1, sizeof *c);
  c->nx = d->nx; c->ny = d->ny;
  *out = c;
  return SWE_OK;
}
void swe_destroy(swe_ctx *c) { free(c); }
double swe_js_pow(double x, double y) { COUNT(); (void)y; return x; }
void swe_hllc(double out[3], double a, double b, double c, double d, double e, double f, double g, double h) {
  COUNT(); (void)a; (void)b; (void)c; (void)d; (void)e; (void)f; (void)g; (void)h;
  out[0] = out[1] = out[2] = 0.0;
}
int swe_host_alloc(size_t bytes, void **out) { *out = calloc(1, bytes ? bytes : 1); return *out ? SWE_OK : SWE_ENOMEM; }
void swe_host_free(void *p) { free(p); }
int swe_set_field(swe_ctx *c, int f, const double *h) { (void)c; (void)f; (void)h; COUNT(); return SWE_OK; }
int swe_get_field(swe_ctx *c, int f, double *h) { (void)c; (void)f; (void)h; COUNT(); return SWE_OK; }
int swe_get_ghosts(swe_ctx *c, int f, double *h) { (void)c; (void)f; (void)h; COUNT(); return SWE_OK; }
int swe_set_row(swe_ctx *c, int r, const double *h) { (void)c; (void)r; (void)h; COUNT(); return SWE_OK; }
int swe_set_bc(swe_ctx *c, int s, int k) { (void)c; (void)s; (void)k; COUNT(); return SWE_OK; }
int swe_set_scalar(swe_ctx *c, int w, double v) { (void)c; (void)w; (void)v; COUNT(); return SWE_OK; }
int swe_get_scalar(swe_ctx *c, int w, double *v) { (void)c; (void)w; COUNT(); *v = 0.0; return SWE_OK; }
int swe_apply_bc(swe_ctx *c, const double e[8]) { (void)c; (void)e; COUNT(); return SWE_OK; }
int swe_residual(swe_ctx *c, double dt) { (void)c; (void)dt; COUNT(); return SWE_OK; }
int swe_apply_friction(swe_ctx *c, double dt) { (void)c; (void)dt; COUNT(); return SWE_OK; }
int swe_dry_clean(swe_ctx *c, double *a) { (void)c; COUNT(); *a = 0.0; return SWE_OK; }
int swe_max_dt(swe_ctx *c, double *o) { (void)c; COUNT(); *o = 1.0; return SWE_OK; }
int swe_step(swe_ctx *c, double dt, const double e[8], int wt, int wr, double *t, double *m) {
  (void)c; (void)e; (void)wt; (void)wr; COUNT(); *t = dt; *m = 0.0; return SWE_OK;
}
int swe_step_begin(swe_ctx *c, double dt, const double e[8], double *t) { (void)c; (void)e; COUNT(); *t = dt; return SWE_OK; }
int swe_step_end(swe_ctx *c, const double e[8], int wt, int wr, double *m) {
  (void)c; (void)e; (void)wt; (void)wr; COUNT(); *m = 0.0; return SWE_OK;
}
int swe_keep_residual(swe_ctx *c, int on) { (void)c; (void)on; COUNT(); return SWE_OK; }
int swe_get_ext(swe_ctx *c, double e[8]) { (void)c; COUNT(); memset(e, 0, 8 * sizeof(double)); return SWE_OK; }
int swe_advance(swe_ctx *c, int64_t n, int m, double v, double te, int64_t *t) {
  (void)c; (void)m; (void)v; (void)te; COUNT(); *t = n; return SWE_OK;
}
int swe_set_tide(swe_ctx *c, int n, const int32_t *m, const double *a, const double *o, const double *p,
                 const double *la, const double *lo) {
  (void)c; (void)n; (void)m; (void)a; (void)o; (void)p; (void)la; (void)lo; COUNT(); return SWE_OK;
}
int swe_set_harmonic(swe_ctx *c, int s, int ne, double msl, const double *a, const double *o, const double *p, int nu,
                     double u0, const double *ua, const double *uo, const double *up) {
  (void)c; (void)s; (void)ne; (void)msl; (void)a; (void)o; (void)p; (void)nu; (void)u0; (void)ua; (void)uo; (void)up;
  COUNT(); return SWE_OK;
}
int swe_set_harmonic_ramp(swe_ctx *c, int s, double r, int f, double u) {
  (void)c; (void)s; (void)r; (void)f; (void)u; COUNT(); return SWE_OK;
}
int swe_diagnostics(swe_ctx *c, double r, double o[SWE_D_COUNT]) { (void)c; (void)r; (void)o; COUNT(); return SWE_OK; }
/* a recorder and a projection that always report 5 samples */
int swe_set_recorder(swe_ctx *c, int n, const int32_t *cells, int e, double t0, double dt, int64_t cap) {
  (void)c; (void)n; (void)cells; (void)e; (void)t0; (void)dt; (void)cap; COUNT(); return SWE_OK;
}
int swe_get_record(swe_ctx *c, double *eta, double *t, int64_t *n) { (void)c; (void)eta; (void)t; COUNT(); *n = 5; return SWE_OK; }
int swe_set_projection(swe_ctx *c, const double *coef, int64_t cap) { (void)c; (void)coef; (void)cap; COUNT(); return SWE_OK; }
int swe_get_projection(swe_ctx *c, double *v, double *t, int64_t *n) { (void)c; (void)v; (void)t; COUNT(); *n = 5; return SWE_OK; }
int swe_adj_create(swe_ctx *c, int64_t n, int k, swe_adj **out) {
  (void)c; (void)n; (void)k; COUNT(); *out = (swe_adj *)calloc(1, sizeof(swe_adj)); return SWE_OK;
}
int swe_adj_create_levels(swe_ctx *c, int64_t n, int64_t i, int64_t o, swe_adj **out) {
  (void)c; (void)n; (void)i; (void)o; COUNT(); *out = (swe_adj *)calloc(1, sizeof(swe_adj)); return SWE_OK;
}
void swe_adj_destroy(swe_adj *a) { free(a); }
int swe_adj_set_gauges(swe_adj *a, int n, const int32_t *c, int e, const double *o) {
  (void)a; (void)n; (void)c; (void)e; (void)o; COUNT(); return SWE_OK;
}
int swe_adj_forward(swe_adj *a, const double *h0, const double *nf, double *J) { (void)a; (void)h0; (void)nf; COUNT(); *J = 0.0; return SWE_OK; }
int swe_adj_loss(swe_adj *a, const double *h0, const double *nf, double *J) { (void)a; (void)h0; (void)nf; COUNT(); *J = 0.0; return SWE_OK; }
int swe_adj_backward(swe_adj *a, double *gh, double *gn) { (void)a; (void)gh; (void)gn; COUNT(); return SWE_OK; }
int swe_adj_get_series(swe_adj *a, double *e) { (void)a; (void)e; COUNT(); return SWE_OK; }
int swe_adj_timing(swe_adj *a, double *f, double *b) { (void)a; COUNT(); *f = *b = 0.0; return SWE_OK; }
