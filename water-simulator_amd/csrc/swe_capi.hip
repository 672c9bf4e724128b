// C ABI (include/swe_amd.h) over the HIP kernels of swe_kernels.hip.
// A context owns the device buffers of one basin and one HIP stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "swe_ctx.hpp"
#include "swe_device.hpp"

using namespace swe;

static thread_local std::string g_err;

static int fail(swe_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg; else g_err = msg;
  return code;
}

// entry points that run kernels on the current state wait for a split step to end
#define MIDCHK(c)                                                                                   \
  do {                                                                                              \
    if ((c)->mid) return fail((c), SWE_ESTATE, "a step is in progress (swe_step_begin without swe_step_end)"); \
  } while (0)

static void refresh_params(swe_ctx *c) {
  Params &P = c->P;
  P.friction = c->mfield ? 2 : ((c->manning != 0.0 && c->manning == c->manning) ? 1 : 0);
  P.manning = c->manning;
  P.coriolis = c->coriolis;
  P.coriolis_on = (c->coriolis != 0.0 && c->coriolis == c->coriolis) ? 1 : 0;
  P.sponge = c->sponge ? 1 : 0;
  P.tide = c->tide;
  P.tide_n = c->tide_n;
  Geo &G = P.geo;
  G.dx = c->rows[R_DX]; G.rdx = c->rows[R_RDX]; G.dyN = c->rows[R_DYN]; G.rdyN = c->rows[R_RDYN];
  G.dyS = c->rows[R_DYS]; G.rdyS = c->rows[R_RDYS]; G.dyCFL = c->rows[R_DYCFL]; G.rdyCFL = c->rows[R_RDYCFL];
  G.bedPhi = c->rows[R_BED]; G.rbedPhi = c->rows[R_RBED]; G.tanPhi = c->rows[R_TAN]; G.fRow = c->rows[R_F];
  G.geo = c->rows[R_GEO];
  P.geoA = c->rowsA;
}

static void drop_graphs(swe_ctx *c) {
  for (int i = 0; i < 2; ++i) {
    if (c->gexec[i]) { (void)hipGraphExecDestroy(c->gexec[i]); c->gexec[i] = nullptr; }
    if (c->graph[i]) { (void)hipGraphDestroy(c->graph[i]); c->graph[i] = nullptr; }
  }
}

static int upload_row(swe_ctx *c, int row, const double *host) {
  const int H = c->P.H;
  std::vector<double> v(host, host + H);
  c->hrows[row] = v;
  auto put = [&](int slot, const std::vector<double> &x) -> int {
    HIPCHK(c, hipMemcpyAsync(c->rows[slot], x.data(), sizeof(double) * H, hipMemcpyHostToDevice, c->stream));
    return SWE_OK;
  };
  std::vector<double> r(H);
  for (int j = 0; j < H; ++j) r[j] = 1.0 / v[j];          // correctly rounded reciprocals (1/inf = 0)
  int rc = SWE_OK;
  switch (row) {
    case SWE_ROW_DX: rc = put(R_DX, v); if (!rc) rc = put(R_RDX, r); break;
    case SWE_ROW_DYN: rc = put(R_DYN, v); if (!rc) rc = put(R_RDYN, r); break;
    case SWE_ROW_DYS: rc = put(R_DYS, v); if (!rc) rc = put(R_RDYS, r); break;
    case SWE_ROW_DYCFL: rc = put(R_DYCFL, v); if (!rc) rc = put(R_RDYCFL, r); break;
    case SWE_ROW_BEDPHI: rc = put(R_BED, v); if (!rc) rc = put(R_RBED, r); break;
    case SWE_ROW_TANPHI: rc = put(R_TAN, v); break;
    case SWE_ROW_F: rc = put(R_F, v); break;
    case SWE_ROW_GEO: rc = put(R_GEO, v); break;
    case SWE_ROW_AREA: rc = put(R_AREA, v); break;
    default: return fail(c, SWE_EINVAL, "bad row id");
  }
  if (rc) return rc;
  // the interleaved copy the stage kernels read: every array's value at row j
  // (reciprocals of the divisors as above; rows not set yet hold zeros)
  static const int src_of[R_COUNT] = {SWE_ROW_DX, SWE_ROW_DX, SWE_ROW_DYN, SWE_ROW_DYN, SWE_ROW_DYS,
                                      SWE_ROW_DYS, SWE_ROW_DYCFL, SWE_ROW_DYCFL, SWE_ROW_BEDPHI, SWE_ROW_BEDPHI,
                                      SWE_ROW_TANPHI, SWE_ROW_F, SWE_ROW_GEO, SWE_ROW_AREA};
  static const bool recip[R_COUNT] = {false, true, false, true, false, true, false, true, false, true,
                                      false, false, false, false};
  std::vector<double> a((size_t)H * GEO_AOS, 0.0);
  for (int k = 0; k < R_COUNT; ++k) {
    const std::vector<double> &src = c->hrows[src_of[k]];
    if (src.empty()) continue;
    for (int j = 0; j < H; ++j) a[(size_t)j * GEO_AOS + k] = recip[k] ? 1.0 / src[j] : src[j];
  }
  for (int j = 0; j < H; ++j) a[(size_t)j * GEO_AOS + GEO_TANR] = a[(size_t)j * GEO_AOS + R_TAN] / c->P.R;
  // a captured step graph holds the stage-kernel instance chosen for the old rows
  const int was_uni = c->P.geo_uni;
  set_uniform_rows(c->P, a.data(), H);
  if (c->P.geo_uni != was_uni) drop_graphs(c);
  HIPCHK(c, hipMemcpyAsync(c->rowsA, a.data(), sizeof(double) * a.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

static int default_chunk(const Params &P) {
  const int strips = (P.W + SWEEP_OUT - 1) / SWEEP_OUT;
  // about two rounds of the 256 CU x 4 SIMD x 3 resident wavefronts the
  // stage kernels' registers allow, and the shortest chunks on grids too small
  // to fill them (each wave marches its rows one after another, so a step of a
  // small grid takes time in proportion to the chunk height); swe_autotune
  // refines this on the device
  const int target_waves = 6144;
  int chunks = (target_waves + strips - 1) / strips;
  if (chunks < 1) chunks = 1;
  int rows = (P.H + chunks - 1) / chunks;
  if (rows > P.H) rows = P.H;
  if (rows < DEFAULT_MIN_CHUNK) rows = DEFAULT_MIN_CHUNK;
  return rows;
}

extern "C" {

double swe_js_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  if (x != x || y != y) return x + y;
  if (y == 1.0) return x;
  if (y == 2.0) return x * x;
  if (y == 0.5 && x >= 0.0) return std::sqrt(x);
  if (!(x > 0.0) || std::isinf(x) || std::isinf(y)) return std::pow(x, y);
  return swe::js_pow_pos(x, y);
}

void swe_hllc(double out[3], double hL, double huL, double hvL, double hR, double huR, double hvR,
              double dry, double g) {
  const HllcConst C = {g, 0.5 * g, 1.0 / g, dry, g * dry >= SQRT_N_MIN};
  const Flux3 f = swe::hllc(hL, huL, hvL, hR, huR, hvR, C);
  out[0] = f.m; out[1] = f.n; out[2] = f.t;
}

const char *swe_build_info(void) { return "water-simulator_amd 0.1 gfx950 fp64"; }

int swe_host_alloc(size_t bytes, void **out) {
  if (!out) return fail(nullptr, SWE_EINVAL, "null output pointer");
  *out = nullptr;
  const hipError_t e = hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault);
  if (e != hipSuccess) {
    *out = nullptr;
    return fail(nullptr, e == hipErrorOutOfMemory ? SWE_ENOMEM : SWE_EHIP,
                std::string("hipHostMalloc: ") + hipGetErrorString(e));
  }
  return SWE_OK;
}

void swe_host_free(void *p) {
  if (p) (void)hipHostFree(p);
}

const char *swe_last_error(const swe_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int swe_create(const swe_desc *d, swe_ctx **out) {
  if (!d || !out) return fail(nullptr, SWE_EINVAL, "null argument");
  *out = nullptr;
  if (d->nx < 2 || d->ny < 1) return fail(nullptr, SWE_EINVAL, "nx must be >= 2 and ny >= 1");
  if (d->order != 1 && d->order != 2) return fail(nullptr, SWE_EINVAL, "order must be 1 or 2");
  if (d->limiter < 0 || d->limiter > 2) return fail(nullptr, SWE_EINVAL, "unknown limiter");
  if (d->kind != SWE_CARTESIAN && d->kind != SWE_SPHERE) return fail(nullptr, SWE_EINVAL, "unknown kind");
  for (int r = 0; r < SWE_NROWS; ++r)
    if (!d->rows[r] && r != SWE_ROW_AREA) return fail(nullptr, SWE_EINVAL, "missing geometry row");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, SWE_EHIP, "no HIP device");
  swe_ctx *c = new swe_ctx();
  Params &P = c->P;
  P.nx = d->nx; P.ny = d->ny; P.W = d->nx + 4; P.H = d->ny + 4;
  P.sphere = d->kind == SWE_SPHERE; P.order = d->order; P.limiter = d->limiter;
  for (int s = 0; s < 4; ++s) P.bc[s] = P.sphere ? (s < 2 ? BC_PERIODIC : BC_REFLECT) : BC_REFLECT;
  P.g = d->g; P.half_g = 0.5 * d->g; P.neg_half_g = -d->g * 0.5; P.rg = 1.0 / d->g;
  P.min_depth = d->min_depth; P.eps4 = d->eps4; P.cfl = d->cfl; P.R = d->R;
  c->manning = d->manning; c->coriolis = d->coriolis;
  c->N = (size_t)P.W * P.H;
  const size_t bytes = c->N * sizeof(double);
  auto bail = [&](const char *what, hipError_t e) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    swe_destroy(c);
    return fail(nullptr, e == hipErrorOutOfMemory ? SWE_ENOMEM : SWE_EHIP, m);
  };
  hipError_t e;
  if ((e = hipGetDevice(&c->device)) != hipSuccess) return bail("hipGetDevice", e);
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return bail("stream", e);
  if ((e = hipMalloc(&c->b, bytes)) != hipSuccess) return bail("hipMalloc b", e);
  // each state buffer's h, hu, hv contiguous, and all rows in one block: the
  // stage kernels' layout contract (Params::fstride, rstride)
  for (int i = 0; i < 3; ++i) {
    if ((e = hipMalloc(&c->st[i][0], 3 * bytes)) != hipSuccess) return bail("hipMalloc state", e);
    c->st[i][1] = c->st[i][0] + c->N;
    c->st[i][2] = c->st[i][0] + 2 * c->N;
  }
  if ((e = hipMalloc(&c->rows[0], sizeof(double) * P.H * R_COUNT)) != hipSuccess) return bail("hipMalloc rows", e);
  for (int r = 1; r < R_COUNT; ++r) c->rows[r] = c->rows[0] + (size_t)r * P.H;
  if ((e = hipMalloc(&c->rowsA, sizeof(double) * P.H * GEO_AOS)) != hipSuccess) return bail("hipMalloc rows", e);
  P.fstride = (int)c->N;
  P.rstride = P.H;
  if ((e = hipMalloc(&c->dctl, sizeof(Ctl))) != hipSuccess) return bail("hipMalloc ctl", e);
  if ((e = hipHostMalloc((void **)&c->hctl, sizeof(Ctl), hipHostMallocDefault)) != hipSuccess) return bail("pinned ctl", e);
  const int strips = (P.W + SWEEP_OUT - 1) / SWEEP_OUT;
  c->max_parts = strips * ((P.H + MIN_CHUNK - 1) / MIN_CHUNK);
  if ((e = hipMalloc(&c->mf_part, sizeof(double) * 2 * c->max_parts)) != hipSuccess) return bail("hipMalloc part", e);
  if ((e = hipEventCreate(&c->ev0)) != hipSuccess) return bail("event", e);
  if ((e = hipEventCreate(&c->ev1)) != hipSuccess) return bail("event", e);
  for (int f = 0; f < 3; ++f)
    for (int i = 0; i < 3; ++i) (void)hipMemsetAsync(c->st[i][f], 0, bytes, c->stream);
  (void)hipMemsetAsync(c->b, 0, bytes, c->stream);
  std::memset(c->hctl, 0, sizeof(Ctl));
  c->hctl->dt_mode = DT_HOST;
  if ((e = hipMemcpyAsync(c->dctl, c->hctl, sizeof(Ctl), hipMemcpyHostToDevice, c->stream)) != hipSuccess)
    return bail("ctl upload", e);
  std::vector<double> zero(P.H, 0.0);
  for (int r = 0; r < SWE_NROWS; ++r) {
    const double *src = d->rows[r] ? d->rows[r] : zero.data();
    int rc = upload_row(c, r, src);
    if (rc) { std::string m = c->err; swe_destroy(c); return fail(nullptr, rc, m); }
  }
  c->chunk = default_chunk(P);
  refresh_params(c);
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail("sync", e);
  *out = c;
  return SWE_OK;
}

void swe_destroy(swe_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graphs(c);
  (void)hipFree(c->b);
  for (auto &s : c->st) (void)hipFree(s[0]);
  (void)hipFree(c->rows[0]);
  (void)hipFree(c->rowsA);
  (void)hipFree(c->mfield); (void)hipFree(c->sponge); (void)hipFree(c->tide);
  (void)hipFree(c->rh); (void)hipFree(c->theta);
  (void)hipFree(c->dctl); (void)hipFree(c->mf_part);
  (void)hipFree(c->rec_cells); (void)hipFree(c->rec_series); (void)hipFree(c->rec_times);
  (void)hipFree(c->proj_coef); (void)hipFree(c->proj_vals); (void)hipFree(c->proj_times);
  (void)hipFree(c->diag);
  if (c->hctl) (void)hipHostFree(c->hctl);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

static int ensure_probe(swe_ctx *c) {
  const size_t bytes = c->N * sizeof(double);
  if (!c->rh) {                       // rh, rhu, rhv contiguous (Params::fstride)
    HIPCHK(c, hipMalloc(&c->rh, 3 * bytes));
    HIPCHK(c, hipMemsetAsync(c->rh, 0, 3 * bytes, c->stream));
    c->rhu = c->rh + c->N;
    c->rhv = c->rh + 2 * c->N;
  }
  if (!c->theta) {
    HIPCHK(c, hipMalloc(&c->theta, bytes));
    HIPCHK(c, hipMemsetAsync(c->theta, 0, bytes, c->stream));
  }
  return SWE_OK;
}

// the state the fields name: the current one, or mid-step the stage-1 state
static int cur_buf(const swe_ctx *c) { return c->mid ? c->Bi : c->A; }

static double *field_ptr(swe_ctx *c, int f) {
  switch (f) {
    case SWE_F_B: return c->b;
    case SWE_F_H: return c->st[cur_buf(c)][0];
    case SWE_F_HU: return c->st[cur_buf(c)][1];
    case SWE_F_HV: return c->st[cur_buf(c)][2];
    case SWE_F_MANNING: return c->mfield;
    case SWE_F_SPONGE: return c->sponge;
    case SWE_F_RH: return c->rh;
    case SWE_F_RHU: return c->rhu;
    case SWE_F_RHV: return c->rhv;
    case SWE_F_THETA: return c->theta;
    default: return nullptr;
  }
}

int swe_set_field(swe_ctx *c, int f, const double *host) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  const size_t bytes = c->N * sizeof(double);
  if (f == SWE_F_MANNING || f == SWE_F_SPONGE) {
    double **p = f == SWE_F_MANNING ? &c->mfield : &c->sponge;
    if (!host) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      HIPCHK(c, hipFree(*p));
      *p = nullptr;
      refresh_params(c);
      drop_graphs(c);
      return SWE_OK;
    }
    if (f == SWE_F_MANNING)
      for (size_t k = 0; k < c->N; ++k)
        if (!(host[k] >= 0) || !std::isfinite(host[k]))
          return fail(c, SWE_EINVAL, "setManningField: Manning n must be finite and >= 0");
    if (!*p) { HIPCHK(c, hipMalloc(p, bytes)); drop_graphs(c); }
    HIPCHK(c, hipMemcpyAsync(*p, host, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    refresh_params(c);
    return SWE_OK;
  }
  if (f < SWE_F_B || f > SWE_F_HV) return fail(c, SWE_EINVAL, "field is not writable");
  if (!host) return fail(c, SWE_EINVAL, "null host pointer");
  HIPCHK(c, hipMemcpyAsync(field_ptr(c, f), host, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_get_field(swe_ctx *c, int f, double *host) {
  if (!c || !host) return fail(c, SWE_EINVAL, "null argument");
  double *p = field_ptr(c, f);
  if (!p) return fail(c, SWE_ESTATE, "field not available (never computed or not set)");
  HIPCHK(c, hipMemcpyAsync(host, p, c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_set_row(swe_ctx *c, int row, const double *host) {
  if (!c || !host) return fail(c, SWE_EINVAL, "null argument");
  if (row < 0 || row >= SWE_NROWS) return fail(c, SWE_EINVAL, "bad row id");
  return upload_row(c, row, host);
}

int swe_set_bc(swe_ctx *c, int side, int kind) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  if (side < 0 || side > 3 || kind < 0 || kind > SWE_BC_STORED) return fail(c, SWE_EINVAL, "bad side or kind");
  if (c->P.bc[side] != kind) drop_graphs(c);
  c->P.bc[side] = kind;
  return SWE_OK;
}

static int pull_ctl(swe_ctx *c);

int swe_set_scalar(swe_ctx *c, int which, double v) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  if (which == SWE_S_T || which == SWE_S_STEPS) {
    // the rest of the control block (counters the device advanced) is kept
    int rc = pull_ctl(c);
    if (rc) return rc;
  }
  switch (which) {
    case SWE_S_MANNING: c->manning = v; break;
    case SWE_S_CORIOLIS: c->coriolis = v; break;
    case SWE_S_CFL: c->P.cfl = v; break;
    case SWE_S_SPONGE_REF: c->P.sponge_ref = v; break;
    case SWE_S_T: c->hctl->t = v; break;
    case SWE_S_STEPS: c->hctl->steps = (long long)v; break;
    default: return fail(c, SWE_EINVAL, "scalar is read-only or unknown");
  }
  refresh_params(c);
  drop_graphs(c);
  if (which == SWE_S_T || which == SWE_S_STEPS) {
    HIPCHK(c, hipMemcpyAsync(c->dctl, c->hctl, sizeof(Ctl), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return SWE_OK;
}

static int pull_ctl(swe_ctx *c) {
  HIPCHK(c, hipMemcpyAsync(c->hctl, c->dctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_get_scalar(swe_ctx *c, int which, double *v) {
  if (!c || !v) return fail(c, SWE_EINVAL, "null argument");
  int rc = pull_ctl(c);
  if (rc) return rc;
  switch (which) {
    case SWE_S_MANNING: *v = c->manning; break;
    case SWE_S_CORIOLIS: *v = c->coriolis; break;
    case SWE_S_CFL: *v = c->P.cfl; break;
    case SWE_S_SPONGE_REF: *v = c->P.sponge_ref; break;
    case SWE_S_T: *v = c->hctl->t; break;
    case SWE_S_STEPS: *v = (double)c->hctl->steps; break;
    case SWE_S_MASS_FLOORED: *v = c->hctl->mf; break;
    case SWE_S_DT_NEXT: *v = c->hctl->dt_next; break;
    case SWE_S_NONFINITE: *v = c->hctl->nonfinite; break;
    case SWE_S_CHUNK_ROWS: *v = c->chunk; break;
    case SWE_S_DIAG_MS: *v = c->diag_ms; break;
    default: return fail(c, SWE_EINVAL, "unknown scalar");
  }
  return SWE_OK;
}

static int push_ctl(swe_ctx *c) {
  HIPCHK(c, hipMemcpyAsync(c->dctl, c->hctl, sizeof(Ctl), hipMemcpyHostToDevice, c->stream));
  return SWE_OK;
}

int swe_apply_bc(swe_ctx *c, const double ext[8]) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  MIDCHK(c);
  int rc = pull_ctl(c);
  if (rc) return rc;
  for (int i = 0; i < 8; ++i) c->hctl->ext[i] = ext ? ext[i] : 0.0;
  if ((rc = push_ctl(c))) return rc;
  StateW S = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
  HIPCHK(c, launch_apply_bc(c->P, S, c->b, c->dctl, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

static SweepArgs base_args(swe_ctx *c) {
  SweepArgs a{};
  a.P = c->P;
  a.b = c->b;
  a.mfield = c->mfield;
  a.sponge = c->sponge;
  a.ctl = c->dctl;
  a.mf_part = c->mf_part;
  a.rows_per_chunk = c->chunk;
  return a;
}

int swe_residual(swe_ctx *c, double dt) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  MIDCHK(c);
  int rc = ensure_probe(c);
  if (rc) return rc;
  if ((rc = pull_ctl(c))) return rc;
  c->hctl->dt = dt;
  if ((rc = push_ctl(c))) return rc;
  if (c->tide_n) HIPCHK(c, launch_tide_times(c->dctl, c->stream));
  SweepArgs a = base_args(c);
  for (int s = 0; s < 4; ++s) a.P.bc[s] = BC_STORED;      // ghosts as the caller left them
  a.in = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
  a.rh = c->rh; a.rhu = c->rhu; a.rhv = c->rhv; a.theta = c->theta;
  HIPCHK(c, launch_sweep(MODE_PROBE, a, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_apply_friction(swe_ctx *c, double dt) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  MIDCHK(c);
  if (c->P.friction == 0) return SWE_OK;                  // `if (!n) return;` (swe.mjs:689)
  StateW S = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
  HIPCHK(c, launch_friction(c->P, S, c->mfield, dt, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_dry_clean(swe_ctx *c, double *added) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  MIDCHK(c);
  int blocks = (c->P.nx * c->P.ny + 255) / 256;
  if (blocks > 2 * c->max_parts) blocks = 2 * c->max_parts;
  if (blocks < 1) blocks = 1;
  StateW S = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
  HIPCHK(c, launch_dry_clean(c->P, S, c->mf_part, blocks, c->stream));
  std::vector<double> part(blocks);
  HIPCHK(c, hipMemcpyAsync(part.data(), c->mf_part, sizeof(double) * blocks, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double s = 0.0;
  for (double p : part) s += p;
  if (added) *added = s;
  return SWE_OK;
}

int swe_max_dt(swe_ctx *c, double *out) {
  if (!c || !out) return fail(c, SWE_EINVAL, "null argument");
  MIDCHK(c);
  int rc = pull_ctl(c);
  if (rc) return rc;
  const int saved = c->hctl->dt_mode;
  c->hctl->dt_mode = DT_HOST;
  c->hctl->cfl_bits = 0;
  if ((rc = push_ctl(c))) return rc;
  StateP S = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
  HIPCHK(c, launch_maxdt(c->P, S, c->dctl, c->stream));
  HIPCHK(c, launch_finalize(c->dctl, c->mf_part, 0, c->P.cfl, 0, c->stream));
  if ((rc = pull_ctl(c))) return rc;
  *out = c->hctl->dt_next;
  c->hctl->dt_mode = saved;
  return push_ctl(c);
}

// One step on the device, stream-ordered; the caller handles the ctl block.
// applyBC before each stage (swe.mjs:784,798) writes the ghost ring of that
// stage's input; a refused step still gets its applyBC, as in the reference.
// Stage 1 (swe.mjs:784-796): applyBC on A, then A -> B.
static int enqueue_stage1(swe_ctx *c, int np) {
  if (np > c->max_parts) return fail(c, SWE_EINVAL, "chunk too small");
  StateW SA = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
  StateW SB = {c->st[c->Bi][0], c->st[c->Bi][1], c->st[c->Bi][2]};
  HIPCHK(c, launch_apply_bc(c->P, SA, c->b, c->dctl, c->stream));
  SweepArgs a = base_args(c);
  a.in = {SA.h, SA.hu, SA.hv};
  a.out = SB;
  HIPCHK(c, launch_sweep(MODE_STAGE1, a, c->stream));
  return SWE_OK;
}

// Stage 2 (swe.mjs:798-812, with friction, sponge and the next maxDt) and the
// finalize: applyBC on B, then (A, B) -> C, and C becomes the current state.
// rec: sample the gauge and projection recorders after the step (device-resident runs).
static int enqueue_stage2(swe_ctx *c, int np, int want_theta, int want_rh, bool rec) {
  StateW SA = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
  StateW SB = {c->st[c->Bi][0], c->st[c->Bi][1], c->st[c->Bi][2]};
  HIPCHK(c, launch_apply_bc(c->P, SB, c->b, c->dctl, c->stream));
  SweepArgs b2 = base_args(c);
  b2.in = {SB.h, SB.hu, SB.hv};
  b2.h0 = {SA.h, SA.hu, SA.hv};
  b2.out = {c->st[c->C][0], c->st[c->C][1], c->st[c->C][2]};
  b2.mf_part = c->mf_part + np;
  if (want_theta || want_rh) {
    int rc = ensure_probe(c);
    if (rc) return rc;
    b2.P.want_theta = want_theta ? 1 : 0;
    b2.P.want_rh = want_rh ? 1 : 0;
    b2.theta = c->theta; b2.rh = c->rh; b2.rhu = c->rhu; b2.rhv = c->rhv;
  }
  HIPCHK(c, launch_sweep(MODE_STAGE2, b2, c->stream));
  if (rec && c->rec_n > 0) {
    const RecArgs r = {c->st[c->C][0], c->b, c->rec_cells, c->rec_series, c->rec_times, c->rec_n, c->rec_cap};
    HIPCHK(c, launch_finalize(c->dctl, c->mf_part, np, c->P.cfl, 1, c->stream, &r));
  } else {
    HIPCHK(c, launch_finalize(c->dctl, c->mf_part, np, c->P.cfl, 1, c->stream));
  }
  if (rec && c->proj_cap > 0) HIPCHK(c, launch_projection(c, c->st[c->C][0], c->b));
  std::swap(c->A, c->C);
  return SWE_OK;
}

static int enqueue_step(swe_ctx *c, int want_theta, int want_rh, bool rec = false) {
  const int np = sweep_waves(c->P, c->chunk);
  int rc = enqueue_stage1(c, np);
  if (rc) return rc;
  return enqueue_stage2(c, np, want_theta, want_rh, rec);
}

// Control block for a step taken by the host's call (swe_step, swe_step_begin):
// the caller's Flather values, and dt from the argument or (NaN) from maxDt.
static int begin_host_step(swe_ctx *c, double dt, const double ext[8]) {
  int rc = pull_ctl(c);
  if (rc) return rc;
  Ctl *h = c->hctl;
  for (int i = 0; i < 8; ++i) h->ext[i] = ext ? ext[i] : 0.0;
  h->harm_on = 0;          // this step uses the caller's values; swe_advance restores the harmonics
  h->cfl_bits = 0;
  const bool self = std::isnan(dt);
  h->dt_mode = self ? DT_SELF : DT_HOST;
  if (!self) h->dt = dt;
  if ((rc = push_ctl(c))) return rc;
  if (!self && c->tide_n) HIPCHK(c, launch_tide_times(c->dctl, c->stream));
  if (self) {
    StateP S = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
    HIPCHK(c, launch_maxdt(c->P, S, c->dctl, c->stream));
    HIPCHK(c, launch_finalize(c->dctl, c->mf_part, 0, c->P.cfl, 0, c->stream));
  }
  return SWE_OK;
}

static int end_host_step(swe_ctx *c, double *dt_taken, double *mass_floored) {
  int rc = pull_ctl(c);
  if (rc) return rc;
  Ctl *h = c->hctl;
  if (dt_taken) *dt_taken = h->dt_last;
  if (mass_floored) *mass_floored = h->mf;
  h->dt_mode = DT_HOST;
  return push_ctl(c);
}

int swe_step(swe_ctx *c, double dt, const double ext[8], int want_theta, int want_rh,
             double *dt_taken, double *mass_floored) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  MIDCHK(c);
  for (int s = 0; s < 4; ++s)
    if (c->P.bc[s] == BC_STORED) return fail(c, SWE_EINVAL, "swe_step needs device boundary kinds on every side");
  int rc = begin_host_step(c, dt, ext);
  if (rc) return rc;
  if ((rc = enqueue_step(c, want_theta || c->keep_resid, want_rh || c->keep_resid))) return rc;
  return end_host_step(c, dt_taken, mass_floored);
}

int swe_step_begin(swe_ctx *c, double dt, const double ext[8], double *dt_taken) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  MIDCHK(c);
  int rc = begin_host_step(c, dt, ext);
  if (rc) return rc;
  const int np = sweep_waves(c->P, c->chunk);
  if ((rc = enqueue_stage1(c, np))) return rc;
  if ((rc = pull_ctl(c))) return rc;
  const double d = c->hctl->dt;
  const bool ok = d > 0 && !std::isinf(d);
  if (dt_taken) *dt_taken = ok ? d : 0.0;
  if (!ok) {                  // refused (swe.mjs:784): the state stays as applyBC left it
    c->hctl->dt_mode = DT_HOST;
    c->hctl->dt_last = 0.0;
    return push_ctl(c);
  }
  c->mid = 1;
  c->mid_np = np;
  return SWE_OK;
}

int swe_step_end(swe_ctx *c, const double ext[8], int want_theta, int want_rh, double *mass_floored) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  if (!c->mid) return fail(c, SWE_ESTATE, "swe_step_end without swe_step_begin");
  int rc = pull_ctl(c);
  if (rc) return rc;
  if (ext)
    for (int i = 0; i < 8; ++i) c->hctl->ext[i] = ext[i];
  if ((rc = push_ctl(c))) return rc;
  c->mid = 0;
  if ((rc = enqueue_stage2(c, c->mid_np, want_theta || c->keep_resid, want_rh || c->keep_resid, false))) return rc;
  return end_host_step(c, nullptr, mass_floored);
}

int swe_keep_residual(swe_ctx *c, int on) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  if (on) {
    int rc = ensure_probe(c);
    if (rc) return rc;
  }
  if ((on != 0) != (c->keep_resid != 0)) drop_graphs(c);
  c->keep_resid = on != 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_get_ext(swe_ctx *c, double ext[8]) {
  if (!c || !ext) return fail(c, SWE_EINVAL, "null argument");
  int rc = pull_ctl(c);
  if (rc) return rc;
  for (int i = 0; i < 8; ++i) ext[i] = c->hctl->ext[i];
  return SWE_OK;
}

int swe_get_ghosts(swe_ctx *c, int f, double *host) {
  if (!c || !host) return fail(c, SWE_EINVAL, "null argument");
  if (f < SWE_F_B || f > SWE_F_HV) return fail(c, SWE_EINVAL, "ghost rings exist for b, h, hu, hv only");
  const double *p = field_ptr(c, f);
  const size_t W = c->P.W, H = c->P.H, row = W * sizeof(double);
  // rows 0-1 and H-2..H-1 whole, then columns 0-1 and W-2..W-1 of rows 2..H-3
  HIPCHK(c, hipMemcpyAsync(host, p, 2 * row, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(host + (H - 2) * W, p + (H - 2) * W, 2 * row, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpy2DAsync(host + 2 * W, row, p + 2 * W, row, 2 * sizeof(double), H - 4,
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpy2DAsync(host + 2 * W + W - 2, row, p + 2 * W + W - 2, row, 2 * sizeof(double), H - 4,
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_set_tide(swe_ctx *c, int n, const int32_t *m, const double *amp, const double *omega, const double *phase,
                 const double *lat, const double *lon) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  if (n < 0 || n > MAX_TIDE) return fail(c, SWE_EINVAL, "tide: 0..4 constituents");
  if (n > 0 && (!c->P.sphere || !(c->P.R > 0) || std::isinf(c->P.R)))
    return fail(c, SWE_EINVAL, "tide: the equilibrium tide needs a spherical grid");
  std::vector<double> T;
  if (n > 0 && (!omega || !phase || tide_table(c->P, n, m, amp, lat, lon, T)))
    return fail(c, SWE_EINVAL, "tide: bad constituents (m must be 1 or 2, amplitudes and angles finite)");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  HIPCHK(c, hipFree(c->tide));
  c->tide = nullptr;
  c->tide_n = 0;
  if (n > 0) {
    HIPCHK(c, hipMalloc(&c->tide, sizeof(double) * T.size()));
    HIPCHK(c, hipMemcpy(c->tide, T.data(), sizeof(double) * T.size(), hipMemcpyHostToDevice));
    c->tide_n = n;
  }
  refresh_params(c);
  int rc = pull_ctl(c);
  if (rc) return rc;
  c->hctl->tide_n = n;
  for (int k = 0; k < n; ++k) { c->hctl->tide_om[k] = omega[k]; c->hctl->tide_ph[k] = phase[k]; }
  if ((rc = push_ctl(c))) return rc;
  if (n) HIPCHK(c, launch_tide_times(c->dctl, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_set_harmonic(swe_ctx *c, int side, int n_eta, double msl, const double *amp, const double *omega,
                     const double *phase, int n_u, double u0, const double *uamp, const double *uomega,
                     const double *uphase) {
  if (!c || side < 0 || side > 3 || n_eta < 0 || n_eta > MAX_HARM || n_u < 0 || n_u > MAX_HARM)
    return fail(c, SWE_EINVAL, "bad harmonic forcing");
  int rc = pull_ctl(c);
  if (rc) return rc;
  Harm &hm = c->hctl->harm[side];
  std::memset(&hm, 0, sizeof(hm));
  hm.n_eta = n_eta; hm.msl = msl; hm.n_u = n_u; hm.u0 = u0;
  for (int k = 0; k < n_eta; ++k) { hm.amp[k] = amp[k]; hm.omega[k] = omega[k]; hm.phase[k] = phase[k]; }
  for (int k = 0; k < n_u; ++k) { hm.uamp[k] = uamp[k]; hm.uomega[k] = uomega[k]; hm.uphase[k] = uphase[k]; }
  c->hctl->harm_on = 1;
  c->harm_enabled = 1;
  rc = push_ctl(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_set_harmonic_ramp(swe_ctx *c, int side, double ramp, int u_from_eta, double u_scale) {
  if (!c || side < 0 || side > 3 || !(ramp >= 0) || !std::isfinite(ramp) || !std::isfinite(u_scale))
    return fail(c, SWE_EINVAL, "bad harmonic ramp");
  int rc = pull_ctl(c);
  if (rc) return rc;
  Harm &hm = c->hctl->harm[side];
  hm.ramp = ramp;
  hm.u_from_eta = u_from_eta ? 1 : 0;
  hm.u_scale = u_scale;
  rc = push_ctl(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_use_graph(swe_ctx *c, int on) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  c->use_graph = on != 0;
  if (!on) drop_graphs(c);
  return SWE_OK;
}

int swe_set_chunk_rows(swe_ctx *c, int rows) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  MIDCHK(c);
  if (rows == 0) rows = default_chunk(c->P);
  if (rows < MIN_CHUNK) return fail(c, SWE_EINVAL, "chunk rows must be >= 1");
  if (rows != c->chunk) drop_graphs(c);
  c->chunk = rows;
  return SWE_OK;
}

// Times stage 1 + stage 2 of the current state for one chunk height without
// touching the state: A is only read, B and C are scratch, t and the step count
// are not advanced (no finalize).
static int time_chunk(swe_ctx *c, int rows, int reps, float *ms) {
  c->chunk = rows;
  const int np = sweep_waves(c->P, rows);
  if (np > c->max_parts) return fail(c, SWE_EINVAL, "chunk too small");
  StateW SA = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
  StateW SB = {c->st[c->Bi][0], c->st[c->Bi][1], c->st[c->Bi][2]};
  SweepArgs a = base_args(c);
  a.in = {SA.h, SA.hu, SA.hv};
  a.out = SB;
  SweepArgs b2 = base_args(c);
  b2.in = {SB.h, SB.hu, SB.hv};
  b2.h0 = {SA.h, SA.hu, SA.hv};
  b2.out = {c->st[c->C][0], c->st[c->C][1], c->st[c->C][2]};
  b2.mf_part = c->mf_part + np;
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < reps; ++r) {
    HIPCHK(c, launch_sweep(MODE_STAGE1, a, c->stream));
    HIPCHK(c, launch_sweep(MODE_STAGE2, b2, c->stream));
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  HIPCHK(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
  return SWE_OK;
}

int swe_autotune(swe_ctx *c, int reps, int *best_rows) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  MIDCHK(c);
  if (reps < 1) reps = 1;
  int rc = pull_ctl(c);
  if (rc) return rc;
  // a valid dt for the trial sweeps: the CFL step of the current state
  const int saved_mode = c->hctl->dt_mode;
  const double saved_dt = c->hctl->dt;
  c->hctl->dt_mode = DT_SELF;
  c->hctl->cfl_bits = 0;
  if ((rc = push_ctl(c))) return rc;
  StateP S = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
  HIPCHK(c, launch_maxdt(c->P, S, c->dctl, c->stream));
  HIPCHK(c, launch_finalize(c->dctl, c->mf_part, 0, c->P.cfl, 0, c->stream));
  StateW SA = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
  HIPCHK(c, launch_apply_bc(c->P, SA, c->b, c->dctl, c->stream));
  // resident wavefront slots: CUs x 4 SIMDs x waves per SIMD the kernel's registers allow
  int cus = 0, per_cu = 0;
  HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep_stage1_ptr(), SWEEP_LANES, 0));
  const int slots = cus * (per_cu > 0 ? per_cu : 4);
  const int strips = (c->P.W + SWEEP_OUT - 1) / SWEEP_OUT;
  std::vector<int> cand = {8, 12, 16, 24, 32, 40, 48, 64, 96};
  // a grid whose 8-row chunks do not fill the resident slots is latency-bound:
  // each wavefront marches its chunk (plus 4 halo rows) one row after another,
  // so shorter chunks finish sooner
  if (sweep_waves(c->P, DEFAULT_MIN_CHUNK) < slots)
    for (int rows : {1, 2, 3, 4, 6}) cand.push_back(rows);
  for (int rounds = 1; rounds <= 4; ++rounds) {
    const int chunks = (rounds * slots + strips - 1) / strips;
    if (chunks >= 1) cand.push_back((c->P.H + chunks - 1) / chunks);
  }
  const int old = c->chunk;
  int best = old;
  float best_ms = 1e30f;
  for (int rows : cand) {
    if (rows < MIN_CHUNK || rows > c->P.H || sweep_waves(c->P, rows) > c->max_parts) continue;
    float ms = 0.f;
    if ((rc = time_chunk(c, rows, 1, &ms))) { c->chunk = old; return rc; }   // warm
    if ((rc = time_chunk(c, rows, reps, &ms))) { c->chunk = old; return rc; }
    if (ms < best_ms) { best_ms = ms; best = rows; }
  }
  c->chunk = best;
  if (best != old) drop_graphs(c);
  if ((rc = pull_ctl(c))) return rc;
  c->hctl->dt_mode = saved_mode;
  c->hctl->dt = saved_dt;
  c->hctl->cfl_bits = 0;
  if ((rc = push_ctl(c))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (best_rows) *best_rows = best;
  return SWE_OK;
}

static int build_graph(swe_ctx *c, int which) {
  // which = parity of the buffer rotation (A index 0 or 2)
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int saveA = c->A, saveC = c->C;
  int rc = enqueue_step(c, c->keep_resid, c->keep_resid, true);
  c->A = saveA; c->C = saveC;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return fail(c, SWE_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
  c->graph[which] = g;
  HIPCHK(c, hipGraphInstantiate(&c->gexec[which], g, nullptr, nullptr, 0));
  return SWE_OK;
}

int swe_advance(swe_ctx *c, int64_t nsteps, int dt_mode, double dt_value, double t_end, int64_t *taken) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  MIDCHK(c);
  if (nsteps < 0 || dt_mode < SWE_DT_SELF || dt_mode > SWE_DT_SAMPLED) return fail(c, SWE_EINVAL, "bad advance args");
  for (int s = 0; s < 4; ++s)
    if (c->P.bc[s] == BC_STORED) return fail(c, SWE_EINVAL, "swe_advance needs device boundary kinds on every side");
  int rc = pull_ctl(c);
  if (rc) return rc;
  Ctl *h = c->hctl;
  if (dt_mode == SWE_DT_SAMPLED && (c->rec_n == 0 || h->rec_every != 0))
    return fail(c, SWE_EINVAL, "SWE_DT_SAMPLED needs a recorder armed with sample times (every = 0)");
  const long long steps0 = h->steps;
  h->harm_on = c->harm_enabled;
  h->dt_mode = dt_mode == SWE_DT_SELF    ? DT_SELF
               : dt_mode == SWE_DT_FIXED ? DT_FIXED
               : dt_mode == SWE_DT_UNTIL ? DT_UNTIL
                                         : DT_SAMPLED;
  h->dt_fixed = dt_value;
  h->t_end = t_end;
  h->cfl_bits = 0;
  if ((rc = push_ctl(c))) return rc;
  StateP S = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
  HIPCHK(c, launch_maxdt(c->P, S, c->dctl, c->stream));
  HIPCHK(c, launch_finalize(c->dctl, c->mf_part, 0, c->P.cfl, 0, c->stream));
  if (c->use_graph && c->graph_chunk != c->chunk) { drop_graphs(c); c->graph_chunk = c->chunk; }
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  // a sampled or UNTIL run ends at t_end: its length is checked every SAMPLED_BATCH steps
  constexpr int64_t SAMPLED_BATCH = 256;
  for (int64_t n = 0; n < nsteps; ++n) {
    if ((dt_mode == SWE_DT_SAMPLED || dt_mode == SWE_DT_UNTIL) && n > 0 && n % SAMPLED_BATCH == 0) {
      if ((rc = pull_ctl(c))) return rc;
      if (!(h->t < t_end)) break;
    }
    if (c->use_graph) {
      const int which = c->A == 0 ? 0 : 1;
      if (!c->gexec[which]) { if ((rc = build_graph(c, which))) return rc; }
      HIPCHK(c, hipGraphLaunch(c->gexec[which], c->stream));
      std::swap(c->A, c->C);
    } else {
      if ((rc = enqueue_step(c, c->keep_resid, c->keep_resid, true))) return rc;
    }
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  if ((rc = pull_ctl(c))) return rc;
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  if (taken) *taken = h->steps - steps0;
  h->dt_mode = DT_HOST;
  return push_ctl(c);
}

int swe_profile_step(swe_ctx *c, int nsteps, double out_ms[3]) {
  if (!c || !out_ms || nsteps < 1) return fail(c, SWE_EINVAL, "bad profile args");
  MIDCHK(c);
  for (int s = 0; s < 4; ++s)
    if (c->P.bc[s] == BC_STORED) return fail(c, SWE_EINVAL, "profiling needs device boundary kinds");
  int rc = pull_ctl(c);
  if (rc) return rc;
  c->hctl->dt_mode = DT_SELF;
  c->hctl->cfl_bits = 0;
  if ((rc = push_ctl(c))) return rc;
  StateP S = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
  HIPCHK(c, launch_maxdt(c->P, S, c->dctl, c->stream));
  HIPCHK(c, launch_finalize(c->dctl, c->mf_part, 0, c->P.cfl, 0, c->stream));
  const int np = sweep_waves(c->P, c->chunk);
  std::vector<hipEvent_t> ev(4 * nsteps);
  for (auto &e : ev) HIPCHK(c, hipEventCreate(&e));
  for (int n = 0; n < nsteps; ++n) {
    SweepArgs a = base_args(c);
    a.in = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
    a.out = {c->st[c->Bi][0], c->st[c->Bi][1], c->st[c->Bi][2]};
    SweepArgs b2 = base_args(c);
    b2.in = {c->st[c->Bi][0], c->st[c->Bi][1], c->st[c->Bi][2]};
    b2.h0 = {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]};
    b2.out = {c->st[c->C][0], c->st[c->C][1], c->st[c->C][2]};
    b2.mf_part = c->mf_part + np;
    HIPCHK(c, launch_apply_bc(c->P, {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]}, c->b, c->dctl, c->stream));
    HIPCHK(c, hipEventRecord(ev[4 * n], c->stream));
    HIPCHK(c, launch_sweep(MODE_STAGE1, a, c->stream));
    HIPCHK(c, hipEventRecord(ev[4 * n + 1], c->stream));
    HIPCHK(c, launch_apply_bc(c->P, {c->st[c->Bi][0], c->st[c->Bi][1], c->st[c->Bi][2]}, c->b, c->dctl, c->stream));
    HIPCHK(c, launch_sweep(MODE_STAGE2, b2, c->stream));
    HIPCHK(c, hipEventRecord(ev[4 * n + 2], c->stream));
    HIPCHK(c, launch_finalize(c->dctl, c->mf_part, np, c->P.cfl, 1, c->stream));
    HIPCHK(c, hipEventRecord(ev[4 * n + 3], c->stream));
    std::swap(c->A, c->C);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double acc[3] = {0, 0, 0};
  for (int n = 0; n < nsteps; ++n)
    for (int k = 0; k < 3; ++k) {
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, ev[4 * n + k], ev[4 * n + k + 1]));
      acc[k] += ms;
    }
  for (auto &e : ev) (void)hipEventDestroy(e);
  for (int k = 0; k < 3; ++k) out_ms[k] = acc[k] / nsteps;
  if ((rc = pull_ctl(c))) return rc;
  c->hctl->dt_mode = DT_HOST;
  return push_ctl(c);
}

int swe_sync(swe_ctx *c) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_device_ptr(swe_ctx *c, int f, double **p) {
  if (!c || !p) return fail(c, SWE_EINVAL, "null argument");
  *p = field_ptr(c, f);
  return *p ? SWE_OK : fail(c, SWE_ESTATE, "field not allocated");
}

int swe_last_advance_ms(swe_ctx *c, double *ms) {
  if (!c || !ms) return fail(c, SWE_EINVAL, "null argument");
  *ms = c->last_ms;
  return SWE_OK;
}

int swe_set_recorder(swe_ctx *c, int ng, const int32_t *cells, int every, double t0, double dt_rec,
                     int64_t capacity) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  if (ng < 0 || (ng > 0 && (!cells || capacity < 1 || every < 0)))
    return fail(c, SWE_EINVAL, "recorder: bad arguments");
  if (ng > 0 && every == 0 && !(dt_rec > 0 && std::isfinite(dt_rec) && std::isfinite(t0)))
    return fail(c, SWE_EINVAL, "recorder: sample times need a finite t0 and dt_rec > 0");
  for (int g = 0; g < ng; ++g)
    if (cells[g] < 0 || (size_t)cells[g] >= c->N) return fail(c, SWE_EINVAL, "recorder: gauge cell out of range");
  if ((double)ng * (double)capacity * sizeof(double) > 64.0 * (1ull << 30))
    return fail(c, SWE_EINVAL, "recorder: more than 64 GiB of samples");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);     // graphs bake the recorder's buffers in
  (void)hipFree(c->rec_cells); (void)hipFree(c->rec_series); (void)hipFree(c->rec_times);
  c->rec_cells = nullptr; c->rec_series = nullptr; c->rec_times = nullptr;
  c->rec_n = 0; c->rec_cap = 0;
  int rc = pull_ctl(c);
  if (rc) return rc;
  Ctl *h = c->hctl;
  h->rec_every = every; h->rec_flag = 0;
  h->rec_count = 0; h->rec_row = 0; h->rec_since = 0;
  h->rec_t0 = t0; h->rec_next = t0; h->rec_dt = dt_rec;
  if (ng > 0) {
    hipError_t e;
    if ((e = hipMalloc(&c->rec_cells, sizeof(int) * ng)) != hipSuccess ||
        (e = hipMalloc(&c->rec_series, sizeof(double) * (size_t)ng * capacity)) != hipSuccess ||
        (e = hipMalloc(&c->rec_times, sizeof(double) * capacity)) != hipSuccess) {
      (void)hipFree(c->rec_cells); (void)hipFree(c->rec_series); (void)hipFree(c->rec_times);
      c->rec_cells = nullptr; c->rec_series = nullptr; c->rec_times = nullptr;
      return fail(c, e == hipErrorOutOfMemory ? SWE_ENOMEM : SWE_EHIP, std::string("recorder: ") + hipGetErrorString(e));
    }
    HIPCHK(c, hipMemcpyAsync(c->rec_cells, cells, sizeof(int) * ng, hipMemcpyHostToDevice, c->stream));
    c->rec_n = ng;
    c->rec_cap = capacity;
  }
  if ((rc = push_ctl(c))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_set_projection(swe_ctx *c, const double *coef, int64_t capacity) {
  if (!c) return fail(nullptr, SWE_EINVAL, "null context");
  if (coef && capacity < 1) return fail(c, SWE_EINVAL, "projection: capacity must be >= 1");
  if (coef)
    for (int j = 0; j < c->P.ny; ++j)
      if (!std::isfinite(coef[j])) return fail(c, SWE_EINVAL, "projection: coefficients must be finite");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);     // graphs bake the recorder's buffers in
  (void)hipFree(c->proj_coef); (void)hipFree(c->proj_vals); (void)hipFree(c->proj_times);
  c->proj_coef = nullptr; c->proj_vals = nullptr; c->proj_times = nullptr;
  c->proj_cap = 0;
  int rc = pull_ctl(c);
  if (rc) return rc;
  c->hctl->proj_count = 0;
  if (coef) {
    if ((rc = ensure_diag(c))) return rc;
    hipError_t e;
    if ((e = hipMalloc(&c->proj_coef, sizeof(double) * c->P.ny)) != hipSuccess ||
        (e = hipMalloc(&c->proj_vals, sizeof(double) * capacity)) != hipSuccess ||
        (e = hipMalloc(&c->proj_times, sizeof(double) * capacity)) != hipSuccess) {
      (void)hipFree(c->proj_coef); (void)hipFree(c->proj_vals); (void)hipFree(c->proj_times);
      c->proj_coef = nullptr; c->proj_vals = nullptr; c->proj_times = nullptr;
      return fail(c, e == hipErrorOutOfMemory ? SWE_ENOMEM : SWE_EHIP, std::string("projection: ") + hipGetErrorString(e));
    }
    HIPCHK(c, hipMemcpyAsync(c->proj_coef, coef, sizeof(double) * c->P.ny, hipMemcpyHostToDevice, c->stream));
    c->proj_cap = capacity;
  }
  if ((rc = push_ctl(c))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SWE_OK;
}

int swe_get_projection(swe_ctx *c, double *vals, double *times, int64_t *count) {
  if (!c || !count) return fail(c, SWE_EINVAL, "null argument");
  int rc = pull_ctl(c);
  if (rc) return rc;
  const long long n = c->hctl->proj_count;
  *count = n;
  const long long rows = n < c->proj_cap ? n : c->proj_cap;
  if (rows > 0 && vals)
    HIPCHK(c, hipMemcpyAsync(vals, c->proj_vals, sizeof(double) * rows, hipMemcpyDeviceToHost, c->stream));
  if (rows > 0 && times)
    HIPCHK(c, hipMemcpyAsync(times, c->proj_times, sizeof(double) * rows, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return n > c->proj_cap ? fail(c, SWE_ESTATE, "projection: more samples than its capacity (the first ones kept)")
                         : SWE_OK;
}

int swe_get_record(swe_ctx *c, double *eta, double *times, int64_t *count) {
  if (!c || !count) return fail(c, SWE_EINVAL, "null argument");
  int rc = pull_ctl(c);
  if (rc) return rc;
  const long long n = c->hctl->rec_count;
  *count = n;
  const long long rows = n < c->rec_cap ? n : c->rec_cap;
  if (rows > 0 && eta)
    HIPCHK(c, hipMemcpyAsync(eta, c->rec_series, sizeof(double) * (size_t)rows * c->rec_n, hipMemcpyDeviceToHost, c->stream));
  if (rows > 0 && times)
    HIPCHK(c, hipMemcpyAsync(times, c->rec_times, sizeof(double) * (size_t)rows, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return n > c->rec_cap ? fail(c, SWE_ESTATE, "recorder: more samples than its capacity (the first ones kept)") : SWE_OK;
}

}  // extern "C"

// ---- library-internal entry points used by the adjoint ----------------------
namespace swe {
int tide_table(const Params &P, int n, const int32_t *m, const double *amp, const double *lat, const double *lon,
               std::vector<double> &out) {
  if (n < 0 || n > MAX_TIDE || (n > 0 && (!m || !amp || !lat || !lon))) return SWE_EINVAL;
  const int W = P.W, H = P.H;
  out.assign((size_t)n * (2 * H + 2 * W), 0.0);
  for (int c = 0; c < n; ++c) {
    if ((m[c] != 1 && m[c] != 2) || !std::isfinite(amp[c])) return SWE_EINVAL;
    double *X = out.data() + (size_t)c * (2 * H + 2 * W), *Y = X + H, *C = Y + H, *S = C + W;
    const double gA = P.g * amp[c] / P.R;
    for (int j = 0; j < H; ++j) {
      const double phi = lat[j];
      if (!std::isfinite(phi)) return SWE_EINVAL;
      // d eta_eq / (R cos(phi) d lon) and d eta_eq / (R d phi), times g, per unit
      // sin resp. cos of the phase
      if (m[c] == 2) { X[j] = -2.0 * gA * std::cos(phi); Y[j] = -gA * std::sin(2.0 * phi); }
      else { X[j] = -2.0 * gA * std::sin(phi); Y[j] = 2.0 * gA * std::cos(2.0 * phi); }
    }
    for (int i = 0; i < W; ++i) {
      if (!std::isfinite(lon[i])) return SWE_EINVAL;
      C[i] = std::cos(m[c] * lon[i]);
      S[i] = std::sin(m[c] * lon[i]);
    }
  }
  return SWE_OK;
}

int ctx_fail(swe_ctx *c, int code, const std::string &msg) { return fail(c, code, msg); }
int ctx_pull(swe_ctx *c) { return pull_ctl(c); }
int ctx_push(swe_ctx *c) { return push_ctl(c); }
SweepArgs ctx_base_args(swe_ctx *c) { return base_args(c); }
int ctx_enqueue_step(swe_ctx *c, int want_theta, int want_rh) { return enqueue_step(c, want_theta, want_rh); }
}  // namespace swe
