// N-API addon: the thin binding between the Node drop-in (js/swe.mjs) and the
// C ABI of include/swe_amd.h.  Every export maps one C function; typed arrays are
// borrowed for the duration of the call only (napi_get_typedarray_info), never
// retained.  Errors from the library become thrown JS Errors.  The C functions
// read or write a fixed number of doubles behind each pointer (W*H per field, H
// per row, samples x gauges per series), so every typed array is checked against
// the sizes the handle records before the call, and a wrong length throws a
// RangeError without touching the library.
#include <node_api.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/swe_amd.h"

namespace {

#define NAPI_OK(env, call)                                                    \
  do {                                                                        \
    if ((call) != napi_ok) {                                                  \
      napi_throw_error((env), nullptr, "swe_napi: N-API call failed: " #call); \
      return nullptr;                                                         \
    }                                                                         \
  } while (0)

// A solver handle: the context and the sizes its arrays must have.
struct Solver {
  swe_ctx *c = nullptr;
  int nx = 0, ny = 0, W = 0, H = 0;
  size_t N = 0;
  int rec_n = 0;              // gauges of the armed recorder
  long long rec_cap = 0;      // its capacity in samples
  long long proj_cap = 0;     // capacity of the projection recorder
};

// An adjoint handle: the adjoint and the sizes of its controls and series.
struct Adj {
  swe_adj *a = nullptr;
  size_t N = 0;
  long long nsteps = 0;
  int ng = 0, every = 0;      // set by adjSetGauges
};

napi_value range_error(napi_env env, const std::string &m) {
  napi_throw_range_error(env, nullptr, m.c_str());
  return nullptr;
}

// "<what>: Float64Array of <want> elements, got <n>"
napi_value bad_length(napi_env env, const char *what, size_t want, size_t n) {
  return range_error(env, std::string(what) + ": expected " + std::to_string(want) + " elements, got " +
                              std::to_string(n));
}

napi_value throw_swe(napi_env env, swe_ctx *ctx, const char *what) {
  std::string m = std::string(what) + ": " + swe_last_error(ctx);
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
  for (size_t i = argc; i < want; ++i) napi_get_undefined(env, &argv[i]);
  return true;
}

Solver *get_solver(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !static_cast<Solver *>(p)->c) {
    napi_throw_type_error(env, nullptr, "swe_napi: expected a solver handle");
    return nullptr;
  }
  return static_cast<Solver *>(p);
}

double get_double(napi_env env, napi_value v) {
  double d = NAN;
  napi_get_value_double(env, v, &d);
  return d;
}

int get_int(napi_env env, napi_value v) {
  int32_t i = 0;
  napi_get_value_int32(env, v, &i);
  return i;
}

// Float64Array -> data pointer and length; null/undefined -> nullptr
bool get_f64(napi_env env, napi_value v, double **data, size_t *len) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_null || t == napi_undefined) { *data = nullptr; *len = 0; return true; }
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_ta) return false;
  napi_typedarray_type tt;
  void *p = nullptr;
  napi_value ab;
  size_t off = 0;
  if (napi_get_typedarray_info(env, v, &tt, len, &p, &ab, &off) != napi_ok || tt != napi_float64_array) return false;
  *data = static_cast<double *>(p);
  return true;
}

napi_value num(napi_env env, double d) {
  napi_value r;
  napi_create_double(env, d, &r);
  return r;
}

void finalize_solver(napi_env, void *data, void *) {
  Solver *h = static_cast<Solver *>(data);
  swe_destroy(h->c);
  delete h;
}

// create({nx, ny, kind, order, limiter, g, minDepth, eps4, cfl, coriolis, manning, R, rows:[9 Float64Array]})
napi_value Create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  napi_value o = argv[0];
  auto prop = [&](const char *name) {
    napi_value v;
    napi_get_named_property(env, o, name, &v);
    return v;
  };
  swe_desc d;
  std::memset(&d, 0, sizeof(d));
  d.nx = get_int(env, prop("nx"));
  d.ny = get_int(env, prop("ny"));
  if (d.nx < 2 || d.ny < 1 || d.nx > (1 << 20) || d.ny > (1 << 20))
    return range_error(env, "swe_napi.create: nx must be in [2, 2^20] and ny in [1, 2^20]");
  d.kind = get_int(env, prop("kind"));
  d.order = get_int(env, prop("order"));
  d.limiter = get_int(env, prop("limiter"));
  d.g = get_double(env, prop("g"));
  d.min_depth = get_double(env, prop("minDepth"));
  d.eps4 = get_double(env, prop("eps4"));
  d.cfl = get_double(env, prop("cfl"));
  d.coriolis = get_double(env, prop("coriolis"));
  d.manning = get_double(env, prop("manning"));
  d.R = get_double(env, prop("R"));
  napi_value rows = prop("rows");
  for (uint32_t r = 0; r < SWE_NROWS; ++r) {
    napi_value e;
    NAPI_OK(env, napi_get_element(env, rows, r, &e));
    double *p = nullptr;
    size_t n = 0;
    if (!get_f64(env, e, &p, &n) || (!p && r != SWE_ROW_AREA)) {
      napi_throw_type_error(env, nullptr, "swe_napi.create: rows must be Float64Array(ny+4)");
      return nullptr;
    }
    if (p && n != (size_t)(d.ny + 4)) return bad_length(env, "swe_napi.create: geometry row", d.ny + 4, n);
    d.rows[r] = p;
  }
  swe_ctx *ctx = nullptr;
  if (swe_create(&d, &ctx) != SWE_OK) return throw_swe(env, nullptr, "swe_create");
  Solver *h = new Solver();
  h->c = ctx;
  h->nx = d.nx; h->ny = d.ny; h->W = d.nx + 4; h->H = d.ny + 4;
  h->N = (size_t)h->W * h->H;
  napi_value ext;
  if (napi_create_external(env, h, finalize_solver, nullptr, &ext) != napi_ok) {
    swe_destroy(ctx);
    delete h;
    napi_throw_error(env, nullptr, "swe_napi: napi_create_external failed");
    return nullptr;
  }
  return ext;
}

// setField(h, id, Float64Array(W*H)|null) / getField(h, id, Float64Array(W*H))
napi_value SetField(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double *p; size_t n;
  if (!get_f64(env, argv[2], &p, &n)) { napi_throw_type_error(env, nullptr, "setField: Float64Array or null"); return nullptr; }
  if (p && n != h->N) return bad_length(env, "setField", h->N, n);
  if (swe_set_field(h->c, get_int(env, argv[1]), p) != SWE_OK) return throw_swe(env, h->c, "swe_set_field");
  return nullptr;
}

napi_value GetField(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double *p; size_t n;
  if (!get_f64(env, argv[2], &p, &n) || !p) { napi_throw_type_error(env, nullptr, "getField: Float64Array"); return nullptr; }
  if (n != h->N) return bad_length(env, "getField", h->N, n);
  if (swe_get_field(h->c, get_int(env, argv[1]), p) != SWE_OK) return throw_swe(env, h->c, "swe_get_field");
  return nullptr;
}

// getGhosts(h, id, Float64Array(W*H)): the ghost ring only
napi_value GetGhosts(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double *p; size_t n;
  if (!get_f64(env, argv[2], &p, &n) || !p) { napi_throw_type_error(env, nullptr, "getGhosts: Float64Array"); return nullptr; }
  if (n != h->N) return bad_length(env, "getGhosts", h->N, n);
  if (swe_get_ghosts(h->c, get_int(env, argv[1]), p) != SWE_OK) return throw_swe(env, h->c, "swe_get_ghosts");
  return nullptr;
}

// setRow(h, id, Float64Array(H))
napi_value SetRow(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double *p; size_t n;
  if (!get_f64(env, argv[2], &p, &n) || !p) { napi_throw_type_error(env, nullptr, "setRow: Float64Array"); return nullptr; }
  if (n != (size_t)h->H) return bad_length(env, "setRow", h->H, n);
  if (swe_set_row(h->c, get_int(env, argv[1]), p) != SWE_OK) return throw_swe(env, h->c, "swe_set_row");
  return nullptr;
}

napi_value SetBc(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  if (swe_set_bc(h->c, get_int(env, argv[1]), get_int(env, argv[2])) != SWE_OK) return throw_swe(env, h->c, "swe_set_bc");
  return nullptr;
}

napi_value SetScalar(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  if (swe_set_scalar(h->c, get_int(env, argv[1]), get_double(env, argv[2])) != SWE_OK)
    return throw_swe(env, h->c, "swe_set_scalar");
  return nullptr;
}

napi_value GetScalar(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double v = 0;
  if (swe_get_scalar(h->c, get_int(env, argv[1]), &v) != SWE_OK) return throw_swe(env, h->c, "swe_get_scalar");
  return num(env, v);
}

// the eight Flather values (eta, u per side); a shorter array is zero-filled, a
// longer one read only to its eighth element
bool get_ext(napi_env env, napi_value v, double ext[8]) {
  double *p; size_t n;
  if (!get_f64(env, v, &p, &n)) return false;
  for (int i = 0; i < 8; ++i) ext[i] = (p && (size_t)i < n) ? p[i] : 0.0;
  return true;
}

napi_value ApplyBc(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double ext[8];
  if (!get_ext(env, argv[1], ext)) { napi_throw_type_error(env, nullptr, "applyBc: ext Float64Array(8)"); return nullptr; }
  if (swe_apply_bc(h->c, ext) != SWE_OK) return throw_swe(env, h->c, "swe_apply_bc");
  return nullptr;
}

napi_value Residual(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  if (swe_residual(h->c, get_double(env, argv[1])) != SWE_OK) return throw_swe(env, h->c, "swe_residual");
  return nullptr;
}

napi_value ApplyFriction(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  if (swe_apply_friction(h->c, get_double(env, argv[1])) != SWE_OK) return throw_swe(env, h->c, "swe_apply_friction");
  return nullptr;
}

napi_value DryClean(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double added = 0;
  if (swe_dry_clean(h->c, &added) != SWE_OK) return throw_swe(env, h->c, "swe_dry_clean");
  return num(env, added);
}

napi_value MaxDt(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double v = 0;
  if (swe_max_dt(h->c, &v) != SWE_OK) return throw_swe(env, h->c, "swe_max_dt");
  return num(env, v);
}

napi_value pair(napi_env env, double a, double b) {
  napi_value arr;
  NAPI_OK(env, napi_create_array_with_length(env, 2, &arr));
  NAPI_OK(env, napi_set_element(env, arr, 0, num(env, a)));
  NAPI_OK(env, napi_set_element(env, arr, 1, num(env, b)));
  return arr;
}

// step(h, dt (NaN = self-timed), ext, wantTheta, wantRh) -> [dtTaken, massFloored]
napi_value Step(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double ext[8];
  if (!get_ext(env, argv[2], ext)) { napi_throw_type_error(env, nullptr, "step: ext Float64Array(8)"); return nullptr; }
  bool wt = false, wr = false;
  napi_get_value_bool(env, argv[3], &wt);
  napi_get_value_bool(env, argv[4], &wr);
  double taken = 0, mf = 0;
  if (swe_step(h->c, get_double(env, argv[1]), ext, wt, wr, &taken, &mf) != SWE_OK) return throw_swe(env, h->c, "swe_step");
  return pair(env, taken, mf);
}

// stepBegin(h, dt (NaN = self-timed), ext) -> dtTaken (0: refused)
napi_value StepBegin(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double ext[8];
  if (!get_ext(env, argv[2], ext)) { napi_throw_type_error(env, nullptr, "stepBegin: ext Float64Array(8)"); return nullptr; }
  double taken = 0;
  if (swe_step_begin(h->c, get_double(env, argv[1]), ext, &taken) != SWE_OK) return throw_swe(env, h->c, "swe_step_begin");
  return num(env, taken);
}

// stepEnd(h, wantTheta, wantRh) -> massFloored
napi_value StepEnd(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  bool wt = false, wr = false;
  napi_get_value_bool(env, argv[1], &wt);
  napi_get_value_bool(env, argv[2], &wr);
  double mf = 0;
  if (swe_step_end(h->c, nullptr, wt, wr, &mf) != SWE_OK) return throw_swe(env, h->c, "swe_step_end");
  return num(env, mf);
}

// keepResidual(h, on)
napi_value KeepResidual(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  bool on = false;
  napi_get_value_bool(env, argv[1], &on);
  if (swe_keep_residual(h->c, on ? 1 : 0) != SWE_OK) return throw_swe(env, h->c, "swe_keep_residual");
  return nullptr;
}

// getExt(h, Float64Array(8))
napi_value GetExt(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double *p; size_t n;
  if (!get_f64(env, argv[1], &p, &n) || !p) { napi_throw_type_error(env, nullptr, "getExt: Float64Array"); return nullptr; }
  if (n != 8) return bad_length(env, "getExt", 8, n);
  if (swe_get_ext(h->c, p) != SWE_OK) return throw_swe(env, h->c, "swe_get_ext");
  return nullptr;
}

// read up to 4 harmonic terms from three Float64Arrays of equal length (or nulls: none)
bool get_terms(napi_env env, napi_value a, napi_value o, napi_value p, double *amp, double *om, double *ph, int *n,
               const char *what) {
  double *x[3]; size_t len[3];
  napi_value v[3] = {a, o, p};
  for (int q = 0; q < 3; ++q)
    if (!get_f64(env, v[q], &x[q], &len[q])) {
      napi_throw_type_error(env, nullptr, (std::string(what) + ": Float64Array or null").c_str());
      return false;
    }
  if (len[1] != len[0] || len[2] != len[0] || len[0] > 4) {
    range_error(env, std::string(what) + ": amplitude, frequency and phase arrays of one length, at most 4");
    return false;
  }
  *n = (int)len[0];
  for (int k = 0; k < *n; ++k) { amp[k] = x[0][k]; om[k] = x[1][k]; ph[k] = x[2][k]; }
  return true;
}

// setHarmonic(h, side, msl, amp, omega, phase, u0, uamp, uomega, uphase): the Flather
// forcing device-resident runs evaluate, eta = msl + sum amp cos(omega t - phase)
napi_value SetHarmonic(napi_env env, napi_callback_info info) {
  napi_value argv[10];
  if (!get_args(env, info, 10, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double a[4], o[4], p[4], ua[4], uo[4], up[4];
  int ne = 0, nu = 0;
  if (!get_terms(env, argv[3], argv[4], argv[5], a, o, p, &ne, "setHarmonic eta")) return nullptr;
  if (!get_terms(env, argv[7], argv[8], argv[9], ua, uo, up, &nu, "setHarmonic u")) return nullptr;
  if (swe_set_harmonic(h->c, get_int(env, argv[1]), ne, get_double(env, argv[2]), a, o, p, nu, get_double(env, argv[6]),
                       ua, uo, up) != SWE_OK)
    return throw_swe(env, h->c, "swe_set_harmonic");
  return nullptr;
}

// setHarmonicRamp(h, side, ramp, uFromEta, uScale)
napi_value SetHarmonicRamp(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  bool fe = false;
  napi_get_value_bool(env, argv[3], &fe);
  if (swe_set_harmonic_ramp(h->c, get_int(env, argv[1]), get_double(env, argv[2]), fe ? 1 : 0,
                            get_double(env, argv[4])) != SWE_OK)
    return throw_swe(env, h->c, "swe_set_harmonic_ramp");
  return nullptr;
}

// advance(h, nsteps, mode, dtValue, tEnd) -> steps taken
napi_value Advance(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  int64_t taken = 0;
  if (swe_advance(h->c, (int64_t)get_double(env, argv[1]), get_int(env, argv[2]), get_double(env, argv[3]),
                  get_double(env, argv[4]), &taken) != SWE_OK)
    return throw_swe(env, h->c, "swe_advance");
  return num(env, (double)taken);
}

// diagnostics(h, ref, Float64Array out[5]): volume, energy, maxSurfaceDeviation, maxSpeed, finite
napi_value Diagnostics(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double *out = nullptr;
  size_t n = 0;
  if (!get_f64(env, argv[2], &out, &n) || !out) {
    napi_throw_type_error(env, nullptr, "diagnostics: Float64Array(5)");
    return nullptr;
  }
  if (n != SWE_D_COUNT) return bad_length(env, "diagnostics", SWE_D_COUNT, n);
  if (swe_diagnostics(h->c, get_double(env, argv[1]), out) != SWE_OK) return throw_swe(env, h->c, "swe_diagnostics");
  return nullptr;
}

bool get_i32(napi_env env, napi_value v, int32_t **data, size_t *len);

// setRecorder(h, Int32Array cells, every, t0, dtRec, capacity)
napi_value SetRecorder(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  int32_t *cells = nullptr;
  size_t n = 0;
  if (!get_i32(env, argv[1], &cells, &n)) {
    napi_throw_type_error(env, nullptr, "setRecorder: Int32Array cells");
    return nullptr;
  }
  if (n > (size_t)INT32_MAX) return range_error(env, "setRecorder: too many cells");
  const int64_t cap = (int64_t)get_double(env, argv[5]);
  if (swe_set_recorder(h->c, (int)n, cells, get_int(env, argv[2]), get_double(env, argv[3]), get_double(env, argv[4]),
                       cap) != SWE_OK)
    return throw_swe(env, h->c, "swe_set_recorder");
  h->rec_n = (int)n;
  h->rec_cap = n > 0 ? cap : 0;
  return nullptr;
}

// rows a series read may copy: min(count, capacity)
long long held(long long count, long long cap) { return count < cap ? count : cap; }

// getRecord(h, Float64Array eta|null, Float64Array times|null) -> samples recorded
napi_value GetRecord(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double *eta = nullptr, *times = nullptr;
  size_t n1 = 0, n2 = 0;
  if (!get_f64(env, argv[1], &eta, &n1) || !get_f64(env, argv[2], &times, &n2)) {
    napi_throw_type_error(env, nullptr, "getRecord: Float64Array or null");
    return nullptr;
  }
  int64_t count = 0;
  const int rc0 = swe_get_record(h->c, nullptr, nullptr, &count);
  if (rc0 != SWE_OK && rc0 != SWE_ESTATE) return throw_swe(env, h->c, "swe_get_record");
  const long long rows = held(count, h->rec_cap);
  if (eta && n1 < (size_t)(rows * h->rec_n)) return bad_length(env, "getRecord eta", (size_t)(rows * h->rec_n), n1);
  if (times && n2 < (size_t)rows) return bad_length(env, "getRecord times", (size_t)rows, n2);
  if (swe_get_record(h->c, eta, times, &count) != SWE_OK) return throw_swe(env, h->c, "swe_get_record");
  return num(env, (double)count);
}

// setProjection(h, Float64Array(ny) coef|null, capacity)
napi_value SetProjection(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double *coef = nullptr;
  size_t n = 0;
  if (!get_f64(env, argv[1], &coef, &n)) {
    napi_throw_type_error(env, nullptr, "setProjection: Float64Array coef or null");
    return nullptr;
  }
  if (coef && n != (size_t)h->ny) return bad_length(env, "setProjection coef", h->ny, n);
  const int64_t cap = (int64_t)get_double(env, argv[2]);
  if (swe_set_projection(h->c, coef, cap) != SWE_OK) return throw_swe(env, h->c, "swe_set_projection");
  h->proj_cap = coef ? cap : 0;
  return nullptr;
}

// getProjection(h, Float64Array vals|null, Float64Array times|null) -> samples recorded
napi_value GetProjection(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  double *vals = nullptr, *times = nullptr;
  size_t n1 = 0, n2 = 0;
  if (!get_f64(env, argv[1], &vals, &n1) || !get_f64(env, argv[2], &times, &n2)) {
    napi_throw_type_error(env, nullptr, "getProjection: Float64Array or null");
    return nullptr;
  }
  int64_t count = 0;
  const int rc0 = swe_get_projection(h->c, nullptr, nullptr, &count);
  if (rc0 != SWE_OK && rc0 != SWE_ESTATE) return throw_swe(env, h->c, "swe_get_projection");
  const long long rows = held(count, h->proj_cap);
  if (vals && n1 < (size_t)rows) return bad_length(env, "getProjection vals", (size_t)rows, n1);
  if (times && n2 < (size_t)rows) return bad_length(env, "getProjection times", (size_t)rows, n2);
  if (swe_get_projection(h->c, vals, times, &count) != SWE_OK) return throw_swe(env, h->c, "swe_get_projection");
  return num(env, (double)count);
}

// hllc(hL, huL, hvL, hR, huR, hvR, dry, g) -> [m, n, t]
napi_value Hllc(napi_env env, napi_callback_info info) {
  napi_value argv[8];
  if (!get_args(env, info, 8, argv)) return nullptr;
  double a[8];
  for (int i = 0; i < 8; ++i) a[i] = get_double(env, argv[i]);
  double out[3];
  swe_hllc(out, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]);
  napi_value arr;
  NAPI_OK(env, napi_create_array_with_length(env, 3, &arr));
  for (uint32_t i = 0; i < 3; ++i) NAPI_OK(env, napi_set_element(env, arr, i, num(env, out[i])));
  return arr;
}

napi_value JsPow(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  return num(env, swe_js_pow(get_double(env, argv[0]), get_double(env, argv[1])));
}

// ---- equilibrium tide and the discrete adjoint ----------------------------------
bool get_i32(napi_env env, napi_value v, int32_t **data, size_t *len) {
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_ta) return false;
  napi_typedarray_type tt;
  void *p = nullptr;
  napi_value ab;
  size_t off = 0;
  if (napi_get_typedarray_info(env, v, &tt, len, &p, &ab, &off) != napi_ok || tt != napi_int32_array) return false;
  *data = static_cast<int32_t *>(p);
  return true;
}

// setTide(h, Int32Array m, Float64Array amp, omega, phase (one per constituent), lat(H), lon(W))
napi_value SetTide(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  int32_t *m = nullptr;
  size_t n = 0, k[5];
  double *a[5];
  if (!get_i32(env, argv[1], &m, &n)) { napi_throw_type_error(env, nullptr, "setTide: m Int32Array"); return nullptr; }
  for (int q = 0; q < 5; ++q)
    if (!get_f64(env, argv[2 + q], &a[q], &k[q])) { napi_throw_type_error(env, nullptr, "setTide: Float64Array"); return nullptr; }
  if (n > 4) return range_error(env, "setTide: at most 4 constituents");
  for (int q = 0; q < 3; ++q)
    if (k[q] != n) return bad_length(env, "setTide amp/omega/phase", n, k[q]);
  if (n > 0 && k[3] != (size_t)h->H) return bad_length(env, "setTide lat", h->H, k[3]);
  if (n > 0 && k[4] != (size_t)h->W) return bad_length(env, "setTide lon", h->W, k[4]);
  if (swe_set_tide(h->c, (int)n, m, a[0], a[1], a[2], a[3], a[4]) != SWE_OK) return throw_swe(env, h->c, "swe_set_tide");
  return nullptr;
}

Adj *get_adj(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !static_cast<Adj *>(p)->a) {
    napi_throw_type_error(env, nullptr, "swe_napi: expected an adjoint handle");
    return nullptr;
  }
  return static_cast<Adj *>(p);
}

napi_value throw_adj(napi_env env, swe_adj *a, const char *what) {
  std::string m = std::string(what) + ": " + swe_adj_last_error(a);
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

void finalize_adj(napi_env, void *data, void *) {
  Adj *h = static_cast<Adj *>(data);
  swe_adj_destroy(h->a);
  delete h;
}

void finalize_pinned(napi_env env, void *data, void *hint) {
  int64_t change = 0;
  napi_adjust_external_memory(env, -(int64_t)(reinterpret_cast<uintptr_t>(hint)), &change);
  swe_host_free(data);
}

// allocPinned(n) -> Float64Array of n zeros in page-locked host memory (freed when
// the array is collected; V8 is told of the external bytes)
napi_value AllocPinned(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  const double nd = get_double(env, argv[0]);
  if (!(nd >= 0) || nd > 9.0e15) {
    napi_throw_range_error(env, nullptr, "swe_napi: bad array length");
    return nullptr;
  }
  const size_t n = (size_t)nd, bytes = n * sizeof(double);
  void *p = nullptr;
  if (swe_host_alloc(bytes, &p) != SWE_OK) return throw_swe(env, nullptr, "swe_host_alloc");
  std::memset(p, 0, bytes);
  napi_value ab, ta;
  if (napi_create_external_arraybuffer(env, p, bytes, finalize_pinned, reinterpret_cast<void *>((uintptr_t)bytes),
                                       &ab) != napi_ok) {
    swe_host_free(p);
    napi_throw_error(env, nullptr, "swe_napi: napi_create_external_arraybuffer failed");
    return nullptr;
  }
  int64_t change = 0;
  napi_adjust_external_memory(env, (int64_t)bytes, &change);
  NAPI_OK(env, napi_create_typedarray(env, napi_float64_array, n, ab, 0, &ta));
  return ta;
}

// adjCreate(h, nsteps, checkpointEvery, outerEvery = 0) -> adjoint handle; a
// nonzero outerEvery (a multiple of checkpointEvery) selects two-level checkpointing
napi_value AdjCreate(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return nullptr;
  Solver *h = get_solver(env, argv[0]);
  if (!h) return nullptr;
  napi_valuetype t4;
  NAPI_OK(env, napi_typeof(env, argv[3], &t4));
  const int64_t outer = t4 == napi_number ? (int64_t)get_double(env, argv[3]) : 0;
  const int64_t nsteps = (int64_t)get_double(env, argv[1]);
  const int inner = get_int(env, argv[2]);
  swe_adj *a = nullptr;
  const int rc = outer > 0 ? swe_adj_create_levels(h->c, nsteps, inner, outer, &a) : swe_adj_create(h->c, nsteps, inner, &a);
  if (rc != SWE_OK) return throw_swe(env, h->c, "swe_adj_create");
  Adj *ah = new Adj();
  ah->a = a;
  ah->N = h->N;
  ah->nsteps = nsteps;
  napi_value ext;
  if (napi_create_external(env, ah, finalize_adj, nullptr, &ext) != napi_ok) {
    swe_adj_destroy(a);
    delete ah;
    napi_throw_error(env, nullptr, "swe_napi: napi_create_external failed");
    return nullptr;
  }
  return ext;
}

// gauge samples of a run: nsteps / every rows of ng values
size_t series_len(const Adj *a) { return a->every > 0 ? (size_t)(a->nsteps / a->every) * (size_t)a->ng : 0; }

// adjSetGauges(a, Int32Array cells, every, Float64Array obs((nsteps / every) * cells))
napi_value AdjSetGauges(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return nullptr;
  Adj *a = get_adj(env, argv[0]);
  if (!a) return nullptr;
  int32_t *cells = nullptr;
  size_t n = 0, no = 0;
  double *obs = nullptr;
  if (!get_i32(env, argv[1], &cells, &n) || !get_f64(env, argv[3], &obs, &no) || !obs) {
    napi_throw_type_error(env, nullptr, "adjSetGauges: Int32Array cells, Float64Array obs");
    return nullptr;
  }
  const int every = get_int(env, argv[2]);
  if (every < 1) return range_error(env, "adjSetGauges: every must be >= 1");
  if (n > (size_t)INT32_MAX) return range_error(env, "adjSetGauges: too many gauges");
  const size_t want = (size_t)(a->nsteps / every) * n;
  if (no != want) return bad_length(env, "adjSetGauges obs", want, no);
  if (swe_adj_set_gauges(a->a, (int)n, cells, every, obs) != SWE_OK) return throw_adj(env, a->a, "swe_adj_set_gauges");
  a->ng = (int)n;
  a->every = every;
  return nullptr;
}

// adjForward(a, h0(W*H), nfield(W*H)|null) -> J ; adjLoss likewise (replays the recorded dt)
napi_value adj_run(napi_env env, napi_callback_info info, bool record) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Adj *a = get_adj(env, argv[0]);
  if (!a) return nullptr;
  double *h0 = nullptr, *nf = nullptr;
  size_t n1 = 0, n2 = 0;
  if (!get_f64(env, argv[1], &h0, &n1) || !h0 || !get_f64(env, argv[2], &nf, &n2)) {
    napi_throw_type_error(env, nullptr, "adjoint: Float64Array h0, Float64Array|null n");
    return nullptr;
  }
  if (n1 != a->N) return bad_length(env, "adjoint h0", a->N, n1);
  if (nf && n2 != a->N) return bad_length(env, "adjoint Manning field", a->N, n2);
  double J = 0;
  const int rc = record ? swe_adj_forward(a->a, h0, nf, &J) : swe_adj_loss(a->a, h0, nf, &J);
  if (rc != SWE_OK) return throw_adj(env, a->a, record ? "swe_adj_forward" : "swe_adj_loss");
  return num(env, J);
}
napi_value AdjForward(napi_env env, napi_callback_info info) { return adj_run(env, info, true); }
napi_value AdjLoss(napi_env env, napi_callback_info info) { return adj_run(env, info, false); }

// adjBackward(a, Float64Array(W*H) gradH0|null, Float64Array(W*H) gradN|null)
napi_value AdjBackward(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  Adj *a = get_adj(env, argv[0]);
  if (!a) return nullptr;
  double *gh = nullptr, *gn = nullptr;
  size_t n1 = 0, n2 = 0;
  if (!get_f64(env, argv[1], &gh, &n1) || !get_f64(env, argv[2], &gn, &n2)) {
    napi_throw_type_error(env, nullptr, "adjBackward: Float64Array or null");
    return nullptr;
  }
  if (gh && n1 != a->N) return bad_length(env, "adjBackward gradH0", a->N, n1);
  if (gn && n2 != a->N) return bad_length(env, "adjBackward gradN", a->N, n2);
  if (swe_adj_backward(a->a, gh, gn) != SWE_OK) return throw_adj(env, a->a, "swe_adj_backward");
  return nullptr;
}

// adjSeries(a, Float64Array((nsteps / every) * gauges))
napi_value AdjSeries(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  Adj *a = get_adj(env, argv[0]);
  if (!a) return nullptr;
  double *p = nullptr;
  size_t n = 0;
  if (!get_f64(env, argv[1], &p, &n) || !p) { napi_throw_type_error(env, nullptr, "adjSeries: Float64Array"); return nullptr; }
  if (a->every < 1) return range_error(env, "adjSeries: no gauges set (adjSetGauges)");
  if (n != series_len(a)) return bad_length(env, "adjSeries", series_len(a), n);
  if (swe_adj_get_series(a->a, p) != SWE_OK) return throw_adj(env, a->a, "swe_adj_get_series");
  return nullptr;
}

// adjTiming(a) -> [forwardMs, backwardMs]
napi_value AdjTiming(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  Adj *a = get_adj(env, argv[0]);
  if (!a) return nullptr;
  double f = 0, b = 0;
  if (swe_adj_timing(a->a, &f, &b) != SWE_OK) return throw_adj(env, a->a, "swe_adj_timing");
  return pair(env, f, b);
}

napi_value BuildInfo(napi_env env, napi_callback_info) {
  napi_value s;
  napi_create_string_utf8(env, swe_build_info(), NAPI_AUTO_LENGTH, &s);
  return s;
}

napi_value Init(napi_env env, napi_value exports) {
  struct { const char *name; napi_callback cb; } fns[] = {
      {"create", Create}, {"setField", SetField}, {"getField", GetField}, {"setRow", SetRow},
      {"setBc", SetBc}, {"setScalar", SetScalar}, {"getScalar", GetScalar}, {"applyBc", ApplyBc},
      {"residual", Residual}, {"maxDt", MaxDt}, {"applyFriction", ApplyFriction}, {"dryClean", DryClean}, {"step", Step}, {"advance", Advance}, {"hllc", Hllc},
      {"jsPow", JsPow}, {"buildInfo", BuildInfo}, {"setTide", SetTide},
      {"adjCreate", AdjCreate}, {"adjSetGauges", AdjSetGauges}, {"adjForward", AdjForward}, {"adjLoss", AdjLoss},
      {"adjBackward", AdjBackward}, {"adjSeries", AdjSeries}, {"adjTiming", AdjTiming},
      {"allocPinned", AllocPinned}, {"diagnostics", Diagnostics}, {"setRecorder", SetRecorder},
      {"getRecord", GetRecord}, {"setProjection", SetProjection}, {"getProjection", GetProjection},
      {"getGhosts", GetGhosts}, {"stepBegin", StepBegin}, {"stepEnd", StepEnd}, {"keepResidual", KeepResidual},
      {"getExt", GetExt}, {"setHarmonic", SetHarmonic}, {"setHarmonicRamp", SetHarmonicRamp},
  };
  for (auto &f : fns) {
    napi_value fn;
    NAPI_OK(env, napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn));
    NAPI_OK(env, napi_set_named_property(env, exports, f.name, fn));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
