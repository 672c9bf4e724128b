// N-API addon: the thin binding between the Node drop-in (js/swe.mjs) and the
// C ABI of include/swe_amd.h.  Every export maps one C function; typed arrays are
// borrowed for the duration of the call only (napi_get_typedarray_info), never
// retained.  Errors from the library become thrown JS Errors.
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

swe_ctx *get_ctx(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "swe_napi: expected a solver handle");
    return nullptr;
  }
  return static_cast<swe_ctx *>(p);
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

void finalize_ctx(napi_env, void *data, void *) { swe_destroy(static_cast<swe_ctx *>(data)); }

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
    if (!get_f64(env, e, &p, &n) || (p && n != (size_t)(d.ny + 4))) {
      napi_throw_type_error(env, nullptr, "swe_napi.create: rows must be Float64Array(ny+4)");
      return nullptr;
    }
    d.rows[r] = p;
  }
  swe_ctx *ctx = nullptr;
  if (swe_create(&d, &ctx) != SWE_OK) return throw_swe(env, nullptr, "swe_create");
  napi_value ext;
  NAPI_OK(env, napi_create_external(env, ctx, finalize_ctx, nullptr, &ext));
  return ext;
}

// setField(h, id, Float64Array|null) / getField(h, id, Float64Array)
napi_value SetField(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  double *p; size_t n;
  if (!get_f64(env, argv[2], &p, &n)) { napi_throw_type_error(env, nullptr, "setField: Float64Array or null"); return nullptr; }
  if (swe_set_field(c, get_int(env, argv[1]), p) != SWE_OK) return throw_swe(env, c, "swe_set_field");
  return nullptr;
}

napi_value GetField(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  double *p; size_t n;
  if (!get_f64(env, argv[2], &p, &n) || !p) { napi_throw_type_error(env, nullptr, "getField: Float64Array"); return nullptr; }
  if (swe_get_field(c, get_int(env, argv[1]), p) != SWE_OK) return throw_swe(env, c, "swe_get_field");
  return nullptr;
}

napi_value SetRow(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  double *p; size_t n;
  if (!get_f64(env, argv[2], &p, &n) || !p) { napi_throw_type_error(env, nullptr, "setRow: Float64Array"); return nullptr; }
  if (swe_set_row(c, get_int(env, argv[1]), p) != SWE_OK) return throw_swe(env, c, "swe_set_row");
  return nullptr;
}

napi_value SetBc(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  if (swe_set_bc(c, get_int(env, argv[1]), get_int(env, argv[2])) != SWE_OK) return throw_swe(env, c, "swe_set_bc");
  return nullptr;
}

napi_value SetScalar(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  if (swe_set_scalar(c, get_int(env, argv[1]), get_double(env, argv[2])) != SWE_OK)
    return throw_swe(env, c, "swe_set_scalar");
  return nullptr;
}

napi_value GetScalar(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  double v = 0;
  if (swe_get_scalar(c, get_int(env, argv[1]), &v) != SWE_OK) return throw_swe(env, c, "swe_get_scalar");
  return num(env, v);
}

bool get_ext(napi_env env, napi_value v, double ext[8]) {
  double *p; size_t n;
  if (!get_f64(env, v, &p, &n)) return false;
  for (int i = 0; i < 8; ++i) ext[i] = (p && (size_t)i < n) ? p[i] : 0.0;
  return true;
}

napi_value ApplyBc(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  double ext[8];
  if (!get_ext(env, argv[1], ext)) { napi_throw_type_error(env, nullptr, "applyBc: ext Float64Array(8)"); return nullptr; }
  if (swe_apply_bc(c, ext) != SWE_OK) return throw_swe(env, c, "swe_apply_bc");
  return nullptr;
}

napi_value Residual(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  if (swe_residual(c, get_double(env, argv[1])) != SWE_OK) return throw_swe(env, c, "swe_residual");
  return nullptr;
}

napi_value ApplyFriction(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  if (swe_apply_friction(c, get_double(env, argv[1])) != SWE_OK) return throw_swe(env, c, "swe_apply_friction");
  return nullptr;
}

napi_value DryClean(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  double added = 0;
  if (swe_dry_clean(c, &added) != SWE_OK) return throw_swe(env, c, "swe_dry_clean");
  return num(env, added);
}

napi_value MaxDt(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  double v = 0;
  if (swe_max_dt(c, &v) != SWE_OK) return throw_swe(env, c, "swe_max_dt");
  return num(env, v);
}

// step(h, dt (NaN = self-timed), ext, wantTheta, wantRh) -> [dtTaken, massFloored]
napi_value Step(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  double ext[8];
  if (!get_ext(env, argv[2], ext)) { napi_throw_type_error(env, nullptr, "step: ext Float64Array(8)"); return nullptr; }
  bool wt = false, wr = false;
  napi_get_value_bool(env, argv[3], &wt);
  napi_get_value_bool(env, argv[4], &wr);
  double taken = 0, mf = 0;
  if (swe_step(c, get_double(env, argv[1]), ext, wt, wr, &taken, &mf) != SWE_OK) return throw_swe(env, c, "swe_step");
  napi_value arr;
  NAPI_OK(env, napi_create_array_with_length(env, 2, &arr));
  NAPI_OK(env, napi_set_element(env, arr, 0, num(env, taken)));
  NAPI_OK(env, napi_set_element(env, arr, 1, num(env, mf)));
  return arr;
}

// advance(h, nsteps, mode, dtValue, tEnd) -> steps taken
napi_value Advance(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  int64_t taken = 0;
  if (swe_advance(c, (int64_t)get_double(env, argv[1]), get_int(env, argv[2]), get_double(env, argv[3]),
                  get_double(env, argv[4]), &taken) != SWE_OK)
    return throw_swe(env, c, "swe_advance");
  return num(env, (double)taken);
}

// diagnostics(h, ref, Float64Array out[5]): volume, energy, maxSurfaceDeviation, maxSpeed, finite
napi_value Diagnostics(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  double *out = nullptr;
  size_t n = 0;
  if (!get_f64(env, argv[2], &out, &n) || !out || n < SWE_D_COUNT) {
    napi_throw_type_error(env, nullptr, "diagnostics: Float64Array(5)");
    return nullptr;
  }
  if (swe_diagnostics(c, get_double(env, argv[1]), out) != SWE_OK) return throw_swe(env, c, "swe_diagnostics");
  return nullptr;
}

bool get_i32(napi_env env, napi_value v, int32_t **data, size_t *len);

// setRecorder(h, Int32Array cells, every, t0, dtRec, capacity)
napi_value SetRecorder(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  int32_t *cells = nullptr;
  size_t n = 0;
  if (!get_i32(env, argv[1], &cells, &n)) {
    napi_throw_type_error(env, nullptr, "setRecorder: Int32Array cells");
    return nullptr;
  }
  if (swe_set_recorder(c, (int)n, cells, get_int(env, argv[2]), get_double(env, argv[3]), get_double(env, argv[4]),
                       (int64_t)get_double(env, argv[5])) != SWE_OK)
    return throw_swe(env, c, "swe_set_recorder");
  return nullptr;
}

// getRecord(h, Float64Array eta|null, Float64Array times|null) -> samples recorded
napi_value GetRecord(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  double *eta = nullptr, *times = nullptr;
  size_t n1 = 0, n2 = 0;
  if (!get_f64(env, argv[1], &eta, &n1) || !get_f64(env, argv[2], &times, &n2)) {
    napi_throw_type_error(env, nullptr, "getRecord: Float64Array or null");
    return nullptr;
  }
  int64_t count = 0;
  if (swe_get_record(c, eta, times, &count) != SWE_OK) return throw_swe(env, c, "swe_get_record");
  return num(env, (double)count);
}

// setProjection(h, Float64Array coef|null, capacity)
napi_value SetProjection(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  double *coef = nullptr;
  size_t n = 0;
  if (!get_f64(env, argv[1], &coef, &n)) {
    napi_throw_type_error(env, nullptr, "setProjection: Float64Array coef or null");
    return nullptr;
  }
  if (swe_set_projection(c, coef, (int64_t)get_double(env, argv[2])) != SWE_OK)
    return throw_swe(env, c, "swe_set_projection");
  return nullptr;
}

// getProjection(h, Float64Array vals|null, Float64Array times|null) -> samples recorded
napi_value GetProjection(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  double *vals = nullptr, *times = nullptr;
  size_t n1 = 0, n2 = 0;
  if (!get_f64(env, argv[1], &vals, &n1) || !get_f64(env, argv[2], &times, &n2)) {
    napi_throw_type_error(env, nullptr, "getProjection: Float64Array or null");
    return nullptr;
  }
  int64_t count = 0;
  if (swe_get_projection(c, vals, times, &count) != SWE_OK) return throw_swe(env, c, "swe_get_projection");
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

// setTide(h, Int32Array m, Float64Array amp, omega, phase, lat(H), lon(W))
napi_value SetTide(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return nullptr;
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  int32_t *m = nullptr;
  size_t n = 0, k;
  double *a[5];
  if (!get_i32(env, argv[1], &m, &n)) { napi_throw_type_error(env, nullptr, "setTide: m Int32Array"); return nullptr; }
  for (int q = 0; q < 5; ++q)
    if (!get_f64(env, argv[2 + q], &a[q], &k)) { napi_throw_type_error(env, nullptr, "setTide: Float64Array"); return nullptr; }
  if (swe_set_tide(c, (int)n, m, a[0], a[1], a[2], a[3], a[4]) != SWE_OK) return throw_swe(env, c, "swe_set_tide");
  return nullptr;
}

swe_adj *get_adj(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "swe_napi: expected an adjoint handle");
    return nullptr;
  }
  return static_cast<swe_adj *>(p);
}

napi_value throw_adj(napi_env env, swe_adj *a, const char *what) {
  std::string m = std::string(what) + ": " + swe_adj_last_error(a);
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

void finalize_adj(napi_env, void *data, void *) { swe_adj_destroy(static_cast<swe_adj *>(data)); }

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
  swe_ctx *c = get_ctx(env, argv[0]);
  if (!c) return nullptr;
  napi_valuetype t4;
  NAPI_OK(env, napi_typeof(env, argv[3], &t4));
  const int64_t outer = t4 == napi_number ? (int64_t)get_double(env, argv[3]) : 0;
  const int64_t nsteps = (int64_t)get_double(env, argv[1]);
  const int inner = get_int(env, argv[2]);
  swe_adj *a = nullptr;
  const int rc = outer > 0 ? swe_adj_create_levels(c, nsteps, inner, outer, &a) : swe_adj_create(c, nsteps, inner, &a);
  if (rc != SWE_OK) return throw_swe(env, c, "swe_adj_create");
  napi_value ext;
  NAPI_OK(env, napi_create_external(env, a, finalize_adj, nullptr, &ext));
  return ext;
}

// adjSetGauges(a, Int32Array cells, every, Float64Array obs)
napi_value AdjSetGauges(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return nullptr;
  swe_adj *a = get_adj(env, argv[0]);
  if (!a) return nullptr;
  int32_t *cells = nullptr;
  size_t n = 0, no = 0;
  double *obs = nullptr;
  if (!get_i32(env, argv[1], &cells, &n) || !get_f64(env, argv[3], &obs, &no)) {
    napi_throw_type_error(env, nullptr, "adjSetGauges: Int32Array cells, Float64Array obs");
    return nullptr;
  }
  if (swe_adj_set_gauges(a, (int)n, cells, get_int(env, argv[2]), obs) != SWE_OK)
    return throw_adj(env, a, "swe_adj_set_gauges");
  return nullptr;
}

// adjForward(a, h0, nfield|null) -> J ; adjLoss likewise (replays the recorded dt)
napi_value adj_run(napi_env env, napi_callback_info info, bool record) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  swe_adj *a = get_adj(env, argv[0]);
  if (!a) return nullptr;
  double *h0 = nullptr, *nf = nullptr;
  size_t n1 = 0, n2 = 0;
  if (!get_f64(env, argv[1], &h0, &n1) || !h0 || !get_f64(env, argv[2], &nf, &n2)) {
    napi_throw_type_error(env, nullptr, "adjoint: Float64Array h0, Float64Array|null n");
    return nullptr;
  }
  double J = 0;
  const int rc = record ? swe_adj_forward(a, h0, nf, &J) : swe_adj_loss(a, h0, nf, &J);
  if (rc != SWE_OK) return throw_adj(env, a, record ? "swe_adj_forward" : "swe_adj_loss");
  return num(env, J);
}
napi_value AdjForward(napi_env env, napi_callback_info info) { return adj_run(env, info, true); }
napi_value AdjLoss(napi_env env, napi_callback_info info) { return adj_run(env, info, false); }

// adjBackward(a, Float64Array gradH0|null, Float64Array gradN|null)
napi_value AdjBackward(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  swe_adj *a = get_adj(env, argv[0]);
  if (!a) return nullptr;
  double *gh = nullptr, *gn = nullptr;
  size_t n1 = 0, n2 = 0;
  if (!get_f64(env, argv[1], &gh, &n1) || !get_f64(env, argv[2], &gn, &n2)) {
    napi_throw_type_error(env, nullptr, "adjBackward: Float64Array or null");
    return nullptr;
  }
  if (swe_adj_backward(a, gh, gn) != SWE_OK) return throw_adj(env, a, "swe_adj_backward");
  return nullptr;
}

// adjSeries(a, Float64Array out)
napi_value AdjSeries(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  swe_adj *a = get_adj(env, argv[0]);
  if (!a) return nullptr;
  double *p = nullptr;
  size_t n = 0;
  if (!get_f64(env, argv[1], &p, &n) || !p) { napi_throw_type_error(env, nullptr, "adjSeries: Float64Array"); return nullptr; }
  if (swe_adj_get_series(a, p) != SWE_OK) return throw_adj(env, a, "swe_adj_get_series");
  return nullptr;
}

// adjTiming(a) -> [forwardMs, backwardMs]
napi_value AdjTiming(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  swe_adj *a = get_adj(env, argv[0]);
  if (!a) return nullptr;
  double f = 0, b = 0;
  if (swe_adj_timing(a, &f, &b) != SWE_OK) return throw_adj(env, a, "swe_adj_timing");
  napi_value arr;
  NAPI_OK(env, napi_create_array_with_length(env, 2, &arr));
  NAPI_OK(env, napi_set_element(env, arr, 0, num(env, f)));
  NAPI_OK(env, napi_set_element(env, arr, 1, num(env, b)));
  return arr;
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
