/*
 * swe_amd.h -- C ABI of the MI355X (gfx950) shallow-water time step.
 *
 * This is the drop-in boundary for ShallowWater of src/swe.mjs in
 * Syntaxswine/water-simulator.  The reference has no FFI: its "operator API" is
 * the JS class.  Each entry point below replaces one method or one hot loop of
 * that class (file:line of the reference in the comment); the N-API addon and
 * the JS module in water-simulator_amd/js bind them one to one.
 *
 * Conventions: every function returns 0 (SWE_OK) or a negative SWE_E* code and
 * never throws; swe_last_error() gives the text.  Arrays are host pointers to
 * row-major doubles of W*H = (nx+4)*(ny+4) cells (the reference's layout,
 * swe.mjs:196-199) or H per-row values.  Calls are synchronous: when one
 * returns, its results are in the host buffers it was given.  A context is not
 * thread-safe; one stream per context.
 */
#ifndef SWE_AMD_H
#define SWE_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWE_OK 0
#define SWE_EINVAL (-1)
#define SWE_EHIP (-2)
#define SWE_ENOMEM (-3)
#define SWE_ESTATE (-4)

enum swe_kind { SWE_CARTESIAN = 0, SWE_SPHERE = 1 };
enum swe_limiter { SWE_LIM_MINMOD = 0, SWE_LIM_MC = 1, SWE_LIM_CENTRAL = 2 };
/* boundary kinds: reflect swe.mjs:1026, outflow :1042, periodic :1119, flather :1084;
 * STORED = ghost cells supplied by the caller (an unrecognised JS boundary). */
enum swe_bc { SWE_BC_REFLECT = 0, SWE_BC_OUTFLOW = 1, SWE_BC_PERIODIC = 2, SWE_BC_FLATHER = 3,
              SWE_BC_STORED = 4 };
enum swe_side { SWE_WEST = 0, SWE_EAST = 1, SWE_SOUTH = 2, SWE_NORTH = 3 };
/* fields (W*H doubles each) */
enum swe_field { SWE_F_B = 0, SWE_F_H = 1, SWE_F_HU = 2, SWE_F_HV = 3, SWE_F_MANNING = 4,
                 SWE_F_RH = 5, SWE_F_RHU = 6, SWE_F_RHV = 7, SWE_F_THETA = 8, SWE_F_SPONGE = 9 };
/* geometry rows (H doubles each), the arrays of geometry.mjs:301-309 the step reads */
enum swe_row { SWE_ROW_DX = 0, SWE_ROW_DYN = 1, SWE_ROW_DYS = 2, SWE_ROW_DYCFL = 3,
               SWE_ROW_BEDPHI = 4, SWE_ROW_TANPHI = 5, SWE_ROW_F = 6, SWE_ROW_GEO = 7,
               SWE_ROW_AREA = 8, SWE_NROWS = 9 };
/* scalars settable after creation */
enum swe_scalar { SWE_S_MANNING = 0, SWE_S_CORIOLIS = 1, SWE_S_CFL = 2, SWE_S_T = 3,
                  SWE_S_SPONGE_REF = 4, SWE_S_STEPS = 5, SWE_S_MASS_FLOORED = 6,
                  SWE_S_DT_NEXT = 7, SWE_S_NONFINITE = 8, SWE_S_CHUNK_ROWS = 9,
                  SWE_S_DIAG_MS = 10 };
/* dt policy of swe_advance; SAMPLED: the sample-time schedule of the recorder
 * (swe_set_recorder with every = 0) up to t_end */
enum swe_dt_mode { SWE_DT_SELF = 0, SWE_DT_FIXED = 1, SWE_DT_UNTIL = 2, SWE_DT_SAMPLED = 3 };
/* results of swe_diagnostics */
enum swe_diag { SWE_D_VOLUME = 0, SWE_D_ENERGY = 1, SWE_D_MAX_SURFACE_DEVIATION = 2, SWE_D_MAX_SPEED = 3,
                SWE_D_FINITE = 4, SWE_D_COUNT = 5 };

typedef struct swe_desc {
  int32_t nx, ny;           /* interior cells */
  int32_t kind;             /* swe_kind */
  int32_t order;            /* 1 or 2 */
  int32_t limiter;          /* swe_limiter */
  int32_t reserved;
  double g;                 /* gravity; the reference's G = 9.80665 (swe.mjs:85) */
  double min_depth;         /* minDepth (swe.mjs:172) */
  double eps4;              /* max(minDepth,1e-12)**4 as the host engine computed it (swe.mjs:190) */
  double cfl;
  double coriolis;          /* Cartesian f (swe.mjs:172) */
  double manning;           /* scalar n */
  double R;                 /* sphere radius (geometry.mjs:191); ignored for Cartesian */
  const double *rows[SWE_NROWS];   /* host rows, H each; AREA may be NULL */
} swe_desc;

typedef struct swe_ctx swe_ctx;

/* new ShallowWater(opts) minus the closure sampling (swe.mjs:170-244): allocates
 * device buffers for a W*H grid and uploads the geometry rows. */
int swe_create(const swe_desc *desc, swe_ctx **out);
void swe_destroy(swe_ctx *ctx);
const char *swe_last_error(const swe_ctx *ctx);
/* x**y for finite x > 0 exactly as the device computes it (fdlibm e_pow, the
 * algorithm behind the JS engine's Math.pow to within 1 ulp); the host uses it
 * for eps4 = max(minDepth,1e-12)**4 (swe.mjs:190). */
double swe_js_pow(double x, double y);
/* hllc(out, hL, huL, hvL, hR, huR, hvR, dry) (swe.mjs:974-1011): the interface flux
 * the kernels use, compiled for the host; the JS module's `hllc` export calls it. */
void swe_hllc(double out[3], double hL, double huL, double hvL, double hR, double huR, double hvR,
              double dry, double g);
/* version string and the gfx target the library was built for */
const char *swe_build_info(void);

/* Page-locked host buffers (no reference counterpart: the reference's arrays are
 * ordinary Float64Arrays).  Fields kept in them move between host and device at
 * full DMA speed in swe_set_field / swe_get_field; the Node drop-in backs the
 * public arrays of ShallowWater with them.  Contents are not cleared. */
int swe_host_alloc(size_t bytes, void **out);
void swe_host_free(void *p);

/* Host <-> device copies of whole fields (the JS object's Float64Array mirrors).
 * SWE_F_MANNING: setManningField (swe.mjs:713-731); NULL returns to the scalar.
 * SWE_F_SPONGE: per-cell makeSponge strengths (swe.mjs:1174-1186); NULL removes it. */
int swe_set_field(swe_ctx *ctx, int field, const double *host);
int swe_get_field(swe_ctx *ctx, int field, double *host);
/* replace one geometry row (verify-sphere mutates sim.geom between steps) */
int swe_set_row(swe_ctx *ctx, int row, const double *host);
int swe_set_bc(swe_ctx *ctx, int side, int kind);
int swe_set_scalar(swe_ctx *ctx, int which, double value);
int swe_get_scalar(swe_ctx *ctx, int which, double *value);

/* applyBC() (swe.mjs:289-303). ext[2*side], ext[2*side+1] = etaExt(t), uExt(t) of a
 * flather side, evaluated by the caller at the current t. */
int swe_apply_bc(swe_ctx *ctx, const double ext[8]);
/* residual(dt) (swe.mjs:308-629) on the stored state, ghost cells as stored;
 * fills SWE_F_RH/RHU/RHV/THETA. */
int swe_residual(swe_ctx *ctx, double dt);
/* applyFriction(dt) (swe.mjs:659-704) on the stored state: semi-implicit Manning,
 * scalar n or the per-cell field. */
int swe_apply_friction(swe_ctx *ctx, double dt);
/* dryClean() (swe.mjs:734-745) on the stored state; *added = mass invented by
 * flooring h (summed per block, then in block order). */
int swe_dry_clean(swe_ctx *ctx, double *added);
/* maxDt() (swe.mjs:632-647); +inf when every interior cell is dry */
int swe_max_dt(swe_ctx *ctx, double *out);
/* step(dt) (swe.mjs:754-814) without the host forcing hook: dt NaN = self-timed.
 * *dt_taken = 0 when the step was refused (dt not > 0 or not finite).
 * want_theta/want_rh: also fill THETA / RH* with stage 2's values (what the
 * reference leaves in sim.theta / sim.rh after a step). */
int swe_step(swe_ctx *ctx, double dt, const double ext[8], int want_theta, int want_rh,
             double *dt_taken, double *mass_floored);
/* step(dt) split at its second applyBC (swe.mjs:784-812), for boundaries the
 * device does not implement (a caller's own boundary function, run on the host
 * over the caller's arrays; SWE_BC_STORED sides keep the ghost cells the caller
 * uploaded).  swe_step_begin: dt (NaN: maxDt of the current state), applyBC of
 * the device sides and stage 1 (swe.mjs:784-796).  Until swe_step_end the fields
 * SWE_F_H/HU/HV are the stage-1 state, every cell as the reference leaves it
 * (ghosts h + dt*rh), for the caller to download, fill the ghost ring of and
 * upload.  swe_step_end: applyBC of the device sides, stage 2, friction, sponge
 * and the next CFL step (swe.mjs:798-812); the result becomes the current state.
 * ext: the flather values of the step (NULL in swe_step_end keeps swe_step_begin's).
 * *dt_taken = 0: the step was refused (swe.mjs:784) and no swe_step_end follows.
 * Between the two calls only field copies, scalars and swe_step_end are accepted. */
int swe_step_begin(swe_ctx *ctx, double dt, const double ext[8], double *dt_taken);
int swe_step_end(swe_ctx *ctx, const double ext[8], int want_theta, int want_rh, double *mass_floored);
/* 1: every later step (swe_step, swe_step_end, swe_advance) also leaves stage 2's
 * theta and rh/rhu/rhv in SWE_F_THETA/RH/RHU/RHV, as the reference's step leaves
 * them in sim.theta and sim.rh after a step (swe.mjs:798-806). */
int swe_keep_residual(swe_ctx *ctx, int on);
/* The Flather (eta_ext, u_ext) per side the next step will use: the caller's
 * values after swe_step / swe_apply_bc, the device-evaluated harmonics
 * (swe_set_harmonic) during and after swe_advance. */
int swe_get_ext(swe_ctx *ctx, double ext[8]);
/* Copy only the ghost ring (two cells deep) of SWE_F_B/H/HU/HV into a W*H host
 * array, leaving its interior untouched: what applyBC (swe.mjs:289-303) changes. */
int swe_get_ghosts(swe_ctx *ctx, int field, double *host);
/* Device-resident run of nsteps steps (no host round trip): dt policy SELF
 * (maxDt every step), FIXED (dt_value) or UNTIL (min(maxDt, t_end - t)); the
 * flather forcing comes from swe_set_harmonic.  An UNTIL (or SAMPLED) run ends
 * once t reaches t_end: the steps launched past it leave the state untouched,
 * and the host stops launching within 256 steps. */
int swe_advance(swe_ctx *ctx, int64_t nsteps, int dt_mode, double dt_value, double t_end,
                int64_t *steps_taken);
/* Equilibrium-tide body force (no reference counterpart: the reference leaves
 * such forcing to a host hook, swe.mjs:242).  n <= 4 constituents of order m:
 * 2 (semidiurnal) eta_eq = amp cos^2(lat) cos(omega t + 2 lon + phase), or
 * 1 (diurnal) eta_eq = amp sin(2 lat) cos(omega t + lon + phase).  Both stages
 * add g h grad(eta_eq) to the momentum residual, at the stage times t and t + dt,
 * so a resting ocean relaxes to eta = eta_eq + const.  lat: H cell-centre
 * latitudes, lon: W cell-centre longitudes, radians.  Spherical grids only;
 * n = 0 removes it. */
int swe_set_tide(swe_ctx *ctx, int n, const int32_t *m, const double *amp, const double *omega,
                 const double *phase, const double *lat, const double *lon);
/* Flather forcing for swe_advance: eta(t) = msl + sum amp_k cos(omega_k t - phase_k),
 * u(t) = u0 + sum uamp_k cos(uomega_k t - uphase_k), k < 4 (Tide.eta, tide.mjs:103-107). */
int swe_set_harmonic(swe_ctx *ctx, int side, int n_eta, double msl, const double *amp,
                     const double *omega, const double *phase, int n_u, double u0,
                     const double *uamp, const double *uomega, const double *uphase);
/* The wavemaker form of a side's harmonic forcing (src/waves.mjs:163-175,
 * regularWave): every amplitude faded in by min(1, t / ramp) (ramp > 0; 0 =
 * none) and, with u_from_eta, u(t) = u_scale (eta(t) - msl) in place of the
 * u harmonics.  Call after swe_set_harmonic for that side. */
int swe_set_harmonic_ramp(swe_ctx *ctx, int side, double ramp, int u_from_eta, double u_scale);
/* capture one device-resident step in a HIP graph and replay it (1 = on) */
int swe_use_graph(swe_ctx *ctx, int on);
/* rows of grid per wavefront chunk of the fused stage kernel (autotuned; 0 = default) */
int swe_set_chunk_rows(swe_ctx *ctx, int rows);
/* Measurement hook (no reference counterpart): run nsteps self-timed device
 * steps with HIP events around each kernel on the context's stream and return
 * the mean duration in ms of [stage-1 sweep, stage-2 sweep, finalize]. */
int swe_profile_step(swe_ctx *ctx, int nsteps, double out_ms[3]);
/* Autotuner (no reference counterpart): time the two stage kernels on the
 * current state for a set of row-chunk heights (fixed ones and ones that fill
 * whole rounds of the device's resident wavefront slots), keep the fastest and
 * return it.  Leaves the state, t and the step count untouched. */
int swe_autotune(swe_ctx *ctx, int reps, int *best_rows);
/* block until all work on the context's stream is done */
int swe_sync(swe_ctx *ctx);
/* device pointer of a field of the current state (for zero-copy callers) */
int swe_device_ptr(swe_ctx *ctx, int field, double **dptr);
/* elapsed ms of the last swe_advance measured with HIP events on the context's stream */
int swe_last_advance_ms(swe_ctx *ctx, double *ms);
/* The reference's diagnostics of the current device state, computed on the
 * device with no copy of the fields (volume swe.mjs:819-837, energy :864-879,
 * maxSurfaceDeviation(eta_ref) :882-894, maxSpeed :897-909, finite :912-917;
 * out[SWE_D_FINITE] is 1 or 0).  Energy, the spherical volume, both maxima and
 * finite equal the reference's bit for bit (the same per-row and row orders);
 * the Cartesian volume, one running sum over all cells in the reference, is the
 * exactly rounded sum here.  SWE_S_DIAG_MS gives the kernels' duration. */
int swe_diagnostics(swe_ctx *ctx, double eta_ref, double out[SWE_D_COUNT]);
/* swe_diagnostics on fields the caller owns, allocating nothing.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * b, h, hu, hv: device pointers to W*H = (nx+4)*(ny+4) doubles in the layout of
 * the fields above; area: device pointer to the H row areas (SWE_SPHERE only,
 * else NULL); dx, dy: the Cartesian spacing; min_depth, eps4, g as in swe_desc.
 * Same kernels and the same results as swe_diagnostics.  Its scratch (row sums,
 * block partials, maxima) is a static device array of the library sized for
 * ny <= SWE_DIAG_NOALLOC_MAX_ROWS (about 300 KB per GPU, held until the process
 * ends); a larger ny is refused with SWE_EINVAL.  Runs on the null stream and
 * returns when out[] is filled; the caller's fields must be ready (every other
 * swe_* call returns with its work done).  Calls that overlap on one GPU race on
 * that array; errors are read with swe_last_error(NULL). */
#define SWE_DIAG_NOALLOC_MAX_ROWS 16384
int swe_diagnostics_noalloc(int32_t nx, int32_t ny, int32_t kind, const double *b, const double *h,
                            const double *hu, const double *hv, const double *area, double dx, double dy,
                            double min_depth, double eps4, double g, double eta_ref, double out[SWE_D_COUNT]);
/* Gauge recorder of swe_advance (no reference counterpart: tools/waves.mjs:211-232
 * records sim.eta(i, j) = b + h at chosen cells from JS after each step).  Arms
 * the recorder with ng cells (indices into W*H) and room for `capacity` samples:
 * every > 0 samples after every every-th step of later swe_advance calls; every
 * = 0 samples at the times t0, t0 + dt_rec, ... as waves.mjs does, each sample
 * taken after the first step that ends at or after its time, and
 * swe_advance(SWE_DT_SAMPLED, t_end) shortens steps to land on those times
 * (min(maxDt, max(1e-6, min(t_end, next) - t)), waves.mjs:222-223).  ng = 0
 * disarms it.  Clears earlier samples. */
int swe_set_recorder(swe_ctx *ctx, int ng, const int32_t *cells, int every, double t0, double dt_rec,
                     int64_t capacity);
/* The samples so far: eta (min(count, capacity) rows of ng values) and their
 * times t; either may be NULL.  SWE_ESTATE when more samples were due than fit
 * (the first `capacity` are returned). */
int swe_get_record(swe_ctx *ctx, double *eta, double *times, int64_t *count);
/* Projection recorder of swe_advance (no counterpart in swe.mjs: the mode
 * projection of src/globe.mjs:1417-1424, which measureModePeriod evaluates on
 * the host after every step): after each step that runs, s = sum over interior
 * rows j of coef[j] * (sum over the row of b + h), each sum left to right in
 * the reference's order, is appended with the step's end time.  coef: ny
 * values (globe.mjs uses (wRow_j / nx) * P_n(sin lat_j)); NULL disarms.
 * Clears earlier samples. */
int swe_set_projection(swe_ctx *ctx, const double *coef, int64_t capacity);
/* The projection samples so far (min(count, capacity) of them) and their times;
 * either may be NULL.  SWE_ESTATE when more were due than fit. */
int swe_get_projection(swe_ctx *ctx, double *vals, double *times, int64_t *count);

/* ---- batched ensembles (no reference counterpart: the reference runs one
 * ShallowWater per JS object; these advance many uneven basins in one launch
 * per kernel phase, each bit-identical to its own plain run) ------------------ */
typedef struct swe_batch swe_batch;
int swe_batch_create(const swe_desc *descs, int nbasins, swe_batch **out);
void swe_batch_destroy(swe_batch *b);
const char *swe_batch_last_error(const swe_batch *b);
int swe_batch_set_field(swe_batch *b, int basin, int field, const double *host);
int swe_batch_get_field(swe_batch *b, int basin, int field, double *host);
int swe_batch_set_bc(swe_batch *b, int basin, int side, int kind);
int swe_batch_set_harmonic(swe_batch *b, int basin, int side, int n_eta, double msl, const double *amp,
                           const double *omega, const double *phase, int n_u, double u0,
                           const double *uamp, const double *uomega, const double *uphase);
int swe_batch_get_scalar(swe_batch *b, int basin, int which, double *value);
/* swe_get_ext for one basin of the batch */
int swe_batch_get_ext(swe_batch *b, int basin, double ext[8]);
/* swe_set_tide for one basin of the batch */
int swe_batch_set_tide(swe_batch *b, int basin, int n, const int32_t *m, const double *amp, const double *omega,
                       const double *phase, const double *lat, const double *lon);
/* rows per task and persistent-grid wavefronts (nwaves <= 0 keeps the current) */
int swe_batch_set_chunk_rows(swe_batch *b, int rows, int nwaves);
int swe_batch_use_graph(swe_batch *b, int on);
/* nsteps self-timed steps of every basin (each its own dt); *min_steps = fewest
 * steps any basin has taken in total */
int swe_batch_advance(swe_batch *b, int64_t nsteps, int64_t *min_steps);
int swe_batch_last_advance_ms(swe_batch *b, double *ms);
/* autotuner: times `steps` steps (the faster of two runs, after one untimed step)
 * for each (rows, wavefronts) candidate on a copy of the current state, restores
 * the state, keeps the fastest */
int swe_batch_autotune(swe_batch *b, int steps, int *best_rows, int *best_waves);

/* ---- discrete adjoint (no reference counterpart: the reference has no
 * gradients; these differentiate nsteps self-timed step() calls, swe.mjs:754-814,
 * with the dt sequence held at the forward's values) ---------------------------
 * Loss: J = sum over samples s and gauges g of (eta_g(t_s) - obs[s][g])^2, eta
 * sampled after every `every` steps.  Gradients: dJ/dh0 (initial depth, zero
 * momentum; dJ/d eta0 on wet cells) and dJ/dn (per-cell Manning n, interior).
 * Boundaries: reflect, outflow, periodic and flather (its far-field values as the
 * forward used them: swe_set_harmonic's, evaluated on the device, or the last
 * ones given); the makeSponge forcing is differentiated too.  The adjoint drives the
 * context it is created on: its state after any swe_adj_* call is that of the
 * last forward step run, and a scalar Manning n becomes a per-cell field. */
typedef struct swe_adj swe_adj;
/* ckpt: steps between checkpoints; 0 chooses round(sqrt(nsteps)) when that plan
 * fits in 90% of the free device memory, else two levels (swe_adj_create_levels)
 * with inner up to round(cbrt(nsteps)) (smaller until the plan fits) and
 * outer = inner^2 */
int swe_adj_create(swe_ctx *ctx, int64_t nsteps, int ckpt, swe_adj **out);
/* Two-level checkpointing: the forward keeps the state every `outer` steps (a
 * multiple of `inner`); the backward recomputes, per outer segment, the state
 * every `inner` steps, then replays and reverses those inner segments.  Costs one
 * more forward pass than one level (outer == inner) and needs far less memory. */
int swe_adj_create_levels(swe_ctx *ctx, int64_t nsteps, int64_t inner, int64_t outer, swe_adj **out);
/* Device bytes an adjoint of nsteps steps on a grid of ncells = W*H cells needs
 * with the given intervals (nothing is allocated) */
int swe_adj_plan(int64_t nsteps, int64_t ncells, int64_t inner, int64_t outer, double *bytes);
void swe_adj_destroy(swe_adj *a);
const char *swe_adj_last_error(const swe_adj *a);
/* cells: k = j*W + i of interior cells; obs: nsteps/every samples of ng values */
int swe_adj_set_gauges(swe_adj *a, int ng, const int32_t *cells, int every, const double *obs);
/* self-timed forward from (h0, hu = hv = 0) with Manning field nfield (NULL keeps
 * the current one); records dt per step and checkpoints; returns J */
int swe_adj_forward(swe_adj *a, const double *h0, const double *nfield, double *loss);
/* J of a forward that replays the recorded dt sequence (finite-difference checks) */
int swe_adj_loss(swe_adj *a, const double *h0, const double *nfield, double *loss);
/* dJ/dh0 and dJ/dn of the last swe_adj_forward (W*H each; either may be NULL) */
int swe_adj_backward(swe_adj *a, double *grad_h0, double *grad_n);
int swe_adj_get_series(swe_adj *a, double *eta);   /* nsteps/every * ng gauge values */
int swe_adj_get_dt(swe_adj *a, double *dt);        /* nsteps recorded dt */
/* HIP-event ms of the last swe_adj_forward and swe_adj_backward */
int swe_adj_timing(swe_adj *a, double *fwd_ms, double *bwd_ms);
/* The reverse passes run on host arrays (no GPU needed), for checking them:
 * lh += (d residual / d state)^T lr for residual(dt) of a state with stored
 * ghosts (swe.mjs:308-629); BC^T in place (reflect/outflow/periodic/flather,
 * swe.mjs:289-303, :1084-1116: a flather side differentiates at the state
 * h, hu, hv, b applyBC read with far-field values ext, which may be NULL
 * without one); the adjoint of the pre-dryClean state of a stage's end (stage 1:
 * dryClean; stage 2: dryClean, then friction with Manning field nfield or the
 * scalar, accumulating dJ/dn into lN, then the sponge relaxation with per-cell
 * strengths `sponge` or none, swe.mjs:1188-1197); hllc^T (swe.mjs:974-1011). */
int swe_adj_host_residual_rev(const swe_desc *d, const double *h, const double *hu, const double *hv,
                              const double *b, double dt, const double *lrh, const double *lrhu,
                              const double *lrhv, double *lh, double *lhu, double *lhv);
int swe_adj_host_bc_rev(const swe_desc *d, const int32_t bc[4], const double ext[8], const double *h,
                        const double *hu, const double *hv, const double *b, double *lh, double *lhu, double *lhv);
/* The same transpose in the gather form the device runs where no side is Flather,
 * the south and north sides are not periodic and nx, ny >= 4 (SWE_EINVAL
 * otherwise): one work item per source column or row, adding its ghosts' folds in
 * the loop's order.  No counterpart in the reference: a host check of the device
 * pass. */
int swe_adj_host_bc_rev_gather(const swe_desc *d, const int32_t bc[4], const double ext[8], const double *h,
                               const double *hu, const double *hv, const double *b, double *lh, double *lhu,
                               double *lhv);
/* The same transpose folded on read, as the device's reverse epilogues apply it
 * in the configurations of the gather form (SWE_EINVAL otherwise): each interior
 * cell adds the folds of the ghosts whose source it is, in the loop's order, from
 * the ghosts as given; the ghosts are then cleared.  ext, h, hu, hv and b are not
 * read.  No counterpart in the reference: a host check of the device pass. */
int swe_adj_host_bc_rev_onread(const swe_desc *d, const int32_t bc[4], const double ext[8], const double *h,
                               const double *hu, const double *hv, const double *b, double *lh, double *lhu,
                               double *lhv);
int swe_adj_host_epilogue_rev(const swe_desc *d, int stage, const double *pre_h, const double *pre_hu,
                              const double *pre_hv, const double *nfield, const double *sponge, double dt,
                              const double *lout_h,
                              const double *lout_hu, const double *lout_hv, double *lpre_h, double *lpre_hu,
                              double *lpre_hv, double *lN);
void swe_adj_host_hllc_rev(double out[6], double hL, double huL, double hvL, double hR, double huR, double hvR,
                           double dry, double g, double lm, double ln, double lt);

#ifdef __cplusplus
}
#endif
#endif
