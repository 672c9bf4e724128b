// Kernel-side data layout and launch interface for the shallow-water step.
//
// Layout in HBM (one basin): every field is a row-major array of W*H doubles,
// W = nx+4, H = ny+4 (two ghost cells per side), cell k = j*W + i, exactly the
// reference's indexing (swe.mjs:196-199).  The state lives in three rotating
// buffers {h,hu,hv}[3]: stage 1 reads A and writes B, stage 2 reads B (stencil)
// and A (pointwise) and writes C, then A and C swap.  b is shared.  Geometry is
// 13 per-row arrays of length H (divisors and their correctly rounded
// reciprocals).  Before each stage k_apply_bc fills the ghost ring of the
// stage's input state (the reference's applyBC, swe.mjs:289-303, in its order),
// and the stage kernels read the ghosts from memory like any other cell.
#pragma once
#include <hip/hip_runtime.h>

namespace swe {

enum BcKind { BC_REFLECT = 0, BC_OUTFLOW = 1, BC_PERIODIC = 2, BC_FLATHER = 3, BC_STORED = 4 };
enum Mode { MODE_STAGE1 = 1, MODE_STAGE2 = 2, MODE_PROBE = 3 };
enum DtMode { DT_SELF = 0, DT_FIXED = 1, DT_UNTIL = 2, DT_HOST = 3, DT_SAMPLED = 4 };
enum { MAX_HARM = 4, MAX_TIDE = 4 };

struct Geo {
  const double *dx, *rdx, *dyN, *rdyN, *dyS, *rdyS, *dyCFL, *rdyCFL, *bedPhi, *rbedPhi,
      *tanPhi, *fRow, *geo;
};

// Harmonic boundary forcing evaluated on the device for device-resident runs:
// eta(t) = msl + sum_k amp_k cos(omega_k t - phase_k); u(t) likewise.
// With ramp > 0 every amplitude is scaled by min(1, t / ramp) (a wavemaker
// fading in); with u_from_eta, u(t) = u_scale (eta(t) - msl) instead of its own
// harmonics (the shallow-water characteristic of src/waves.mjs:163-175).
struct Harm {
  int n_eta, n_u, u_from_eta, pad1;
  double msl, u0;
  double amp[MAX_HARM], omega[MAX_HARM], phase[MAX_HARM];
  double uamp[MAX_HARM], uomega[MAX_HARM], uphase[MAX_HARM];
  double ramp, u_scale;
};

// Device-resident control block of one basin.
struct Ctl {
  double dt;           // dt of the step about to run (or just run)
  double t;            // step-start time
  double ext[8];       // Flather (eta_ext, u_ext) per side at t
  double dt_next;      // cfl / max wave speed of the state after the last step (inf if dry)
  double mf;           // massFloored of the last step
  double dt_fixed, t_end;
  unsigned long long cfl_bits;   // max-reduction accumulator (bit pattern of a non-negative double)
  long long steps;
  int dt_mode, nonfinite, harm_on;
  int idle;            // DT_SAMPLED / DT_UNTIL run at t_end: the remaining launched steps do nothing
  double dt_last;      // dt the last step ran with (0 if it was refused)
  Harm harm[4];
  // equilibrium-tide constituents: angular frequency and phase, and
  // (cos, sin)(omega t_s + phase) at the two stage times t_s = t, t + dt of the
  // step about to run (written wherever dt is set on the device)
  int tide_n, tide_pad;
  double tide_om[MAX_TIDE], tide_ph[MAX_TIDE];
  double tide_cs[2][MAX_TIDE][2];
  // gauge recorder of device-resident runs: after a step that ran, eta = b + h at
  // the recorder's cells goes to sample row rec_row when rec_flag is set; samples
  // are taken every rec_every-th step, or (rec_every == 0) at the first step
  // ending at or after each sample time rec_next = rec_t0, rec_t0 + rec_dt, ...
  int rec_every, rec_flag;
  long long rec_count, rec_row, rec_since;
  double rec_next, rec_t0, rec_dt;
  long long proj_count;   // samples of the projection recorder so far
};

// Gauge recorder arguments of k_finalize (cells == nullptr: no recorder).
struct RecArgs {
  const double *h, *b;   // the state after the step
  const int *cells;
  double *series;        // rec_cap rows of n values
  double *times;         // rec_cap sample times
  int n;
  long long cap;
};

struct Params {
  int nx, ny, W, H;
  int sphere, order, limiter, coriolis_on;
  int bc[4];
  int friction;        // 0 none, 1 scalar n, 2 per-cell field
  int sponge;          // per-cell sponge strengths present
  int want_theta, want_rh;
  double g, half_g, neg_half_g, rg, min_depth, eps4, cfl, coriolis, manning, R, sponge_ref;
  Geo geo;
  // equilibrium-tide body force: tide_n constituents, each 2H + 2W doubles in
  // `tide` (per row X, Y coefficients; per column cos, sin of m lon); a probe
  // sweep evaluates it at the second stage time when probe_stage2 is set
  int tide_n, probe_stage2;
  const double *tide;
  // Layout contract of the stage kernels: the hu and hv of a state buffer (and
  // rhu, rhv of the residual outputs) sit fstride doubles after its h; the
  // per-row arrays sit rstride doubles apart in the order dx, rdx, dyN, rdyN,
  // dyS, rdyS, dyCFL, rdyCFL, bedPhi, rbedPhi, tanPhi, fRow, geo starting at
  // geo.dx.  The kernels then carry one base pointer per buffer instead of
  // three, and one for all the rows.
  int fstride, rstride;
  // The same 13 per-row values interleaved for the stage kernels:
  // geoA[j * GEO_AOS + k] is array k (the order above) at row j, so a row's
  // values are one scalar base plus immediate offsets (no per-array offsets
  // held in scalar registers across the row loop).
  const double *geoA;
  // geo_uni: every row of geoA holds the same values (a Cartesian grid): the
  // stage kernels then read them once per launch instead of once per row
  int geo_uni, geo_pad;
};
constexpr int GEO_AOS = 16;
// slot of geoA beside the 13 arrays above (and the host's area row): tanPhi[j] / R,
// which the reference forms once per row (swe.mjs:610) and the stage kernels read
constexpr int GEO_TANR = 14;

// Params::geo_uni from the interleaved rows a[j * GEO_AOS + k], j < H.  Never
// set on a sphere (whose rows differ by latitude anyway): the uniform-row stage
// kernels leave out the spherical terms and the equilibrium tide.
inline void set_uniform_rows(Params &P, const double *a, int H) {
  bool uni = !P.sphere;
  for (int j = 1; j < H && uni; ++j)
    if (__builtin_memcmp(&a[(size_t)j * GEO_AOS], a, sizeof(double) * GEO_AOS) != 0) uni = false;
  P.geo_uni = uni ? 1 : 0;
}

struct StateP { const double *h, *hu, *hv; };
struct StateW { double *h, *hu, *hv; };

struct SweepArgs {
  Params P;
  StateP in;          // stencil input (stage 1: A, stage 2: B, probe: A with stored ghosts)
  StateP h0;          // stage 2: A (pointwise)
  StateW out;         // stage 1: B, stage 2: C
  double *b;          // bed (ghost cells rewritten by stage 2)
  double *rh, *rhu, *rhv, *theta;   // probe outputs (or optional stage-2 outputs)
  const double *mfield, *sponge;
  Ctl *ctl;
  double *mf_part;    // one partial per wave
  double *rec;        // stage 2, optional: the dryCleaned, pre-friction interior state (adjoint replay)
  // optional (adjoint replay): per cell, fstride apart, its velocities U, V and
  // surface, its limited slopes (eta, U, V, b; x then y each: sEx sEy sUx sUy
  // sVx sVy sBx sBy), the M, N+pL, N+pR, T flux terms of its east x-face, those
  // of its north y-face, its positivity factor theta and outgoing flux sum (zero /
  // theta = 1 where the stage has none): REC_FIELDS fields
  double *recf;
  int rows_per_chunk;
};

// One basin of a batched ensemble as the batch kernels see it.  The three state
// buffers rotate in lockstep for every basin, so each basin carries the stage
// arguments for both rotation parities: sw[parity][stage-1].  bc_in[parity][s] is
// the state whose ghost ring is filled before stage s+1.
struct BasinDev {
  SweepArgs sw[2][2];
  StateW bc_in[2][2];
  Ctl *ctl;
  int nparts;        // wavefronts (tasks) of one stage of this basin
  int pad;
};

// A unit of batched work: one wavefront's (strip, chunk) of one basin; `wave`
// is its massFloored slot in that basin.
struct Task { int basin, strip, chunk, wave; };

// (cos, sin)(omega t_s + phase) of every tide constituent at the stage times
// t_s = t, t + dt of the step about to run (t, dt: the control block's, passed
// in by the caller), one value per lane: lane q < 4 MAX_TIDE takes constituent
// q / 4, stage time (q / 2) & 1, cos or sin by q & 1, and returns that entry
// (whether or not the constituent is in use) for callers that keep a copy
__device__ __forceinline__ double tide_times_lane(Ctl *ctl, double t, double dt, int lane) {
  if (lane >= 4 * MAX_TIDE) return 0.0;
  const int c = lane >> 2, s = (lane >> 1) & 1, f = lane & 1;
  if (c < ctl->tide_n) {
    const double a = s ? ctl->tide_om[c] * (t + dt) + ctl->tide_ph[c] : ctl->tide_om[c] * t + ctl->tide_ph[c];
    ctl->tide_cs[s][c][f] = f ? sin(a) : cos(a);
  }
  return ctl->tide_cs[s][c][f];
}

constexpr int REC_FIELDS = 21;      // fields of SweepArgs::recf
constexpr int SWEEP_LANES = 64;
constexpr int SWEEP_HALO = 3;
constexpr int SWEEP_OUT = SWEEP_LANES - 2 * SWEEP_HALO;   // 58 output columns per wave

}  // namespace swe
