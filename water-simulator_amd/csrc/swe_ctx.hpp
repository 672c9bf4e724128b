// Private to the library: the plain-basin context behind swe_ctx*, shared by
// the C ABI (swe_capi.hip) and the adjoint (swe_adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/swe_amd.h"
#include "swe_kernels.hpp"

namespace swe {
hipError_t launch_sweep(int mode, const SweepArgs &a, hipStream_t st);
int sweep_waves(const Params &P, int rows_per_chunk);
const void *k_sweep_stage1_ptr();
hipError_t launch_apply_bc(const Params &P, StateW S, double *b, const Ctl *ctl, hipStream_t st);
hipError_t launch_maxdt(const Params &P, StateP S, Ctl *ctl, hipStream_t st);
// rec: the gauge recorder of a device-resident step (nullptr: none)
hipError_t launch_finalize(Ctl *ctl, const double *mf_part, int nparts, double cfl, int phase, hipStream_t st,
                           const RecArgs *rec = nullptr);
hipError_t launch_friction(const Params &P, StateW S, const double *mfield, double dt, hipStream_t st);
hipError_t launch_dry_clean(const Params &P, StateW S, double *part, int blocks, hipStream_t st);
hipError_t launch_tide_times(Ctl *ctl, hipStream_t st);
// the per-constituent coefficient table of the equilibrium-tide body force
// (2H + 2W doubles each; see Params::tide), SWE_EINVAL on bad arguments
int tide_table(const Params &P, int n, const int32_t *m, const double *amp, const double *lat, const double *lon,
               std::vector<double> &out);

// rows per wavefront chunk of the stage kernels: any height from 1 gives the
// same results; the default never goes below DEFAULT_MIN_CHUNK, the autotuner
// tries shorter chunks on grids too small to fill the device
constexpr int MIN_CHUNK = 1;
constexpr int DEFAULT_MIN_CHUNK = 8;
// device row arrays of a context: the divisors, their correctly rounded
// reciprocals, and the coefficients
enum { R_DX, R_RDX, R_DYN, R_RDYN, R_DYS, R_RDYS, R_DYCFL, R_RDYCFL, R_BED, R_RBED, R_TAN, R_F, R_GEO,
       R_AREA, R_COUNT };
}  // namespace swe

struct swe_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  swe::Params P{};
  size_t N = 0;
  double *rows[swe::R_COUNT] = {};
  double *rowsA = nullptr;      // the rows interleaved per row (Params::geoA)
  std::vector<double> hrows[SWE_NROWS];
  double *b = nullptr;
  double *st[3][3] = {};       // [buffer][h,hu,hv]
  int A = 0, Bi = 1, C = 2;     // A = current state
  double *mfield = nullptr, *sponge = nullptr;
  double *rh = nullptr, *rhu = nullptr, *rhv = nullptr, *theta = nullptr;
  swe::Ctl *dctl = nullptr;
  swe::Ctl *hctl = nullptr;     // pinned host shadow
  double *mf_part = nullptr;
  int max_parts = 0;
  int chunk = 0;
  double manning = 0.0, coriolis = 0.0;
  double *tide = nullptr;      // equilibrium-tide coefficient table (tide_n constituents)
  int tide_n = 0;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  hipGraph_t graph[2] = {nullptr, nullptr};
  bool use_graph = false;
  int graph_chunk = -1;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  // gauge recorder of swe_advance: cells, rec_cap sample rows of rec_n values, times
  int *rec_cells = nullptr;
  double *rec_series = nullptr, *rec_times = nullptr;
  int rec_n = 0;
  long long rec_cap = 0;
  // device diagnostics: per-row sums, per-block partials, maxima, result
  // projection recorder of swe_advance: ny coefficients, proj_cap samples and times
  double *proj_coef = nullptr, *proj_vals = nullptr, *proj_times = nullptr;
  long long proj_cap = 0;
  double *diag = nullptr;
  int diag_blocks = 0;
  double diag_ms = 0.0;       // duration of the last swe_diagnostics' kernels
  // Flather harmonics set by swe_set_harmonic: device-resident runs evaluate them
  // every step; swe_step uses the caller's values for its own step only
  int harm_enabled = 0;
  // a step split around a host ghost fill (swe_step_begin .. swe_step_end) is in
  // progress: the current fields are the stage-1 state in buffer Bi, and the
  // stage kernels' massFloored partials of stage 1 (mid_np of them) wait for stage 2
  int mid = 0, mid_np = 0;
  // every step also leaves stage 2's theta and residual in the probe fields
  int keep_resid = 0;
  std::string err;
};

#define HIPCHK(ctx, expr)                                                        \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);            \
      return SWE_EHIP;                                                           \
    }                                                                            \
  } while (0)

namespace swe {
int ctx_fail(swe_ctx *c, int code, const std::string &msg);
int ctx_pull(swe_ctx *c);
int ctx_push(swe_ctx *c);
SweepArgs ctx_base_args(swe_ctx *c);
int ctx_enqueue_step(swe_ctx *c, int want_theta, int want_rh);
// the diagnostics' device scratch (allocated on first use)
int ensure_diag(swe_ctx *c);
// the projection recorder's kernels on a state (after a step's finalize)
hipError_t launch_projection(swe_ctx *c, const double *h, const double *b);
}  // namespace swe
