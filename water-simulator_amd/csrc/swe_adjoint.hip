// Discrete adjoint of the shallow-water step (swe_adj_* in include/swe_amd.h).
//
// One forward step at fixed dt (swe.mjs:754-814), with interior values the only
// ones that carry from step to step:
//   A' = BC(A); R1 = residual(A'); B = dryClean(A' + dt R1)          (stage 1)
//   B' = BC(B); R2 = residual(B'); C0 = 0.5 (A' + B' + dt R2)
//   C  = friction(dryClean(C0), n)                                      (stage 2)
// Its reverse, given the adjoint of C on the interior: friction^T and dryClean^T,
// the RK combination^T, residual^T of stage 2, BC^T, then the same for stage 1.
// residual^T runs as one-thread-per-cell passes over the grid with their
// intermediates in HBM (swe_adjoint.hpp): on the host all seven, on the device
// two tiled ones (accumulation + positivity^T + face^T, gather + limiter^T),
// since the replayed stage sweeps record the velocities, surface, slopes, face
// flux terms and positivity factors that the other passes recompute.  BC^T runs
// as passes of its own, or folded on read by the next epilogue (adj_bc_fold_into)
// where no side is Flather and south and north are not periodic.
//
// Time: the forward records dt per step on the device and keeps the state every
// K steps (K ~ sqrt(nsteps)); the reverse sweep re-runs each segment from its
// checkpoint with the forward's own stage kernels, writing each step's
// BC-applied stage inputs A', B' and stage 2's dryCleaned, pre-friction state
// straight into a segment store together with both stages' face flux terms and
// positivity factors, then steps back through the segment with no further
// recomputation.  The replay of segment s - 1 runs on a second stream into the
// other of two stores while segment s is reversed.  Memory: nsteps/K states of
// 3 W H doubles, two stores of K steps of 51 W H, plus 22 + 24 W H doubles of
// pass intermediates and scratch (about 56 GB at 1024x1024 and 4000 steps).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "swe_adjoint.hpp"
#include "swe_ctx.hpp"

using namespace swe;

namespace {

// ---- one-thread-per-cell kernels over the shared bodies -----------------------
// Workgroups are dealt to the 8 XCDs round-robin (block b to XCD b mod 8), and
// each XCD has its own L2.  The grids are a multiple of 8 blocks, and block b
// takes the cells of logical block (b mod 8) * (nblocks / 8) + b / 8, so each XCD
// sweeps one contiguous band of rows and finds the rows above and below a cell
// (which every gather pass reads) in its own L2.
constexpr int NXCD = 8;
__device__ __forceinline__ bool cell_of(const Params &P, int &i, int &j) {
  const int per = gridDim.x / NXCD;
  const int blk = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
  const int k = blk * blockDim.x + threadIdx.x;
  i = k % P.W;
  j = k / P.W;
  return j < P.H;
}

// Accumulation^T + positivity^T and face^T over AF_X x AF_Y tiles, one thread
// per cell: a cell's face flux-term adjoints stay in registers from the one pass
// to the other, and only theta^T's lout crosses cells (face^T of a face whose
// flow leaves the east or north neighbour reads that neighbour's): the tile keeps
// its cells' lout in LDS and evaluates the accumulation pass again, the same
// operations and so the same doubles, for the row north and the column east of
// it.  Writes theta^T's lh and the face-state adjoints; tiles are dealt to the
// XCDs in bands of tile rows as cell_of deals blocks.
constexpr int AF_X = 64, AF_Y = 8;
__global__ __launch_bounds__(AF_X * AF_Y) void k_adj_accum_faces_tiled(Params P, StateP S, const double *b,
                                                                      double dt, StateP lR, Work w) {
  __shared__ double lo[AF_Y + 1][AF_X + 1];   // lout, rows j0 .. j0 + AF_Y, columns i0 .. i0 + AF_X
  const int tiles_x = (P.W + AF_X - 1) / AF_X, tiles_y = (P.H + AF_Y - 1) / AF_Y;
  const int per = gridDim.x / NXCD;
  const int blk = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
  const int tx = blk % tiles_x, ty = blk / tiles_x;
  const int t = threadIdx.x, lx = t % AF_X, ly = t / AF_X;
  const int i0 = tx * AF_X, j0 = ty * AF_Y, i = i0 + lx, j = j0 + ly;
  const bool tile = ty < tiles_y;
  const bool own = tile && i < P.W && j < P.H;
  AccumOut a = {};
  if (own) {
    a = adj_accum_eval(P, S, dt, lR, w, i, j);
    lo[ly][lx] = a.lout;
    w.s[S_LHTH][j * P.W + i] = a.lh;
  }
  // the AF_X + AF_Y cells of the row north and the column east of the tile, on
  // the threads of the last wavefronts (as few as hold them)
  const int hq = t - (AF_X * AF_Y - 64 * ((AF_X + AF_Y + 63) / 64));
  if (tile && hq >= 0 && hq < AF_X + AF_Y) {
    const bool north = hq < AF_X;
    const int hi = north ? i0 + hq : i0 + AF_X, hj = north ? j0 + AF_Y : j0 + hq - AF_X;
    if (hi < P.W && hj < P.H) {
      const double v = adj_accum_eval(P, S, dt, lR, w, hi, hj).lout;
      if (north) lo[AF_Y][hq] = v; else lo[hq - AF_X][AF_X] = v;
    }
  }
  __syncthreads();
  if (!own) return;
  const double lf[8] = {a.lx[0], a.lx[1], a.lx[2], a.lx[3], a.ly[0], a.ly[1], a.ly[2], a.ly[3]};
  double r[12];
  // a neighbour's lout is read only for a face on the grid, whose neighbour is a grid cell
  adj_faces_rev_eval(P, b, w, i, j, lf, lo[ly][lx], lo[ly][lx + 1], lo[ly + 1][lx], r);
  const int k = j * P.W + i;
  for (int q = 0; q < 12; ++q) w.s[S_LXETAL + q][k] = r[q];
}
// Gather^T + limiter^T over 64 x 4 tiles of cells, one thread per cell: each
// cell's gather (cells_rev) runs once and leaves its slope adjoints in LDS, the
// tile's one-cell halo along each axis gets that axis's slope adjoints from
// slope_adj (the same operations, so the same doubles), and limiter^T + vel^T
// then read a neighbour's slope adjoints from LDS instead of recomputing them
// from memory.  Tiles are dealt to the XCDs in bands of tile rows as cell_of
// deals blocks.
constexpr int CT_X = 64, CT_Y = 4;
__global__ __launch_bounds__(CT_X * CT_Y) void k_adj_cells_slopes_tiled(Params P, StateP S, const double *b,
                                                                       StateP lR, Work w, const double *tcs,
                                                                       StateW lS) {
  __shared__ double sx[CT_Y][CT_X + 2][3];   // x-slope adjoints (eta, U, V), columns i0 - 1 .. i0 + CT_X
  __shared__ double sy[CT_Y + 2][CT_X][3];   // y-slope adjoints, rows j0 - 1 .. j0 + CT_Y
  const int tiles_x = (P.W + CT_X - 1) / CT_X, tiles_y = (P.H + CT_Y - 1) / CT_Y;
  const int per = gridDim.x / NXCD;
  const int blk = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
  const int tx = blk % tiles_x, ty = blk / tiles_x;
  const int t = threadIdx.x, lx = t % CT_X, ly = t / CT_X;
  const int i0 = tx * CT_X, j0 = ty * CT_Y, i = i0 + lx, j = j0 + ly;
  const bool tile = ty < tiles_y;
  const bool own = tile && i < P.W && j < P.H;
  CellRev c = {};
  if (own) {
    c = cells_rev(P, S, b, lR, w, tcs, i, j);
    sx[ly][lx + 1][0] = c.lsE[0]; sx[ly][lx + 1][1] = c.lsU[0]; sx[ly][lx + 1][2] = c.lsV[0];
    sy[ly + 1][lx][0] = c.lsE[1]; sy[ly + 1][lx][1] = c.lsU[1]; sy[ly + 1][lx][2] = c.lsV[1];
  }
  if (tile && t < 2 * CT_Y) {                           // x halo: row t / 2, west or east
    const int r = t >> 1, east = t & 1, hi = east ? i0 + CT_X : i0 - 1, hj = j0 + r;
    if (hi >= 0 && hi < P.W && hj < P.H) {
      const SlopeAdj a = slope_adj(P, b, lR, w, hi, hj, 0);
      double *q = sx[r][east ? CT_X + 1 : 0];
      q[0] = a.e; q[1] = a.u; q[2] = a.v;
    }
  } else if (tile && t < 2 * CT_Y + 2 * CT_X) {         // y halo: column (t - 2 CT_Y) / 2, south or north
    const int col = (t - 2 * CT_Y) >> 1, north = t & 1, hi = i0 + col, hj = north ? j0 + CT_Y : j0 - 1;
    if (hi < P.W && hj >= 0 && hj < P.H) {
      const SlopeAdj a = slope_adj(P, b, lR, w, hi, hj, 1);
      double *q = sy[north ? CT_Y + 1 : 0][col];
      q[0] = a.e; q[1] = a.u; q[2] = a.v;
    }
  }
  __syncthreads();
  if (!own) return;
  const int k = j * P.W + i;
  // limiter^T asks only for the slopes of grid cells that have one
  slopes_rev_core(P, S, w, lS, i, j, c.leta, c.lU, c.lV, w.s[S_LHTH][k], c.lhuC, c.lhvC, [&](int d, int o) {
    const double *q = d == 0 ? sx[ly][lx + 1 + o] : sy[ly + 1 + o][lx];
    SlopeAdj r = {q[0], q[1], q[2]};
    return r;
  });
}
__global__ void k_adj_bc(Params P, StateW l, int pass, StateP S, const double *b, const double *ext) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (pass == 0 ? 4 * P.W : 4 * P.ny)) adj_bc_ghost(P, l, pass, t, S, b, ext);
}
// one pass of BC^T in gather form (adj_bc_gather_ok): one work item per source
// column (pass 0) or row (pass 1), no atomics, the host loop's result
__global__ void k_adj_bc_gather(Params P, StateW l, int pass) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < (pass == 0 ? 2 * (P.W - 4) : 2 * P.ny)) adj_bc_gather(P, l, pass, q);
}
__global__ void k_adj_epilogue(Params P, StateP pre, StateP lOut, const double *nfield, const double *sponge,
                               double dt, double *lN, int stage, StateW lA, StateW lB, StateW lR, int fold) {
  int i, j;
  if (cell_of(P, i, j)) adj_epilogue_rev_at(P, pre, lOut, nfield, sponge, dt, lN, stage, lA, lB, lR, i, j, fold != 0);
}

__global__ void k_fill(size_t n, double *a, double v) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = v;
}
// the step about to run takes dt = dts[n] (and starts at t = ts[n] when ts is
// given) and the recorded Flather values exts[8 n ..]; the stage-time tide
// factors follow, and are also kept in tcs_out
__global__ void k_ctl_dt(Ctl *ctl, const double *dts, const double *ts, const double *exts, int64_t n,
                         double *tcs_out) {
  const int lane = threadIdx.x;
  const double dt = dts[n], t = ts ? ts[n] : ctl->t;
  if (lane == 0) {
    ctl->dt = dt;
    if (ts) ctl->t = t;
    for (int q = 0; q < 8; ++q) ctl->ext[q] = exts[8 * n + q];
    ctl->dt_mode = DT_HOST;
    ctl->cfl_bits = 0ull;
  }
  // the tide factors, one per lane; lane q's entry is tide_cs[s][c][f] with
  // c = q / 4, s = (q / 2) & 1, f = q & 1
  const double v = tide_times_lane(ctl, t, dt, lane);
  if (tcs_out && lane < 4 * MAX_TIDE) tcs_out[((lane >> 1) & 1) * 2 * MAX_TIDE + (lane >> 2) * 2 + (lane & 1)] = v;
}
// self-timed steps from here on, from t = 0, the Flather harmonics evaluated
// on the device when the context has them; the max-reduction starts from zero
__global__ void k_ctl_self(Ctl *ctl, int harm) {
  if (threadIdx.x == 0) {
    ctl->dt_mode = DT_SELF; ctl->cfl_bits = 0ull; ctl->t = 0.0; ctl->steps = 0; ctl->harm_on = harm;
  }
}
__global__ void k_ctl_record(const Ctl *ctl, double *table, int64_t n) {
  if (threadIdx.x == 0) table[n] = ctl->dt_last;
}
// the Flather values the step about to run uses (set by the previous finalize)
__global__ void k_ctl_save_ext(const Ctl *ctl, double *exts, int64_t n) {
  if (threadIdx.x < 8) exts[8 * n + threadIdx.x] = ctl->ext[threadIdx.x];
}
__global__ void k_gauge_sample(int ng, const int *cells, const double *h, const double *b, double *out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < ng) out[g] = b[cells[g]] + h[cells[g]];
}
// d/dh of sum_g (eta_g - obs_g)^2 at one sample; one lane, since gauges may share a cell
__global__ void k_gauge_inject(int ng, const int *cells, const double *eta, const double *obs, double *lh) {
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int g = 0; g < ng; ++g) lh[cells[g]] += 2.0 * (eta[g] - obs[g]);
}

// 256-thread blocks covering n, rounded up to a multiple of the XCD count
dim3 grid_for(size_t n) {
  const size_t nb = (n + 255) / 256;
  return dim3((unsigned)((nb + NXCD - 1) / NXCD * NXCD));
}

StateW st3(double *base, size_t N) { return {base, base + N, base + 2 * N}; }
StateP sp3(const double *base, size_t N) { return {base, base + N, base + 2 * N}; }
StateP sp(const StateW &s) { return {s.h, s.hu, s.hv}; }

// residual^T passes on the host: lS += (d residual(S) / d S)^T lR
void residual_rev_host(const Params &P, const StateP &S, const double *b, double dt, const StateP &lR, const Work &w,
                       const StateW &lS) {
  auto each = [&](auto &&f) {
    for (int j = 0; j < P.H; ++j)
      for (int i = 0; i < P.W; ++i) f(i, j);
  };
  each([&](int i, int j) { adj_cells_at(P, S, b, w, i, j); });
  each([&](int i, int j) { adj_faces_at(P, b, w, i, j); });
  each([&](int i, int j) { adj_theta_at(P, S, dt, w, i, j); });
  each([&](int i, int j) { adj_accum_at(P, S, dt, lR, w, i, j); });
  each([&](int i, int j) { adj_faces_rev_at(P, b, w, i, j); });
  each([&](int i, int j) { adj_cells_rev_at(P, S, b, lR, w, nullptr, i, j); });
  each([&](int i, int j) { adj_slopes_rev_at(P, S, w, lS, i, j); });
}

// Params over host rows for the host entry points (geometry rows in the order
// of swe_row; the reverse passes use no reciprocals)
int host_params(const swe_desc *d, const int32_t bc[4], Params &P) {
  if (!d || d->nx < 2 || d->ny < 1 || d->order < 1 || d->order > 2 || d->limiter < 0 || d->limiter > 2)
    return SWE_EINVAL;
  for (int r = 0; r < SWE_ROW_AREA; ++r)
    if (!d->rows[r]) return SWE_EINVAL;
  std::memset(&P, 0, sizeof(P));
  P.nx = d->nx; P.ny = d->ny; P.W = d->nx + 4; P.H = d->ny + 4;
  P.sphere = d->kind == SWE_SPHERE; P.order = d->order; P.limiter = d->limiter;
  for (int s = 0; s < 4; ++s) {
    if (bc && (bc[s] < SWE_BC_REFLECT || bc[s] > SWE_BC_FLATHER)) return SWE_EINVAL;
    P.bc[s] = bc ? bc[s] : BC_REFLECT;
  }
  P.g = d->g; P.half_g = 0.5 * d->g; P.neg_half_g = -d->g * 0.5; P.rg = 1.0 / d->g;
  P.min_depth = d->min_depth; P.eps4 = d->eps4; P.cfl = d->cfl; P.R = d->R;
  P.coriolis = d->coriolis;
  P.coriolis_on = (d->coriolis != 0.0 && d->coriolis == d->coriolis) ? 1 : 0;
  P.manning = d->manning;
  P.friction = (d->manning != 0.0 && d->manning == d->manning) ? 1 : 0;
  P.geo.dx = d->rows[SWE_ROW_DX]; P.geo.dyN = d->rows[SWE_ROW_DYN]; P.geo.dyS = d->rows[SWE_ROW_DYS];
  P.geo.dyCFL = d->rows[SWE_ROW_DYCFL]; P.geo.bedPhi = d->rows[SWE_ROW_BEDPHI];
  P.geo.tanPhi = d->rows[SWE_ROW_TANPHI]; P.geo.fRow = d->rows[SWE_ROW_F]; P.geo.geo = d->rows[SWE_ROW_GEO];
  return SWE_OK;
}

}  // namespace

struct swe_adj {
  swe_ctx *ctx = nullptr;
  int64_t nsteps = 0;
  int K = 1;                       // steps of one replayed segment
  int64_t KO = 1;                  // steps between the checkpoints kept for the whole run (a multiple of K)
  size_t N = 0;
  double *work = nullptr;          // the device reverse passes' own slots, DEV_SLOTS * N
  Work w{};
  double *ckpt = nullptr;          // ceil(nsteps / KO) states of 3N
  double *ickpt = nullptr;         // two levels: the KO / K segment start states of one outer segment
  double *seg[2] = {nullptr, nullptr};      // two segment stores, K steps x STEP_FIELDS N each
  double *tcs_seg[2] = {nullptr, nullptr};  // K steps x 2 stages x MAX_TIDE x (cos, sin), per store
  double *bed = nullptr;           // the bed with its ghost ring, read by the reverse passes
  hipStream_t rs = nullptr;        // the replay stream
  hipEvent_t replayed[2] = {nullptr, nullptr}, reversed[2] = {nullptr, nullptr};
  double *tmp = nullptr;           // stage scratch, T_COUNT states of 3N
  double *lN = nullptr;            // adjoint of the Manning field
  double *dt_dev = nullptr;        // recorded dt per step
  double *t_dev = nullptr;         // start time of every step
  double *ext_dev = nullptr;       // recorded Flather values of every step, 8 per step
  int *gcells = nullptr;
  double *gobs = nullptr, *geta = nullptr;
  int ng = 0, every = 1, nsamp = 0;
  int lam = 0, lap = 0;            // scratch slots holding the running adjoint and lA'
  bool recorded = false;
  // the adjoint slot lam still holds the ghost adjoints of the step after it
  // (BC^T folded on read by the next stage-2 epilogue, adj_bc_fold_into)
  bool pending = false;
  std::vector<double> h_dt;
  std::vector<double> h_eta;
  double loss = 0.0;
  double fwd_ms = 0.0, bwd_ms = 0.0;
  std::string err;
};

namespace {

#define ACHK(a, expr)                                                            \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      (a)->err = std::string(#expr) + ": " + hipGetErrorString(e_);              \
      return SWE_EHIP;                                                           \
    }                                                                            \
  } while (0)

int afail(swe_adj *a, int code, const std::string &m) {
  if (a) a->err = m;
  return code;
}

// scratch slots (T_A0..T_A2: the states of the inner-checkpoint forward)
enum { T_R, T_CUR, T_LBP, T_L0, T_L1, T_A0, T_A1, T_A2, T_COUNT };
// what one step of a segment keeps: the BC-applied stage inputs (B' holds stage
// 1's dryCleaned output in its interior) and stage 2's dryCleaned, pre-friction
// state (3 fields each), then per stage the cell fields, face flux terms and
// positivity factors its sweep computed (REC_FIELDS each, in the order of the
// Work slots S_U .. S_OUT)
enum { G_AP, G_BP, G_DC2, G_COUNT };
constexpr int STEP_FIELDS = 3 * G_COUNT + 2 * REC_FIELDS;
// Work slots the device passes keep in HBM: S_LXM .. S_LYVR
constexpr int DEV_SLOTS = S_LETA - REC_FIELDS;
static_assert(REC_FIELDS == S_OUT - S_U + 1 && S_U == 0, "recorded slots are Work slots S_U .. S_OUT");
StateW tslot(swe_adj *a, int q) { return st3(a->tmp + (size_t)q * 3 * a->N, a->N); }
StateW cur_state(swe_ctx *c) { return {c->st[c->A][0], c->st[c->A][1], c->st[c->A][2]}; }

int copy3_on(swe_adj *a, const StateW &dst, const StateP &src, hipStream_t s) {
  const size_t bytes = a->N * sizeof(double);
  ACHK(a, hipMemcpyAsync(dst.h, src.h, bytes, hipMemcpyDeviceToDevice, s));
  ACHK(a, hipMemcpyAsync(dst.hu, src.hu, bytes, hipMemcpyDeviceToDevice, s));
  ACHK(a, hipMemcpyAsync(dst.hv, src.hv, bytes, hipMemcpyDeviceToDevice, s));
  return SWE_OK;
}

int copy3(swe_adj *a, const StateW &dst, const StateP &src) {
  const size_t bytes = a->N * sizeof(double);
  hipStream_t s = a->ctx->stream;
  ACHK(a, hipMemcpyAsync(dst.h, src.h, bytes, hipMemcpyDeviceToDevice, s));
  ACHK(a, hipMemcpyAsync(dst.hu, src.hu, bytes, hipMemcpyDeviceToDevice, s));
  ACHK(a, hipMemcpyAsync(dst.hv, src.hv, bytes, hipMemcpyDeviceToDevice, s));
  return SWE_OK;
}

int enqueue(swe_adj *a) {
  const int rc = ctx_enqueue_step(a->ctx, 0, 0);
  if (rc) return afail(a, rc, std::string("forward step: ") + swe_last_error(a->ctx));
  return SWE_OK;
}

// forward step n at the recorded dt and Flather values (t accumulates on the
// device as in the recorded run)
int replay_step(swe_adj *a, int64_t n) {
  hipLaunchKernelGGL(k_ctl_dt, dim3(1), dim3(64), 0, a->ctx->stream, a->ctx->dctl, a->dt_dev, (const double *)nullptr,
                     a->ext_dev, n, (double *)nullptr);
  return enqueue(a);
}

// residual^T at stage input S.  rec: the cell fields, face flux terms and
// positivity factors the replayed sweep recorded for this stage (REC_FIELDS
// fields N apart), which stand in for the three forward recompute passes.
int residual_rev(swe_adj *a, const StateP &S, double dt, const StateP &lR, const StateW &lS, const double *tcs,
                 double *rec) {
  swe_ctx *c = a->ctx;
  const Params &P = c->P;
  hipStream_t st = c->stream;
  Work w = a->w;
  for (int q = 0; q < REC_FIELDS; ++q) w.s[S_U + q] = rec + (size_t)q * a->N;
  const size_t atiles = (size_t)((P.W + AF_X - 1) / AF_X) * ((P.H + AF_Y - 1) / AF_Y);
  hipLaunchKernelGGL(k_adj_accum_faces_tiled, dim3((unsigned)((atiles + NXCD - 1) / NXCD * NXCD)),
                     dim3(AF_X * AF_Y), 0, st, P, S, a->bed, dt, lR, w);
  const size_t tiles = (size_t)((P.W + CT_X - 1) / CT_X) * ((P.H + CT_Y - 1) / CT_Y);
  hipLaunchKernelGGL(k_adj_cells_slopes_tiled, dim3((unsigned)((tiles + NXCD - 1) / NXCD * NXCD)),
                     dim3(CT_X * CT_Y), 0, st, P, S, a->bed, lR, w, tcs, lS);
  ACHK(a, hipGetLastError());
  return SWE_OK;
}

// BC^T of step n's applyBC on state S (whose interior Flather sides read)
int bc_rev(swe_adj *a, const StateW &l, const StateP &S, int64_t n) {
  swe_ctx *c = a->ctx;
  const Params &P = c->P;
  const double *ext = a->ext_dev + 8 * n;
  if (adj_bc_gather_ok(P)) {
    hipLaunchKernelGGL(k_adj_bc_gather, grid_for(2 * (P.W - 4)), dim3(256), 0, c->stream, P, l, 0);
    hipLaunchKernelGGL(k_adj_bc_gather, grid_for(2 * P.ny), dim3(256), 0, c->stream, P, l, 1);
  } else {
    hipLaunchKernelGGL(k_adj_bc, grid_for(4 * P.W), dim3(256), 0, c->stream, P, l, 0, S, a->bed, ext);
    hipLaunchKernelGGL(k_adj_bc, grid_for(4 * P.ny), dim3(256), 0, c->stream, P, l, 1, S, a->bed, ext);
  }
  ACHK(a, hipGetLastError());
  return SWE_OK;
}

StateW gslot(swe_adj *a, int buf, int64_t q, int g) {
  return st3(a->seg[buf] + ((size_t)q * STEP_FIELDS + 3 * g) * a->N, a->N);
}
double *recslot(swe_adj *a, int buf, int64_t q, int stage) {
  return a->seg[buf] + ((size_t)q * STEP_FIELDS + 3 * G_COUNT + (size_t)(stage - 1) * REC_FIELDS) * a->N;
}
const double *tcs_of(swe_adj *a, int buf, int64_t q) { return a->tcs_seg[buf] + (size_t)q * 2 * MAX_TIDE * 2; }

// Re-run steps [n0, n1) from state src into segment store `buf` on
// the replay stream, with the forward's stage kernels at the recorded dt: step
// q = n - n0 reads A' = slot q (BC applied in place), writes B' = its stage-1
// output (BC applied in place), stage 2's dryCleaned pre-friction state, and its
// output straight into slot q + 1.  The replay touches only the store, the
// control block and the context's bed, never what the reverse sweep reads.
// The replay of a segment [n0, n1) into store buf on the replay stream, in two
// parts so that its steps can be issued between other work: the start state,
// then step n (in order, n0 first).
int segment_begin(swe_adj *a, const double *src, int buf) {
  return copy3_on(a, gslot(a, buf, 0, G_AP), sp3(src, a->N), a->rs);
}
int segment_step(swe_adj *a, int64_t n0, int64_t n, int64_t n1, int buf) {
  swe_ctx *c = a->ctx;
  const Params &P = c->P;
  hipStream_t st = a->rs;
  const int np = sweep_waves(P, c->chunk);
  const int64_t q = n - n0;
  const StateW Ap = gslot(a, buf, q, G_AP), Bp = gslot(a, buf, q, G_BP), dc = gslot(a, buf, q, G_DC2);
  const StateW next = n + 1 < n1 ? gslot(a, buf, q + 1, G_AP) : tslot(a, T_CUR);
  hipLaunchKernelGGL(k_ctl_dt, dim3(1), dim3(64), 0, st, c->dctl, a->dt_dev, a->t_dev, a->ext_dev, n,
                     a->tcs_seg[buf] + (size_t)q * 2 * MAX_TIDE * 2);
  ACHK(a, launch_apply_bc(P, Ap, c->b, c->dctl, st));
  SweepArgs s1 = ctx_base_args(c);
  s1.in = sp(Ap);
  s1.out = Bp;
  s1.recf = recslot(a, buf, q, 1);
  ACHK(a, launch_sweep(MODE_STAGE1, s1, st));
  ACHK(a, launch_apply_bc(P, Bp, c->b, c->dctl, st));
  SweepArgs s2 = ctx_base_args(c);
  s2.in = sp(Bp);
  s2.h0 = sp(Ap);
  s2.out = next;
  s2.rec = dc.h;
  s2.recf = recslot(a, buf, q, 2);
  s2.mf_part = c->mf_part + np;
  ACHK(a, launch_sweep(MODE_STAGE2, s2, st));
  return SWE_OK;
}
int replay_segment(swe_adj *a, const double *src, int64_t n0, int64_t n1, int buf) {
  if (int rc = segment_begin(a, src, buf)) return rc;
  for (int64_t n = n0; n < n1; ++n)
    if (int rc = segment_step(a, n0, n, n1, buf)) return rc;
  return SWE_OK;
}

// Two levels: the start states of the K-step segments of outer segment so, from
// its checkpoint, by the stage kernels at the recorded dt on the replay stream
// (the states of the forward run bit for bit).
int inner_checkpoints(swe_adj *a, int64_t so) {
  swe_ctx *c = a->ctx;
  const Params &P = c->P;
  hipStream_t st = a->rs;
  const int np = sweep_waves(P, c->chunk);
  const size_t N3 = 3 * a->N;
  const int64_t o0 = so * a->KO, o1 = std::min<int64_t>(a->nsteps, o0 + a->KO);
  const int64_t nin = (o1 - o0 + a->K - 1) / a->K;
  int rc;
  if ((rc = copy3_on(a, st3(a->ickpt, a->N), sp3(a->ckpt + (size_t)so * N3, a->N), st))) return rc;
  if ((rc = copy3_on(a, tslot(a, T_A0), sp3(a->ckpt + (size_t)so * N3, a->N), st))) return rc;
  int cur = 0;
  for (int64_t m = 1; m < nin; ++m) {
    for (int64_t n = o0 + (m - 1) * a->K; n < o0 + m * a->K; ++n) {
      const StateW X = tslot(a, T_A0 + cur), Y = tslot(a, T_A0 + (cur + 1) % 3), Z = tslot(a, T_A0 + (cur + 2) % 3);
      hipLaunchKernelGGL(k_ctl_dt, dim3(1), dim3(64), 0, st, c->dctl, a->dt_dev, a->t_dev, a->ext_dev, n,
                         (double *)nullptr);
      ACHK(a, launch_apply_bc(P, X, c->b, c->dctl, st));
      SweepArgs s1 = ctx_base_args(c);
      s1.in = sp(X);
      s1.out = Y;
      ACHK(a, launch_sweep(MODE_STAGE1, s1, st));
      ACHK(a, launch_apply_bc(P, Y, c->b, c->dctl, st));
      SweepArgs s2 = ctx_base_args(c);
      s2.in = sp(Y);
      s2.h0 = sp(X);
      s2.out = Z;
      s2.mf_part = c->mf_part + np;
      ACHK(a, launch_sweep(MODE_STAGE2, s2, st));
      cur = (cur + 2) % 3;
    }
    if ((rc = copy3_on(a, st3(a->ickpt + (size_t)m * N3, a->N), sp(tslot(a, T_A0 + cur)), st))) return rc;
  }
  return SWE_OK;
}

// Step n back from what replay_segment kept in slot q of store buf: maps the
// adjoint of the step's output (slot a->lam, interior) to the adjoint of its
// start state.
int step_rev(swe_adj *a, int buf, int64_t q, int64_t n) {
  swe_ctx *c = a->ctx;
  const Params &P = c->P;
  const double dt = a->h_dt[n];
  hipStream_t st = c->stream;
  const dim3 g = grid_for(a->N), blk(256);
  const StateW Ap = gslot(a, buf, q, G_AP), Bp = gslot(a, buf, q, G_BP), dc = gslot(a, buf, q, G_DC2);
  const StateW lR = tslot(a, T_R), lBp = tslot(a, T_LBP), lam = tslot(a, a->lam), lAp = tslot(a, a->lap);
  const StateW none = {nullptr, nullptr, nullptr};
  const double *tcs = tcs_of(a, buf, q);
  // BC^T of a stage runs as its own pass, or (the configurations of the gather
  // form) is folded on read by the epilogue that next reads that adjoint
  const bool defer = adj_bc_gather_ok(P);
  int rc;
  // stage 2: friction^T, dryClean^T, C0 = 0.5 (A' + B' + dt R2)^T, residual^T, BC^T
  hipLaunchKernelGGL(k_adj_epilogue, g, blk, 0, st, P, sp(dc), sp(lam), c->mfield, c->sponge, dt, a->lN, 2, lAp,
                     lBp, lR, a->pending ? 1 : 0);
  if ((rc = residual_rev(a, sp(Bp), dt, sp(lR), lBp, tcs + 2 * MAX_TIDE, recslot(a, buf, q, 2)))) return rc;
  if (!defer && (rc = bc_rev(a, lBp, sp(Bp), n))) return rc;
  // stage 1: dryClean^T (B's interior is stage 1's dryCleaned output), B0 = A' + dt R1 ^T,
  // residual^T, BC^T
  hipLaunchKernelGGL(k_adj_epilogue, g, blk, 0, st, P, sp(Bp), sp(lBp), c->mfield, c->sponge, dt, a->lN, 1, lAp,
                     none, lR, defer ? 1 : 0);
  if ((rc = residual_rev(a, sp(Ap), dt, sp(lR), lAp, tcs, recslot(a, buf, q, 1)))) return rc;
  if (!defer && (rc = bc_rev(a, lAp, sp(Ap), n))) return rc;
  a->pending = defer;
  std::swap(a->lam, a->lap);
  return SWE_OK;
}

// Progress lines for long runs: with SWE_ADJ_PROGRESS set in the environment the
// forward and the backward report on stderr every twentieth of their work,
// after waiting for the device to get there (so a report means work done).
int64_t progress_stride(int64_t total) {
  static const bool on = std::getenv("SWE_ADJ_PROGRESS") != nullptr;
  return on ? std::max<int64_t>(1, total / 20) : 0;
}

void progress(hipStream_t st, const char *what, int64_t done, int64_t total) {
  (void)hipStreamSynchronize(st);
  std::fprintf(stderr, "[swe_adj] %s: %lld of %lld\n", what, (long long)done, (long long)total);
  std::fflush(stderr);
}

int set_initial(swe_adj *a, const double *h0, const double *nfield) {
  swe_ctx *c = a->ctx;
  std::vector<double> z(a->N, 0.0);
  int rc = swe_set_field(c, SWE_F_H, h0);
  if (!rc) rc = swe_set_field(c, SWE_F_HU, z.data());
  if (!rc) rc = swe_set_field(c, SWE_F_HV, z.data());
  if (!rc && nfield) rc = swe_set_field(c, SWE_F_MANNING, nfield);
  if (rc) return afail(a, rc, std::string("initial state: ") + swe_last_error(c));
  hipLaunchKernelGGL(k_ctl_self, dim3(1), dim3(64), 0, c->stream, c->dctl, c->harm_enabled);
  ACHK(a, hipGetLastError());
  return SWE_OK;
}

int sample(swe_adj *a, int64_t s) {
  swe_ctx *c = a->ctx;
  if (!a->ng) return SWE_OK;
  hipLaunchKernelGGL(k_gauge_sample, grid_for(a->ng), dim3(256), 0, c->stream, a->ng, a->gcells,
                     c->st[c->A][0], c->b, a->geta + (size_t)s * a->ng);
  ACHK(a, hipGetLastError());
  return SWE_OK;
}

// forward run from the initial state: self-timed (recording dt and keeping
// checkpoints) or replaying the recorded dt sequence
int forward(swe_adj *a, bool record, double *loss) {
  swe_ctx *c = a->ctx;
  const size_t N3 = 3 * a->N;
  int rc;
  if (record) {
    // the first dt from the initial state (maxDt, swe.mjs:632-647)
    ACHK(a, launch_maxdt(c->P, sp(cur_state(c)), c->dctl, c->stream));
    ACHK(a, launch_finalize(c->dctl, c->mf_part, 0, c->P.cfl, 0, c->stream));
  }
  const int64_t pstride = progress_stride(a->nsteps);
  for (int64_t n = 0; n < a->nsteps; ++n) {
    if (pstride && n > 0 && n % pstride == 0) progress(c->stream, record ? "forward steps" : "replayed steps", n, a->nsteps);
    if (record && n % a->KO == 0)
      if ((rc = copy3(a, st3(a->ckpt + (size_t)(n / a->KO) * N3, a->N), sp(cur_state(c))))) return rc;
    if (record) {
      hipLaunchKernelGGL(k_ctl_save_ext, dim3(1), dim3(64), 0, c->stream, c->dctl, a->ext_dev, n);
      if ((rc = enqueue(a))) return rc;
      hipLaunchKernelGGL(k_ctl_record, dim3(1), dim3(64), 0, c->stream, c->dctl, a->dt_dev, n);
    } else if ((rc = replay_step(a, n))) {
      return rc;
    }
    if ((n + 1) % a->every == 0 && (n + 1) / a->every <= a->nsamp)
      if ((rc = sample(a, (n + 1) / a->every - 1))) return rc;
  }
  ACHK(a, hipStreamSynchronize(c->stream));
  if (record) {
    a->h_dt.assign((size_t)a->nsteps, 0.0);
    ACHK(a, hipMemcpy(a->h_dt.data(), a->dt_dev, sizeof(double) * a->nsteps, hipMemcpyDeviceToHost));
    for (int64_t n = 0; n < a->nsteps; ++n)
      if (!(a->h_dt[n] > 0) || std::isinf(a->h_dt[n]))
        return afail(a, SWE_ESTATE, "forward step " + std::to_string(n) + " was refused (no wet cell)");
    // step start times, accumulated as the control block accumulates them
    std::vector<double> ts(a->nsteps);
    double t = 0.0;
    for (int64_t n = 0; n < a->nsteps; ++n) { ts[n] = t; t = t + a->h_dt[n]; }
    ACHK(a, hipMemcpy(a->t_dev, ts.data(), sizeof(double) * a->nsteps, hipMemcpyHostToDevice));
    a->recorded = true;
  }
  const size_t no = (size_t)a->ng * a->nsamp;
  a->h_eta.assign(no, 0.0);
  std::vector<double> obs(no);
  if (no) {
    ACHK(a, hipMemcpy(a->h_eta.data(), a->geta, sizeof(double) * no, hipMemcpyDeviceToHost));
    ACHK(a, hipMemcpy(obs.data(), a->gobs, sizeof(double) * no, hipMemcpyDeviceToHost));
  }
  double J = 0.0;
  for (size_t q = 0; q < no; ++q) {
    const double r = a->h_eta[q] - obs[q];
    J += r * r;
  }
  a->loss = J;
  if (loss) *loss = J;
  return SWE_OK;
}

int inject(swe_adj *a, int64_t n) {   // gauge terms of the state after n steps
  if (!a->ng || n <= 0 || n % a->every != 0 || n / a->every > a->nsamp) return SWE_OK;
  const size_t s = (size_t)(n / a->every - 1) * a->ng;
  hipLaunchKernelGGL(k_gauge_inject, dim3(1), dim3(64), 0, a->ctx->stream, a->ng, a->gcells, a->geta + s,
                     a->gobs + s, tslot(a, a->lam).h);
  ACHK(a, hipGetLastError());
  return SWE_OK;
}

struct Timer {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~Timer() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

}  // namespace

extern "C" {

const char *swe_adj_last_error(const swe_adj *a) { return a ? a->err.c_str() : "null adjoint"; }

void swe_adj_destroy(swe_adj *a) {
  if (!a) return;
  // touches nothing of the context, which may already be gone; hipFree waits
  // for the device work that uses the buffers
  if (a->rs) (void)hipStreamSynchronize(a->rs);
  (void)hipFree(a->work); (void)hipFree(a->ckpt); (void)hipFree(a->ickpt); (void)hipFree(a->tmp);
  for (int q = 0; q < 2; ++q) {
    (void)hipFree(a->seg[q]); (void)hipFree(a->tcs_seg[q]);
    if (a->replayed[q]) (void)hipEventDestroy(a->replayed[q]);
    if (a->reversed[q]) (void)hipEventDestroy(a->reversed[q]);
  }
  (void)hipFree(a->bed);
  if (a->rs) (void)hipStreamDestroy(a->rs);
  (void)hipFree(a->lN); (void)hipFree(a->dt_dev); (void)hipFree(a->t_dev); (void)hipFree(a->ext_dev);
  (void)hipFree(a->gcells); (void)hipFree(a->gobs); (void)hipFree(a->geta);
  delete a;
}

int swe_adj_plan(int64_t nsteps, int64_t ncells, int64_t inner, int64_t outer, double *bytes) {
  if (nsteps < 1 || ncells < 1 || inner < 1 || outer < inner || outer % inner || !bytes) return SWE_EINVAL;
  const double nck = (double)((nsteps + outer - 1) / outer);
  const double nin = outer > inner ? (double)((outer + inner - 1) / inner) : 0.0;
  const double fields = 3.0 * (nck + nin) + 2.0 * (double)inner * STEP_FIELDS + DEV_SLOTS + 3.0 * T_COUNT + 2.0;
  *bytes = 8.0 * (double)ncells * fields + 16.0 * (double)nsteps + 2.0 * 8.0 * (double)inner * 2 * MAX_TIDE * 2;
  return SWE_OK;
}

int swe_adj_create(swe_ctx *c, int64_t nsteps, int ckpt, swe_adj **out) {
  if (!c || !out || nsteps < 1 || ckpt < 0) return ctx_fail(c, SWE_EINVAL, "bad adjoint arguments");
  if (ckpt > 0) return swe_adj_create_levels(c, nsteps, ckpt, ckpt, out);
  // sqrt(nsteps) checkpointing when it fits in 90% of the free device memory,
  // else two levels with segments of up to cbrt(nsteps) steps and checkpoints
  // every segment length squared
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return ctx_fail(c, SWE_EHIP, "hipMemGetInfo failed");
  const int64_t k1 = std::max<int64_t>(1, std::llround(std::sqrt((double)nsteps)));
  double bytes = 0.0;
  swe_adj_plan(nsteps, (int64_t)c->N, k1, k1, &bytes);
  if (bytes <= 0.9 * (double)free_b) return swe_adj_create_levels(c, nsteps, k1, k1, out);
  // two levels: the largest inner interval up to cbrt(nsteps) whose plan fits
  int64_t k2 = std::max<int64_t>(1, std::llround(std::cbrt((double)nsteps)));
  auto outer_of = [&](int64_t k) { return std::max<int64_t>(k, std::min<int64_t>(nsteps, k * k) / k * k); };
  while (k2 > 1) {
    swe_adj_plan(nsteps, (int64_t)c->N, k2, outer_of(k2), &bytes);
    if (bytes <= 0.9 * (double)free_b) break;
    k2 = k2 * 3 / 4;
  }
  return swe_adj_create_levels(c, nsteps, k2, outer_of(k2), out);
}

int swe_adj_create_levels(swe_ctx *c, int64_t nsteps, int64_t inner, int64_t outer, swe_adj **out) {
  if (!c || !out || nsteps < 1 || inner < 1 || inner > (1 << 30) || outer < inner || outer % inner)
    return ctx_fail(c, SWE_EINVAL, "bad adjoint arguments (outer must be a multiple of inner)");
  *out = nullptr;
  if (c->mid) return ctx_fail(c, SWE_ESTATE, "a step is in progress (swe_step_begin without swe_step_end)");
  for (int s = 0; s < 4; ++s)
    if (c->P.bc[s] == BC_STORED)
      return ctx_fail(c, SWE_EINVAL, "the adjoint needs device boundary kinds (reflect, outflow, periodic, flather)");
  swe_adj *a = new swe_adj();
  a->ctx = c;
  a->nsteps = nsteps;
  a->K = (int)inner;
  a->KO = outer;
  a->N = c->N;
  a->lam = T_L0;
  a->lap = T_L1;
  const size_t N = a->N, N3 = 3 * N;
  const size_t nck = (size_t)((nsteps + a->KO - 1) / a->KO);
  auto bail = [&](const char *what, hipError_t e) {
    const std::string m = std::string(what) + ": " + hipGetErrorString(e);
    swe_adj_destroy(a);
    return ctx_fail(c, e == hipErrorOutOfMemory ? SWE_ENOMEM : SWE_EHIP, m);
  };
  hipError_t e;
  // slots S_U .. S_OUT are the recorded stage fields (pointed into the segment
  // store per stage) and the fused gather + limiter^T pass keeps S_LETA .. in
  // registers: only the slots in between are allocated here
  const size_t nslots = DEV_SLOTS;
  if ((e = hipMalloc(&a->work, sizeof(double) * nslots * N)) != hipSuccess) return bail("adjoint passes", e);
  for (int q = 0; q < S_COUNT; ++q)
    a->w.s[q] = (q < REC_FIELDS || q >= S_LETA) ? nullptr : a->work + (size_t)(q - REC_FIELDS) * N;
  if ((e = hipMalloc(&a->ckpt, sizeof(double) * nck * N3)) != hipSuccess) return bail("checkpoints", e);
  if (a->KO > a->K) {
    const size_t nin = (size_t)((a->KO + a->K - 1) / a->K);
    if ((e = hipMalloc(&a->ickpt, sizeof(double) * nin * N3)) != hipSuccess) return bail("inner checkpoints", e);
  }
  for (int q = 0; q < 2; ++q) {
    if ((e = hipMalloc(&a->seg[q], sizeof(double) * (size_t)a->K * STEP_FIELDS * N)) != hipSuccess)
      return bail("segment store", e);
    if ((e = hipMalloc(&a->tcs_seg[q], sizeof(double) * (size_t)a->K * 2 * MAX_TIDE * 2)) != hipSuccess)
      return bail("tide factors", e);
    if ((e = hipEventCreateWithFlags(&a->replayed[q], hipEventDisableTiming)) != hipSuccess) return bail("event", e);
    if ((e = hipEventCreateWithFlags(&a->reversed[q], hipEventDisableTiming)) != hipSuccess) return bail("event", e);
  }
  if ((e = hipMalloc(&a->bed, sizeof(double) * N)) != hipSuccess) return bail("bed copy", e);
  if ((e = hipStreamCreateWithFlags(&a->rs, hipStreamNonBlocking)) != hipSuccess) return bail("replay stream", e);
  if ((e = hipMalloc(&a->tmp, sizeof(double) * T_COUNT * N3)) != hipSuccess) return bail("stage scratch", e);
  if ((e = hipMalloc(&a->lN, sizeof(double) * N)) != hipSuccess) return bail("Manning adjoint", e);
  if ((e = hipMalloc(&a->dt_dev, sizeof(double) * (size_t)nsteps)) != hipSuccess) return bail("dt record", e);
  if ((e = hipMalloc(&a->t_dev, sizeof(double) * (size_t)nsteps)) != hipSuccess) return bail("t record", e);
  if ((e = hipMalloc(&a->ext_dev, sizeof(double) * 8 * (size_t)nsteps)) != hipSuccess) return bail("ext record", e);
  hipLaunchKernelGGL(k_fill, grid_for(nslots * N), dim3(256), 0, c->stream, nslots * N, a->work, 0.0);
  hipLaunchKernelGGL(k_fill, grid_for(T_COUNT * N3), dim3(256), 0, c->stream, T_COUNT * N3, a->tmp, 0.0);
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail("clear", e);
  // the differentiated forward always reads a per-cell Manning field
  if (!c->mfield) {
    std::vector<double> nf(N, c->manning == c->manning ? c->manning : 0.0);
    const int rc = swe_set_field(c, SWE_F_MANNING, nf.data());
    if (rc) {
      swe_adj_destroy(a);
      return rc;
    }
  }
  *out = a;
  return SWE_OK;
}

int swe_adj_set_gauges(swe_adj *a, int ng, const int32_t *cells, int every, const double *obs) {
  if (!a || ng < 0 || every < 1 || (ng > 0 && (!cells || !obs))) return afail(a, SWE_EINVAL, "bad gauges");
  const Params &P = a->ctx->P;
  for (int q = 0; q < ng; ++q) {
    const int k = cells[q];
    if (k < 0 || (size_t)k >= a->N || k % P.W < 2 || k % P.W > P.W - 3 || k / P.W < 2 || k / P.W > P.H - 3)
      return afail(a, SWE_EINVAL, "gauge " + std::to_string(q) + " is not an interior cell");
  }
  (void)hipFree(a->gcells); (void)hipFree(a->gobs); (void)hipFree(a->geta);
  a->gcells = nullptr; a->gobs = nullptr; a->geta = nullptr;
  a->ng = ng;
  a->every = every;
  a->nsamp = (int)(a->nsteps / every);
  const size_t no = (size_t)ng * a->nsamp;
  ACHK(a, hipMalloc(&a->gcells, sizeof(int) * std::max(1, ng)));
  ACHK(a, hipMalloc(&a->gobs, sizeof(double) * std::max<size_t>(1, no)));
  ACHK(a, hipMalloc(&a->geta, sizeof(double) * std::max<size_t>(1, no)));
  if (ng) ACHK(a, hipMemcpy(a->gcells, cells, sizeof(int) * ng, hipMemcpyHostToDevice));
  if (no) ACHK(a, hipMemcpy(a->gobs, obs, sizeof(double) * no, hipMemcpyHostToDevice));
  return SWE_OK;
}

int swe_adj_forward(swe_adj *a, const double *h0, const double *nfield, double *loss) {
  if (!a || !h0) return afail(a, SWE_EINVAL, "null argument");
  int rc = set_initial(a, h0, nfield);
  if (rc) return rc;
  Timer t;
  ACHK(a, hipEventCreate(&t.e0));
  ACHK(a, hipEventCreate(&t.e1));
  ACHK(a, hipEventRecord(t.e0, a->ctx->stream));
  a->recorded = false;
  if ((rc = forward(a, true, loss))) return rc;
  ACHK(a, hipEventRecord(t.e1, a->ctx->stream));
  ACHK(a, hipEventSynchronize(t.e1));
  float ms = 0.f;
  ACHK(a, hipEventElapsedTime(&ms, t.e0, t.e1));
  a->fwd_ms = ms;
  return SWE_OK;
}

int swe_adj_loss(swe_adj *a, const double *h0, const double *nfield, double *loss) {
  if (!a || !h0) return afail(a, SWE_EINVAL, "null argument");
  if (!a->recorded) return afail(a, SWE_ESTATE, "swe_adj_forward has not run");
  const int rc = set_initial(a, h0, nfield);
  if (rc) return rc;
  return forward(a, false, loss);
}

int swe_adj_get_series(swe_adj *a, double *eta) {
  if (!a || !eta) return afail(a, SWE_EINVAL, "null argument");
  std::copy(a->h_eta.begin(), a->h_eta.end(), eta);
  return SWE_OK;
}

int swe_adj_get_dt(swe_adj *a, double *dt) {
  if (!a || !dt) return afail(a, SWE_EINVAL, "null argument");
  if (!a->recorded) return afail(a, SWE_ESTATE, "swe_adj_forward has not run");
  std::copy(a->h_dt.begin(), a->h_dt.end(), dt);
  return SWE_OK;
}

int swe_adj_backward(swe_adj *a, double *grad_h0, double *grad_n) {
  if (!a) return SWE_EINVAL;
  if (!a->recorded) return afail(a, SWE_ESTATE, "swe_adj_forward has not run");
  swe_ctx *c = a->ctx;
  const size_t N = a->N, N3 = 3 * N;
  hipStream_t st = c->stream;
  Timer t;
  ACHK(a, hipEventCreate(&t.e0));
  ACHK(a, hipEventCreate(&t.e1));
  ACHK(a, hipEventRecord(t.e0, st));
  int rc;
  hipLaunchKernelGGL(k_fill, grid_for(N3), dim3(256), 0, st, N3, tslot(a, a->lam).h, 0.0);
  hipLaunchKernelGGL(k_fill, grid_for(N), dim3(256), 0, st, N, a->lN, 0.0);
  a->pending = false;
  if ((rc = inject(a, a->nsteps))) return rc;
  // the reverse passes read a private copy of the bed (ghost ring as the forward
  // leaves it); the replay stream keeps writing the context's own
  ACHK(a, hipMemcpyAsync(a->bed, c->b, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
  // Segments of K steps are reversed last to first.  While the reverse passes of
  // one segment (HBM-bound gathers, main stream) run, the replay of the segment
  // before it (VALU-bound stage sweeps, replay stream) fills the other store; with
  // two levels the replay stream first recomputes the start states of an outer
  // segment's K-step segments, when it reaches that outer segment.
  struct Seg { int64_t so, m, n0, n1; };
  std::vector<Seg> segs;
  const bool two = a->KO > a->K;
  const int64_t nout = (a->nsteps + a->KO - 1) / a->KO;
  for (int64_t so = nout - 1; so >= 0; --so) {
    const int64_t o0 = so * a->KO, o1 = std::min<int64_t>(a->nsteps, o0 + a->KO);
    for (int64_t m = (o1 - o0 + a->K - 1) / a->K - 1; m >= 0; --m)
      segs.push_back({so, m, o0 + m * a->K, std::min<int64_t>(o1, o0 + (m + 1) * a->K)});
  }
  // a replay that starts an outer segment first recomputes its inner checkpoints
  auto new_outer = [&](size_t i) { return two && (i == 0 || segs[i - 1].so != segs[i].so); };
  auto src_of = [&](size_t i) {
    return two ? a->ickpt + (size_t)segs[i].m * N3 : a->ckpt + (size_t)segs[i].so * N3;
  };
  auto issue_replay = [&](size_t i) -> int {
    const Seg &g = segs[i];
    if (new_outer(i))
      if (int r = inner_checkpoints(a, g.so)) return r;
    return replay_segment(a, src_of(i), g.n0, g.n1, (int)(i % 2));
  };
  for (int q = 0; q < 2; ++q) ACHK(a, hipEventRecord(a->reversed[q], st));
  ACHK(a, hipStreamWaitEvent(a->rs, a->reversed[0], 0));
  if ((rc = issue_replay(0))) return rc;
  ACHK(a, hipEventRecord(a->replayed[0], a->rs));
  const int64_t pstride = progress_stride((int64_t)segs.size());
  for (size_t i = 0; i < segs.size(); ++i) {
    if (pstride && i > 0 && (int64_t)i % pstride == 0) progress(st, "backward segments", (int64_t)i, (int64_t)segs.size());
    const int buf = (int)(i % 2);
    ACHK(a, hipStreamWaitEvent(st, a->replayed[buf], 0));
    // The replay of the next segment into the other store runs beside this
    // segment's reverse passes, once the segment before this one (which read
    // that store) is reversed.  Its steps are issued one after each reverse
    // step, so that it starts with this segment's first reverse pass and
    // neither stream waits while the host issues the other's launches.  One
    // that first recomputes inner checkpoints (many steps) is issued whole,
    // after this segment's reverse passes.
    const bool next = i + 1 < segs.size();
    const bool split = next && !new_outer(i + 1);
    int64_t rn = next ? segs[i + 1].n0 : 0;   // the next segment's first step not yet issued
    if (split) {
      ACHK(a, hipStreamWaitEvent(a->rs, a->reversed[1 - buf], 0));
      if ((rc = segment_begin(a, src_of(i + 1), 1 - buf))) return rc;
    }
    for (int64_t n = segs[i].n1 - 1; n >= segs[i].n0; --n) {
      if ((rc = step_rev(a, buf, n - segs[i].n0, n))) return rc;
      if ((rc = inject(a, n))) return rc;
      if (split && rn < segs[i + 1].n1) {
        if ((rc = segment_step(a, segs[i + 1].n0, rn, segs[i + 1].n1, 1 - buf))) return rc;
        ++rn;
      }
    }
    ACHK(a, hipEventRecord(a->reversed[buf], st));
    if (split) {
      for (; rn < segs[i + 1].n1; ++rn)
        if ((rc = segment_step(a, segs[i + 1].n0, rn, segs[i + 1].n1, 1 - buf))) return rc;
      ACHK(a, hipEventRecord(a->replayed[1 - buf], a->rs));
    } else if (next) {
      ACHK(a, hipStreamWaitEvent(a->rs, a->reversed[1 - buf], 0));
      if ((rc = issue_replay(i + 1))) return rc;
      ACHK(a, hipEventRecord(a->replayed[1 - buf], a->rs));
    }
  }
  // the first step's BC^T, when folded on read, has no epilogue after it
  if (a->pending) {
    const StateW lam = tslot(a, a->lam);
    if ((rc = bc_rev(a, lam, sp(lam), 0))) return rc;
    a->pending = false;
  }
  ACHK(a, hipEventRecord(t.e1, st));
  ACHK(a, hipEventSynchronize(t.e1));
  float ms = 0.f;
  ACHK(a, hipEventElapsedTime(&ms, t.e0, t.e1));
  a->bwd_ms = ms;
  if (grad_h0) ACHK(a, hipMemcpy(grad_h0, tslot(a, a->lam).h, sizeof(double) * N, hipMemcpyDeviceToHost));
  if (grad_n) ACHK(a, hipMemcpy(grad_n, a->lN, sizeof(double) * N, hipMemcpyDeviceToHost));
  return SWE_OK;
}

int swe_adj_timing(swe_adj *a, double *fwd_ms, double *bwd_ms) {
  if (!a) return SWE_EINVAL;
  if (fwd_ms) *fwd_ms = a->fwd_ms;
  if (bwd_ms) *bwd_ms = a->bwd_ms;
  return SWE_OK;
}

// ---- host entry points: the same passes on host arrays ------------------------
int swe_adj_host_residual_rev(const swe_desc *d, const double *h, const double *hu, const double *hv,
                              const double *b, double dt, const double *lrh, const double *lrhu,
                              const double *lrhv, double *lh, double *lhu, double *lhv) {
  Params P;
  if (host_params(d, nullptr, P) || !h || !hu || !hv || !b || !lrh || !lrhu || !lrhv || !lh || !lhu || !lhv)
    return SWE_EINVAL;
  const size_t N = (size_t)P.W * P.H;
  std::vector<double> work(S_COUNT * N, 0.0);
  Work w;
  for (int q = 0; q < S_COUNT; ++q) w.s[q] = work.data() + (size_t)q * N;
  residual_rev_host(P, {h, hu, hv}, b, dt, {lrh, lrhu, lrhv}, w, {lh, lhu, lhv});
  return SWE_OK;
}

int swe_adj_host_bc_rev(const swe_desc *d, const int32_t bc[4], const double ext[8], const double *h,
                        const double *hu, const double *hv, const double *b, double *lh, double *lhu, double *lhv) {
  Params P;
  if (!bc || host_params(d, bc, P) || !lh || !lhu || !lhv) return SWE_EINVAL;
  bool flather = false;
  for (int s = 0; s < 4; ++s) flather = flather || bc[s] == SWE_BC_FLATHER;
  if (flather && (!ext || !h || !hu || !hv || !b)) return SWE_EINVAL;
  const StateW l = {lh, lhu, lhv};
  const StateP S = {h, hu, hv};
  for (int t = 0; t < 4 * P.W; ++t) adj_bc_ghost(P, l, 0, t, S, b, ext);
  for (int t = 0; t < 4 * P.ny; ++t) adj_bc_ghost(P, l, 1, t, S, b, ext);
  return SWE_OK;
}

// The gather form of the same transpose (adj_bc_gather_ok configurations only),
// for checking it against the loop above on a machine without a GPU.
int swe_adj_host_bc_rev_gather(const swe_desc *d, const int32_t bc[4], const double ext[8], const double *h,
                               const double *hu, const double *hv, const double *b, double *lh, double *lhu,
                               double *lhv) {
  Params P;
  if (!bc || host_params(d, bc, P) || !lh || !lhu || !lhv) return SWE_EINVAL;
  if (!adj_bc_gather_ok(P)) return SWE_EINVAL;
  bool flather = false;
  for (int s = 0; s < 4; ++s) flather = flather || bc[s] == SWE_BC_FLATHER;
  if (flather && (!ext || !h || !hu || !hv || !b)) return SWE_EINVAL;
  const StateW l = {lh, lhu, lhv};
  const StateP S = {h, hu, hv};
  for (int q = 0; q < 2 * (P.W - 4); ++q) adj_bc_gather(P, l, 0, q);
  for (int q = 0; q < 2 * P.ny; ++q) adj_bc_gather(P, l, 1, q);
  return SWE_OK;
}

// The same transpose folded on read (adj_bc_fold_into), as the device's
// epilogues apply it: every interior cell's folds from a snapshot of the ghosts,
// the ghosts then cleared, for checking it against the loop above.
int swe_adj_host_bc_rev_onread(const swe_desc *d, const int32_t bc[4], const double ext[8], const double *h,
                               const double *hu, const double *hv, const double *b, double *lh, double *lhu,
                               double *lhv) {
  (void)ext; (void)h; (void)hu; (void)hv; (void)b;
  Params P;
  if (!bc || host_params(d, bc, P) || !lh || !lhu || !lhv) return SWE_EINVAL;
  if (!adj_bc_gather_ok(P)) return SWE_EINVAL;
  const size_t N = (size_t)P.W * P.H;
  const std::vector<double> sh(lh, lh + N), shu(lhu, lhu + N), shv(lhv, lhv + N);
  const StateP l = {sh.data(), shu.data(), shv.data()};
  for (int j = 0; j < P.H; ++j)
    for (int i = 0; i < P.W; ++i) {
      const size_t k = (size_t)j * P.W + i;
      if (i >= 2 && i <= P.W - 3 && j >= 2 && j <= P.H - 3) {
        adj_bc_fold_into(P, l, i, j, lh[k], lhu[k], lhv[k]);
      } else {
        lh[k] = 0.0; lhu[k] = 0.0; lhv[k] = 0.0;
      }
    }
  return SWE_OK;
}

int swe_adj_host_epilogue_rev(const swe_desc *d, int stage, const double *pre_h, const double *pre_hu,
                              const double *pre_hv, const double *nfield, const double *sponge, double dt,
                              const double *lout_h,
                              const double *lout_hu, const double *lout_hv, double *lpre_h, double *lpre_hu,
                              double *lpre_hv, double *lN) {
  Params P;
  if ((stage != 1 && stage != 2) || host_params(d, nullptr, P) || !pre_h || !pre_hu || !pre_hv || !lout_h ||
      !lout_hu || !lout_hv || !lpre_h || !lpre_hu || !lpre_hv)
    return SWE_EINVAL;
  if (nfield) P.friction = 2;
  P.sponge = sponge ? 1 : 0;
  const size_t N = (size_t)P.W * P.H;
  // dryClean (swe.mjs:734-745) of the given state on the interior
  std::vector<double> dc(3 * N);
  for (size_t k = 0; k < N; ++k) {
    double h = pre_h[k], hu = pre_hu[k], hv = pre_hv[k];
    const int i = (int)(k % P.W), j = (int)(k / P.W);
    if (i >= 2 && i <= P.W - 3 && j >= 2 && j <= P.H - 3) {
      if (h < 0) h = 0.0;
      if (h <= P.min_depth) { hu = 0.0; hv = 0.0; }
    }
    dc[k] = h; dc[N + k] = hu; dc[2 * N + k] = hv;
  }
  // with lA = 0 and a scratch lR, stage 1 leaves lA = the adjoint of pre;
  // stage 2 leaves lA = half of it
  std::vector<double> r(3 * N), z(3 * N);
  const StateW lA = {lpre_h, lpre_hu, lpre_hv}, lB = st3(z.data(), N), lR = st3(r.data(), N);
  for (size_t k = 0; k < N; ++k) { lpre_h[k] = 0.0; lpre_hu[k] = 0.0; lpre_hv[k] = 0.0; }
  for (int j = 0; j < P.H; ++j)
    for (int i = 0; i < P.W; ++i)
      adj_epilogue_rev_at(P, sp3(dc.data(), N), {lout_h, lout_hu, lout_hv}, nfield, sponge, dt, lN, stage, lA, lB,
                          lR, i, j);
  if (stage == 2)
    for (size_t k = 0; k < N; ++k) { lpre_h[k] *= 2.0; lpre_hu[k] *= 2.0; lpre_hv[k] *= 2.0; }
  return SWE_OK;
}

void swe_adj_host_hllc_rev(double out[6], double hL, double huL, double hvL, double hR, double huR, double hvR,
                           double dry, double g, double lm, double ln, double lt) {
  const HllcConst C = {g, 0.5 * g, 1.0 / g, dry, g * dry >= SQRT_N_MIN};
  const HllcIn r = hllc_rev(hL, huL, hvL, hR, huR, hvR, C, lm, ln, lt);
  out[0] = r.hL; out[1] = r.huL; out[2] = r.hvL; out[3] = r.hR; out[4] = r.huR; out[5] = r.hvR;
}

}  // extern "C"
