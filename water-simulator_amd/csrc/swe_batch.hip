// Batched ensembles (include/swe_amd.h, swe_batch_*): many independent basins of
// uneven sizes advanced together, one launch per kernel phase for the whole
// ensemble, each basin with its own dt from its own CFL reduction, its own time,
// boundary kinds, forcing and friction.  The stage kernels run as a persistent
// grid pulling (basin, strip, chunk) tasks from a queue; one step is captured in
// a HIP graph and replayed.  Every basin's numbers are bit-identical to a plain
// run of the same basin (same device code, same per-basin order of operations).
//
// Memory: one pool holds every basin's fields (b, three h/hu/hv buffers, optional
// Manning and sponge fields), one holds the geometry rows and their reciprocals,
// one the massFloored partials; basins are carved out at 256-byte alignment.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/swe_amd.h"
#include "swe_device.hpp"
#include "swe_kernels.hpp"

namespace swe {
hipError_t launch_sweep_batch(int mode, const BasinDev *basins, const Task *tasks, int ntasks, unsigned *qctr,
                              int parity, int nwaves, int lim_all, bool uni_all, bool lean_all, hipStream_t st);
const void *k_sweep_batch_ptr();
hipError_t launch_apply_bc_batch(const BasinDev *basins, int nb, int max_ghost, int parity, int stage, hipStream_t st);
hipError_t launch_maxdt_batch(const BasinDev *basins, int nb, int max_cells, int parity, hipStream_t st);
hipError_t launch_finalize_batch(const BasinDev *basins, int nb, unsigned *qctr, int phase, hipStream_t st);
int tide_table(const Params &P, int n, const int32_t *m, const double *amp, const double *lat, const double *lon,
               std::vector<double> &out);
}  // namespace swe

using namespace swe;

namespace {
constexpr int MIN_CHUNK = 8;
constexpr size_t ALIGN = 32;   // doubles (256 bytes)
enum { R_DX, R_RDX, R_DYN, R_RDYN, R_DYS, R_RDYS, R_DYCFL, R_RDYCFL, R_BED, R_RBED, R_TAN, R_F, R_GEO, R_COUNT };

size_t up(size_t n) { return (n + ALIGN - 1) / ALIGN * ALIGN; }

struct Basin {
  Params P{};
  size_t N = 0;
  size_t off_field = 0;    // into the field pool: b, st[3][3], manning, sponge
  size_t off_rows = 0;     // into the row pool: R_COUNT rows of H
  size_t off_rowsA = 0;    // into the row pool: the rows interleaved, GEO_AOS per row (Params::geoA)
  size_t off_part = 0;     // into the partial pool: 2 * max_parts
  int max_parts = 0;
  double manning = 0.0, coriolis = 0.0;
  bool has_mfield = false, has_sponge = false;
  double *tide = nullptr;  // equilibrium-tide coefficient table of this basin
};
}  // namespace

struct swe_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<Basin> basins;
  double *fields = nullptr, *rows = nullptr, *parts = nullptr;
  Ctl *dctl = nullptr;
  Ctl *hctl = nullptr;
  BasinDev *dbasins = nullptr;
  Task *dtasks = nullptr;
  unsigned *qctr = nullptr;
  int ntasks = 0;
  int parity = 0;             // 0: current state in buffer 0; 1: in buffer 2
  int chunk = 32;
  int nwaves = 0;             // persistent grid size
  int lim_all = -1;           // the slope limiter every basin uses, or -1
  bool uni_all = false;       // every basin's rows are alike (Params::geo_uni)
  bool lean_all = false;      // no basin has a per-cell Manning field or a sponge (set by refresh)
  int max_ghost = 0, max_cells = 0;
  bool dirty = true;
  bool use_graph = true;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  std::string err;
};

static thread_local std::string g_batch_err;

#define BCHK(b, expr)                                                            \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      (b)->err = std::string(#expr) + ": " + hipGetErrorString(e_);              \
      return SWE_EHIP;                                                           \
    }                                                                            \
  } while (0)

static int bfail(swe_batch *b, int code, const std::string &m) {
  if (b) b->err = m; else g_batch_err = m;
  return code;
}

// field slots inside a basin's share of the field pool
enum { FS_B = 0, FS_ST = 1 /* 9 slots */, FS_MANNING = 10, FS_SPONGE = 11, FS_COUNT = 12 };

static double *fslot(swe_batch *b, const Basin &B, int slot) { return b->fields + B.off_field + (size_t)slot * up(B.N); }
static double *state(swe_batch *b, const Basin &B, int buf, int f) { return fslot(b, B, FS_ST + 3 * buf + f); }
static double *rowp(swe_batch *b, const Basin &B, int r) { return b->rows + B.off_rows + (size_t)r * up(B.P.H); }

static void drop_graphs(swe_batch *b) {
  for (int i = 0; i < 2; ++i) {
    if (b->gexec[i]) { (void)hipGraphExecDestroy(b->gexec[i]); b->gexec[i] = nullptr; }
    if (b->graph[i]) { (void)hipGraphDestroy(b->graph[i]); b->graph[i] = nullptr; }
  }
}

static int strips_of(const Params &P) { return (P.W + SWEEP_OUT - 1) / SWEEP_OUT; }
static int chunks_of(const Params &P, int rows) { return (P.H + rows - 1) / rows; }

// rebuild the device basin table and the task list (after a parameter change)
static int refresh(swe_batch *b) {
  if (!b->dirty) return SWE_OK;
  const int nb = (int)b->basins.size();
  std::vector<BasinDev> dev(nb);
  std::vector<Task> tasks;
  // the step graphs hold the stage-2 kernel chosen here; refresh drops them
  b->lean_all = true;
  for (const Basin &B : b->basins)
    if (B.has_mfield || B.has_sponge || B.P.want_theta || B.P.want_rh) b->lean_all = false;
  std::vector<int> order(nb);
  std::iota(order.begin(), order.end(), 0);
  // largest basins first, so the queue ends on small tasks of small basins
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->basins[x].N > b->basins[y].N; });
  for (int i = 0; i < nb; ++i) {
    Basin &B = b->basins[i];
    Params P = B.P;
    P.friction = B.has_mfield ? 2 : ((B.manning != 0.0 && B.manning == B.manning) ? 1 : 0);
    P.manning = B.manning;
    P.coriolis = B.coriolis;
    P.coriolis_on = (B.coriolis != 0.0 && B.coriolis == B.coriolis) ? 1 : 0;
    P.sponge = B.has_sponge ? 1 : 0;
    Geo &G = P.geo;
    G.dx = rowp(b, B, R_DX); G.rdx = rowp(b, B, R_RDX); G.dyN = rowp(b, B, R_DYN); G.rdyN = rowp(b, B, R_RDYN);
    G.dyS = rowp(b, B, R_DYS); G.rdyS = rowp(b, B, R_RDYS); G.dyCFL = rowp(b, B, R_DYCFL);
    G.rdyCFL = rowp(b, B, R_RDYCFL); G.bedPhi = rowp(b, B, R_BED); G.rbedPhi = rowp(b, B, R_RBED);
    G.tanPhi = rowp(b, B, R_TAN); G.fRow = rowp(b, B, R_F); G.geo = rowp(b, B, R_GEO);
    P.geoA = b->rows + B.off_rowsA;
    const int np = strips_of(P) * chunks_of(P, b->chunk);
    if (np > B.max_parts) return bfail(b, SWE_EINVAL, "chunk too small for a basin");
    BasinDev &D = dev[i];
    std::memset(&D, 0, sizeof(D));
    D.ctl = b->dctl + i;
    D.nparts = np;
    for (int par = 0; par < 2; ++par) {
      const int Ai = par == 0 ? 0 : 2, Ci = par == 0 ? 2 : 0, Bi = 1;
      StateW SA = {state(b, B, Ai, 0), state(b, B, Ai, 1), state(b, B, Ai, 2)};
      StateW SB = {state(b, B, Bi, 0), state(b, B, Bi, 1), state(b, B, Bi, 2)};
      StateW SC = {state(b, B, Ci, 0), state(b, B, Ci, 1), state(b, B, Ci, 2)};
      for (int stg = 0; stg < 2; ++stg) {
        SweepArgs &a = D.sw[par][stg];
        a.P = P;
        a.b = fslot(b, B, FS_B);
        a.mfield = B.has_mfield ? fslot(b, B, FS_MANNING) : nullptr;
        a.sponge = B.has_sponge ? fslot(b, B, FS_SPONGE) : nullptr;
        a.ctl = b->dctl + i;
        a.rows_per_chunk = b->chunk;
        a.mf_part = b->parts + B.off_part + (stg == 0 ? 0 : np);
        if (stg == 0) { a.in = {SA.h, SA.hu, SA.hv}; a.out = SB; }
        else { a.in = {SB.h, SB.hu, SB.hv}; a.h0 = {SA.h, SA.hu, SA.hv}; a.out = SC; }
      }
      D.bc_in[par][0] = SA;
      D.bc_in[par][1] = SB;
    }
  }
  // the finalize kernel reads partials through sw[0][0].mf_part; keep it pointing at slot 0
  for (int oi = 0; oi < nb; ++oi) {
    const int i = order[oi];
    const Params &P = b->basins[i].P;
    const int S = strips_of(P), C = chunks_of(P, b->chunk);
    for (int c = 0; c < C; ++c)
      for (int s = 0; s < S; ++s) tasks.push_back({i, s, c, c * S + s});
  }
  (void)hipFree(b->dtasks);
  b->dtasks = nullptr;
  BCHK(b, hipMalloc(&b->dtasks, sizeof(Task) * std::max<size_t>(1, tasks.size())));
  BCHK(b, hipMemcpyAsync(b->dtasks, tasks.data(), sizeof(Task) * tasks.size(), hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipMemcpyAsync(b->dbasins, dev.data(), sizeof(BasinDev) * nb, hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  b->ntasks = (int)tasks.size();
  b->dirty = false;
  drop_graphs(b);
  return SWE_OK;
}

static int pull(swe_batch *b) {
  BCHK(b, hipMemcpyAsync(b->hctl, b->dctl, sizeof(Ctl) * b->basins.size(), hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return SWE_OK;
}
static int push(swe_batch *b) {
  BCHK(b, hipMemcpyAsync(b->dctl, b->hctl, sizeof(Ctl) * b->basins.size(), hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return SWE_OK;
}

static int enqueue_step(swe_batch *b) {
  const int nb = (int)b->basins.size();
  const int par = b->parity;
  BCHK(b, launch_apply_bc_batch(b->dbasins, nb, b->max_ghost, par, 0, b->stream));
  BCHK(b, launch_sweep_batch(MODE_STAGE1, b->dbasins, b->dtasks, b->ntasks, b->qctr, par, b->nwaves, b->lim_all,
                             b->uni_all, false, b->stream));
  BCHK(b, launch_apply_bc_batch(b->dbasins, nb, b->max_ghost, par, 1, b->stream));
  BCHK(b, launch_sweep_batch(MODE_STAGE2, b->dbasins, b->dtasks, b->ntasks, b->qctr, par, b->nwaves, b->lim_all,
                             b->uni_all, b->lean_all, b->stream));
  BCHK(b, launch_finalize_batch(b->dbasins, nb, b->qctr, 1, b->stream));
  b->parity ^= 1;
  return SWE_OK;
}

extern "C" {

const char *swe_batch_last_error(const swe_batch *b) { return b ? b->err.c_str() : g_batch_err.c_str(); }

int swe_batch_create(const swe_desc *descs, int nb, swe_batch **out) {
  if (!descs || nb < 1 || !out) return bfail(nullptr, SWE_EINVAL, "bad arguments");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return bfail(nullptr, SWE_EHIP, "no HIP device");
  swe_batch *b = new swe_batch();
  b->basins.resize(nb);
  size_t fpool = 0, rpool = 0, ppool = 0;
  for (int i = 0; i < nb; ++i) {
    const swe_desc &d = descs[i];
    if (d.nx < 2 || d.ny < 1 || (d.order != 1 && d.order != 2) || d.limiter < 0 || d.limiter > 2 ||
        (d.kind != SWE_CARTESIAN && d.kind != SWE_SPHERE)) {
      delete b;
      return bfail(nullptr, SWE_EINVAL, "basin " + std::to_string(i) + ": invalid description");
    }
    Basin &B = b->basins[i];
    Params &P = B.P;
    P.nx = d.nx; P.ny = d.ny; P.W = d.nx + 4; P.H = d.ny + 4;
    P.sphere = d.kind == SWE_SPHERE; P.order = d.order; P.limiter = d.limiter;
    for (int s = 0; s < 4; ++s) P.bc[s] = P.sphere ? (s < 2 ? BC_PERIODIC : BC_REFLECT) : BC_REFLECT;
    P.g = d.g; P.half_g = 0.5 * d.g; P.neg_half_g = -d.g * 0.5; P.rg = 1.0 / d.g;
    P.min_depth = d.min_depth; P.eps4 = d.eps4; P.cfl = d.cfl; P.R = d.R;
    B.manning = d.manning; B.coriolis = d.coriolis;
    B.N = (size_t)P.W * P.H;
    P.fstride = (int)up(B.N);        // buffer fields and rows sit up(N), up(H) apart in the pools
    P.rstride = (int)up(P.H);
    B.off_field = fpool; fpool += FS_COUNT * up(B.N);
    B.off_rows = rpool; rpool += R_COUNT * up(P.H);
    B.off_rowsA = rpool; rpool += GEO_AOS * up(P.H);
    B.max_parts = strips_of(P) * chunks_of(P, MIN_CHUNK);
    B.off_part = ppool; ppool += up(2 * (size_t)B.max_parts);
    b->max_ghost = std::max(b->max_ghost, 4 * P.W + 4 * (P.H - 4));
    b->max_cells = std::max(b->max_cells, P.nx * P.ny);
  }
  auto bail = [&](const char *what, hipError_t e) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    swe_batch_destroy(b);
    return bfail(nullptr, e == hipErrorOutOfMemory ? SWE_ENOMEM : SWE_EHIP, m);
  };
  hipError_t e;
  if ((e = hipGetDevice(&b->device)) != hipSuccess) return bail("hipGetDevice", e);
  if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return bail("stream", e);
  if ((e = hipMalloc(&b->fields, sizeof(double) * fpool)) != hipSuccess) return bail("field pool", e);
  if ((e = hipMalloc(&b->rows, sizeof(double) * rpool)) != hipSuccess) return bail("row pool", e);
  if ((e = hipMalloc(&b->parts, sizeof(double) * ppool)) != hipSuccess) return bail("partials", e);
  if ((e = hipMalloc(&b->dctl, sizeof(Ctl) * nb)) != hipSuccess) return bail("ctl", e);
  if ((e = hipHostMalloc((void **)&b->hctl, sizeof(Ctl) * nb, hipHostMallocDefault)) != hipSuccess) return bail("pinned ctl", e);
  if ((e = hipMalloc(&b->dbasins, sizeof(BasinDev) * nb)) != hipSuccess) return bail("basin table", e);
  if ((e = hipMalloc(&b->qctr, sizeof(unsigned) * 2)) != hipSuccess) return bail("queue", e);
  if ((e = hipEventCreate(&b->ev0)) != hipSuccess) return bail("event", e);
  if ((e = hipEventCreate(&b->ev1)) != hipSuccess) return bail("event", e);
  (void)hipMemsetAsync(b->fields, 0, sizeof(double) * fpool, b->stream);
  (void)hipMemsetAsync(b->parts, 0, sizeof(double) * ppool, b->stream);
  (void)hipMemsetAsync(b->qctr, 0, sizeof(unsigned) * 2, b->stream);
  std::memset(b->hctl, 0, sizeof(Ctl) * nb);
  for (int i = 0; i < nb; ++i) b->hctl[i].dt_mode = DT_SELF;
  if ((e = hipMemcpyAsync(b->dctl, b->hctl, sizeof(Ctl) * nb, hipMemcpyHostToDevice, b->stream)) != hipSuccess)
    return bail("ctl upload", e);
  // geometry rows and their correctly rounded reciprocals
  const int src_of[R_COUNT] = {SWE_ROW_DX, SWE_ROW_DX, SWE_ROW_DYN, SWE_ROW_DYN, SWE_ROW_DYS, SWE_ROW_DYS,
                               SWE_ROW_DYCFL, SWE_ROW_DYCFL, SWE_ROW_BEDPHI, SWE_ROW_BEDPHI, SWE_ROW_TANPHI,
                               SWE_ROW_F, SWE_ROW_GEO};
  const bool recip[R_COUNT] = {false, true, false, true, false, true, false, true, false, true, false, false, false};
  std::vector<double> host(rpool, 0.0);
  for (int i = 0; i < nb; ++i) {
    const Basin &B = b->basins[i];
    for (int r = 0; r < R_COUNT; ++r) {
      const double *src = descs[i].rows[src_of[r]];
      if (!src) { swe_batch_destroy(b); return bfail(nullptr, SWE_EINVAL, "missing geometry row"); }
      double *dst = host.data() + B.off_rows + (size_t)r * up(B.P.H);
      double *dstA = host.data() + B.off_rowsA + r;
      for (int j = 0; j < B.P.H; ++j) {
        dst[j] = recip[r] ? 1.0 / src[j] : src[j];
        dstA[(size_t)j * GEO_AOS] = dst[j];
      }
    }
    double *A = host.data() + B.off_rowsA;
    for (int j = 0; j < B.P.H; ++j) A[(size_t)j * GEO_AOS + GEO_TANR] = A[(size_t)j * GEO_AOS + R_TAN] / B.P.R;
    set_uniform_rows(b->basins[i].P, A, B.P.H);
  }
  b->lim_all = b->basins[0].P.limiter;
  b->uni_all = true;
  for (const Basin &B : b->basins) {
    if (B.P.limiter != b->lim_all) b->lim_all = -1;
    if (!B.P.geo_uni) b->uni_all = false;
  }
  if ((e = hipMemcpyAsync(b->rows, host.data(), sizeof(double) * rpool, hipMemcpyHostToDevice, b->stream)) != hipSuccess)
    return bail("rows upload", e);
  if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return bail("sync", e);
  int cus = 0, per_cu = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep_batch_ptr(), SWEEP_LANES, 0);
  b->nwaves = std::max(1, cus) * std::max(1, per_cu);
  *out = b;
  return SWE_OK;
}

void swe_batch_destroy(swe_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  for (auto &B : b->basins) (void)hipFree(B.tide);
  (void)hipFree(b->fields); (void)hipFree(b->rows); (void)hipFree(b->parts);
  (void)hipFree(b->dctl); (void)hipFree(b->dbasins); (void)hipFree(b->dtasks); (void)hipFree(b->qctr);
  if (b->hctl) (void)hipHostFree(b->hctl);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

static double *bfield(swe_batch *b, int i, int f) {
  const Basin &B = b->basins[i];
  const int A = b->parity == 0 ? 0 : 2;
  switch (f) {
    case SWE_F_B: return fslot(b, B, FS_B);
    case SWE_F_H: return state(b, B, A, 0);
    case SWE_F_HU: return state(b, B, A, 1);
    case SWE_F_HV: return state(b, B, A, 2);
    case SWE_F_MANNING: return fslot(b, B, FS_MANNING);
    case SWE_F_SPONGE: return fslot(b, B, FS_SPONGE);
    default: return nullptr;
  }
}

int swe_batch_set_field(swe_batch *b, int i, int f, const double *host) {
  if (!b || i < 0 || i >= (int)b->basins.size()) return bfail(b, SWE_EINVAL, "bad basin");
  Basin &B = b->basins[i];
  double *p = bfield(b, i, f);
  if (!p) return bfail(b, SWE_EINVAL, "field is not writable");
  if (f == SWE_F_MANNING || f == SWE_F_SPONGE) {
    bool &flag = f == SWE_F_MANNING ? B.has_mfield : B.has_sponge;
    if (flag != (host != nullptr)) b->dirty = true;
    flag = host != nullptr;
    if (!host) return SWE_OK;
    if (f == SWE_F_MANNING)
      for (size_t k = 0; k < B.N; ++k)
        if (!(host[k] >= 0) || !std::isfinite(host[k])) return bfail(b, SWE_EINVAL, "Manning n must be finite and >= 0");
  } else if (!host) {
    return bfail(b, SWE_EINVAL, "null host pointer");
  }
  BCHK(b, hipMemcpyAsync(p, host, sizeof(double) * B.N, hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return SWE_OK;
}

int swe_batch_get_field(swe_batch *b, int i, int f, double *host) {
  if (!b || !host || i < 0 || i >= (int)b->basins.size()) return bfail(b, SWE_EINVAL, "bad basin");
  double *p = bfield(b, i, f);
  if (!p) return bfail(b, SWE_EINVAL, "field not available");
  BCHK(b, hipMemcpyAsync(host, p, sizeof(double) * b->basins[i].N, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return SWE_OK;
}

int swe_batch_set_bc(swe_batch *b, int i, int side, int kind) {
  if (!b || i < 0 || i >= (int)b->basins.size() || side < 0 || side > 3 || kind < 0 || kind > SWE_BC_FLATHER)
    return bfail(b, SWE_EINVAL, "bad boundary");
  if (b->basins[i].P.bc[side] != kind) b->dirty = true;
  b->basins[i].P.bc[side] = kind;
  return SWE_OK;
}

int swe_batch_set_harmonic(swe_batch *b, int i, int side, int n_eta, double msl, const double *amp,
                           const double *omega, const double *phase, int n_u, double u0, const double *uamp,
                           const double *uomega, const double *uphase) {
  if (!b || i < 0 || i >= (int)b->basins.size() || side < 0 || side > 3 || n_eta < 0 || n_eta > MAX_HARM ||
      n_u < 0 || n_u > MAX_HARM)
    return bfail(b, SWE_EINVAL, "bad harmonic forcing");
  int rc = pull(b);
  if (rc) return rc;
  Harm &hm = b->hctl[i].harm[side];
  std::memset(&hm, 0, sizeof(hm));
  hm.n_eta = n_eta; hm.msl = msl; hm.n_u = n_u; hm.u0 = u0;
  for (int k = 0; k < n_eta; ++k) { hm.amp[k] = amp[k]; hm.omega[k] = omega[k]; hm.phase[k] = phase[k]; }
  for (int k = 0; k < n_u; ++k) { hm.uamp[k] = uamp[k]; hm.uomega[k] = uomega[k]; hm.uphase[k] = uphase[k]; }
  b->hctl[i].harm_on = 1;
  return push(b);
}

int swe_batch_set_tide(swe_batch *b, int i, int n, const int32_t *m, const double *amp, const double *omega,
                       const double *phase, const double *lat, const double *lon) {
  if (!b || i < 0 || i >= (int)b->basins.size()) return bfail(b, SWE_EINVAL, "bad basin");
  Basin &B = b->basins[i];
  if (n < 0 || n > MAX_TIDE) return bfail(b, SWE_EINVAL, "tide: 0..4 constituents");
  if (n > 0 && (!B.P.sphere || !(B.P.R > 0) || std::isinf(B.P.R)))
    return bfail(b, SWE_EINVAL, "tide: the equilibrium tide needs a spherical grid");
  std::vector<double> T;
  if (n > 0 && (!omega || !phase || tide_table(B.P, n, m, amp, lat, lon, T)))
    return bfail(b, SWE_EINVAL, "tide: bad constituents (m must be 1 or 2, amplitudes and angles finite)");
  BCHK(b, hipStreamSynchronize(b->stream));
  drop_graphs(b);
  BCHK(b, hipFree(B.tide));
  B.tide = nullptr;
  B.P.tide = nullptr;
  B.P.tide_n = 0;
  if (n > 0) {
    BCHK(b, hipMalloc(&B.tide, sizeof(double) * T.size()));
    BCHK(b, hipMemcpy(B.tide, T.data(), sizeof(double) * T.size(), hipMemcpyHostToDevice));
    B.P.tide = B.tide;
    B.P.tide_n = n;
  }
  b->dirty = true;
  int rc = pull(b);
  if (rc) return rc;
  b->hctl[i].tide_n = n;
  for (int k = 0; k < n; ++k) { b->hctl[i].tide_om[k] = omega[k]; b->hctl[i].tide_ph[k] = phase[k]; }
  return push(b);
}

int swe_batch_get_scalar(swe_batch *b, int i, int which, double *v) {
  if (!b || !v || i < 0 || i >= (int)b->basins.size()) return bfail(b, SWE_EINVAL, "bad basin");
  int rc = pull(b);
  if (rc) return rc;
  const Ctl &c = b->hctl[i];
  switch (which) {
    case SWE_S_T: *v = c.t; break;
    case SWE_S_STEPS: *v = (double)c.steps; break;
    case SWE_S_MASS_FLOORED: *v = c.mf; break;
    case SWE_S_DT_NEXT: *v = c.dt_next; break;
    case SWE_S_NONFINITE: *v = c.nonfinite; break;
    case SWE_S_CHUNK_ROWS: *v = b->chunk; break;
    default: return bfail(b, SWE_EINVAL, "unknown scalar");
  }
  return SWE_OK;
}

int swe_batch_get_ext(swe_batch *b, int i, double ext[8]) {
  if (!b || !ext || i < 0 || i >= (int)b->basins.size()) return bfail(b, SWE_EINVAL, "bad basin");
  int rc = pull(b);
  if (rc) return rc;
  for (int k = 0; k < 8; ++k) ext[k] = b->hctl[i].ext[k];
  return SWE_OK;
}

int swe_batch_set_chunk_rows(swe_batch *b, int rows, int nwaves) {
  if (!b || rows < MIN_CHUNK) return bfail(b, SWE_EINVAL, "chunk rows below minimum (8)");
  if (rows != b->chunk) b->dirty = true;
  b->chunk = rows;
  if (nwaves > 0 && nwaves != b->nwaves) { b->nwaves = nwaves; drop_graphs(b); }
  return SWE_OK;
}

int swe_batch_use_graph(swe_batch *b, int on) {
  if (!b) return bfail(nullptr, SWE_EINVAL, "null batch");
  b->use_graph = on != 0;
  if (!on) drop_graphs(b);
  return SWE_OK;
}

static int build_graph(swe_batch *b, int which) {
  BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int saved = b->parity;
  int rc = enqueue_step(b);
  b->parity = saved;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return bfail(b, SWE_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
  b->graph[which] = g;
  BCHK(b, hipGraphInstantiate(&b->gexec[which], g, nullptr, nullptr, 0));
  return SWE_OK;
}

int swe_batch_advance(swe_batch *b, int64_t nsteps, int64_t *min_steps) {
  if (!b || nsteps < 0) return bfail(b, SWE_EINVAL, "bad advance args");
  int rc = refresh(b);
  if (rc) return rc;
  const int nb = (int)b->basins.size();
  if ((rc = pull(b))) return rc;
  for (int i = 0; i < nb; ++i) { b->hctl[i].dt_mode = DT_SELF; b->hctl[i].cfl_bits = 0; }
  if ((rc = push(b))) return rc;
  BCHK(b, launch_maxdt_batch(b->dbasins, nb, b->max_cells, b->parity, b->stream));
  BCHK(b, launch_finalize_batch(b->dbasins, nb, b->qctr, 0, b->stream));
  BCHK(b, hipEventRecord(b->ev0, b->stream));
  for (int64_t n = 0; n < nsteps; ++n) {
    if (b->use_graph) {
      const int which = b->parity;
      if (!b->gexec[which] && (rc = build_graph(b, which))) return rc;
      BCHK(b, hipGraphLaunch(b->gexec[which], b->stream));
      b->parity ^= 1;
    } else if ((rc = enqueue_step(b))) {
      return rc;
    }
  }
  BCHK(b, hipEventRecord(b->ev1, b->stream));
  BCHK(b, hipEventSynchronize(b->ev1));
  float ms = 0.f;
  BCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  if ((rc = pull(b))) return rc;
  if (min_steps) {
    long long m = b->hctl[0].steps;
    for (int i = 1; i < nb; ++i) m = std::min(m, b->hctl[i].steps);
    *min_steps = m;
  }
  return SWE_OK;
}

int swe_batch_last_advance_ms(swe_batch *b, double *ms) {
  if (!b || !ms) return bfail(b, SWE_EINVAL, "null argument");
  *ms = b->last_ms;
  return SWE_OK;
}

int swe_batch_autotune(swe_batch *b, int steps, int *best_rows, int *best_waves) {
  if (!b) return bfail(nullptr, SWE_EINVAL, "null batch");
  if (steps < 1) steps = 2;
  // Trials advance a scratch copy: save every basin's state and control block,
  // time candidates, then restore.
  const int nb = (int)b->basins.size();
  int rc = pull(b);
  if (rc) return rc;
  std::vector<Ctl> ctl0(b->hctl, b->hctl + nb);
  const int par0 = b->parity;
  std::vector<std::vector<double>> save(nb);
  for (int i = 0; i < nb; ++i) {
    save[i].resize(4 * b->basins[i].N);
    for (int f = 0; f < 4; ++f)
      if ((rc = swe_batch_get_field(b, i, SWE_F_B + f, save[i].data() + f * b->basins[i].N))) return rc;
  }
  int cus = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  const int base_waves = b->nwaves;
  const int old_rows = b->chunk;
  int best_r = old_rows, best_w = base_waves;
  double best = 1e300;
  for (int rows : {8, 12, 16, 24, 32, 48, 64}) {
    for (int wmul : {1, 2}) {
      if ((rc = swe_batch_set_chunk_rows(b, rows, base_waves * wmul))) continue;
      // restore the starting state for every trial
      b->parity = par0;
      std::memcpy(b->hctl, ctl0.data(), sizeof(Ctl) * nb);
      if ((rc = push(b))) return rc;
      for (int i = 0; i < nb; ++i)
        for (int f = 0; f < 4; ++f)
          if ((rc = swe_batch_set_field(b, i, SWE_F_B + f, save[i].data() + f * b->basins[i].N))) return rc;
      // one untimed step first: it captures the candidate's step graphs, whose
      // host-side capture would otherwise sit inside the timed window; then the
      // faster of two timed runs
      int64_t done = 0;
      if ((rc = swe_batch_advance(b, 1, &done))) return rc;
      double t = 1e300;
      for (int rep = 0; rep < 2; ++rep) {
        if ((rc = swe_batch_advance(b, steps, &done))) return rc;
        t = std::min(t, b->last_ms);
      }
      if (t < best) { best = t; best_r = rows; best_w = base_waves * wmul; }
    }
  }
  b->parity = par0;
  std::memcpy(b->hctl, ctl0.data(), sizeof(Ctl) * nb);
  if ((rc = push(b))) return rc;
  for (int i = 0; i < nb; ++i)
    for (int f = 0; f < 4; ++f)
      if ((rc = swe_batch_set_field(b, i, SWE_F_B + f, save[i].data() + f * b->basins[i].N))) return rc;
  b->nwaves = base_waves;
  if ((rc = swe_batch_set_chunk_rows(b, best_r, best_w))) return rc;
  if (best_rows) *best_rows = best_r;
  if (best_waves) *best_waves = best_w;
  return SWE_OK;
}

}  // extern "C"
