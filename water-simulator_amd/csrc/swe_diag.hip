// Device diagnostics of one grid's current state: the reference's volume,
// energy, maxSurfaceDeviation, maxSpeed and finite (swe.mjs:819-917), read
// where the state lives instead of after a copy of h, hu, hv and b to the host.
//
// Summation.  The reference sums rows left to right and then the rows in
// order (energy on both geometries, volume on the sphere): each wavefront
// owns one row, stages DG_COLS columns of per-cell terms at a time in LDS and
// one lane per sum adds them in the reference's order, so those sums are the
// reference's bit for bit; the row sums are then added in order by one lane.
// The Cartesian volume is one running sum over all interior cells
// (swe.mjs:823-827), which no parallel order reproduces; it is accumulated in
// double-double (TwoSum) in a fixed order and rounded once, i.e. the exactly
// rounded sum, which differs from the reference's sequential rounding by at
// most that rounding's own error.  Maxima and the finite flag do not depend on
// the order and are the reference's bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "swe_ctx.hpp"
#include "swe_device.hpp"

namespace swe {
namespace {

constexpr int DG_WAVES = 16;
constexpr int DG_THREADS = 64 * DG_WAVES;
constexpr int DG_ROWS = 8;          // interior rows per block, two wavefronts each
constexpr int DG_COLS = 128;        // columns per tile, one per lane of each of the two
constexpr int DG_OUT = 8;       // result slots

struct DiagArgs {
  const double *h, *hu, *hv, *b, *area;
  double *row_v, *row_e;        // ny each: the reference-order row sums of h and of the energy terms
  double *blk;                  // 2 per block: the block's double-double sum of h
  unsigned long long *acc;      // maxSurfaceDeviation bits, maxSpeed bits, non-finite count
  double *out;                  // DG_OUT results
  int nx, ny, W, H, ng, sphere, nblocks, pad;
  double min_depth, eps4, g, ref, dx, dy;
};

struct DD { double hi, lo; };

// TwoSum (Knuth): hi + x exactly as hi' + e, the error e collected in lo
__device__ __forceinline__ DD dd_add(DD a, double x) {
  const double s = a.hi + x;
  const double bb = s - a.hi;
  const double e = (a.hi - (s - bb)) + (x - bb);
  return {s, a.lo + e};
}
__device__ __forceinline__ DD dd_merge(DD a, DD b) {
  DD s = dd_add(a, b.hi);
  s.lo = s.lo + b.lo;
  return s;
}

// PROJ: only the row sums of the surface b + h (the mode projection of
// globe.mjs:1417-1424), nothing else.
template <bool PROJ>
__global__ __launch_bounds__(DG_THREADS) void k_diag_rows(DiagArgs a) {
  // two tiles of the block's rows (h and energy terms), alternating, and the
  // running row sums between the wavefronts that take turns adding
  __shared__ double tl[2][DG_ROWS][2][DG_COLS];
  __shared__ double racc[2 * DG_ROWS];
  __shared__ double wred[DG_WAVES][4];
  __shared__ unsigned long long wnf[DG_WAVES];
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int r = w % DG_ROWS;                 // this wavefront's row of the block
  const int j = blockIdx.x * DG_ROWS + r;    // ... in the grid's interior
  const int cc = 64 * (w / DG_ROWS) + l;     // its lane's column of each tile
  if (t < 2 * DG_ROWS) racc[t] = 0.0;
  double m_sd = 0.0, m_sp = 0.0;
  unsigned long long nonfin = 0;
  DD vol = {0.0, 0.0};
  const int ntiles = (a.nx + DG_COLS - 1) / DG_COLS;
  // this lane's cell of the next tile is loaded while the current one is used
  const size_t krow = (size_t)(j + a.ng) * a.W + a.ng;
  double nh = 0.0, nhu = 0.0, nhv = 0.0, nb = 0.0;
  if (j < a.ny && cc < a.nx) {
    nh = a.h[krow + cc]; nhu = a.hu[krow + cc]; nhv = a.hv[krow + cc]; nb = a.b[krow + cc];
  }
  for (int tile = 0; tile < ntiles; ++tile) {
    const int buf = tile & 1, i0 = tile * DG_COLS;
    {
      const int i = i0 + cc;
      const double hk = nh, huk = nhu, hvk = nhv, bk = nb;
      if (j < a.ny && i + DG_COLS < a.nx) {
        const size_t k = krow + i + DG_COLS;
        nh = a.h[k]; nhu = a.hu[k]; nhv = a.hv[k]; nb = a.b[k];
      }
      // x + (-0) == x for every x: a skipped cell or a column past the grid adds nothing
      double tvv = -0.0, tee = -0.0;
      if (PROJ) {
        if (j < a.ny && i < a.nx) tvv = bk + hk;   // globe.mjs:1420: row += b + h
      } else if (j < a.ny && i < a.nx) {
        tvv = hk;
        vol = dd_add(vol, hk);
        const bool eOn = !(hk <= 0);            // swe.mjs:869 (a NaN depth is not skipped)
        const bool mOn = !(hk <= a.min_depth);  // swe.mjs:888, :902
        if (eOn || mOn) {
          // vel() of both components, once (one square root and one reciprocal);
          // a non-finite result is recomputed by the plain division, which keeps
          // the reference's infinities where the reciprocal form would give NaN
          Vel2 uv = vel2(huk, hvk, hk, a.eps4);
          if (!__builtin_isfinite(uv.u)) uv.u = vel1(huk, hk, a.eps4);
          if (!__builtin_isfinite(uv.v)) uv.v = vel1(hvk, hk, a.eps4);
          if (eOn) tee = 0.5 * hk * (uv.u * uv.u + uv.v * uv.v) + 0.5 * a.g * hk * (hk + 2 * bk);
          if (mOn) {
            const double d = fabs(bk + hk - a.ref);
            if (d > m_sd) m_sd = d;
            const double sp = js_hypot(uv.u, uv.v);
            if (sp > m_sp) m_sp = sp;
          }
        }
        if (!__builtin_isfinite(hk) || !__builtin_isfinite(huk) || !__builtin_isfinite(hvk)) ++nonfin;
      }
      tl[buf][r][0][cc] = tvv;
      tl[buf][r][1][cc] = tee;
    }
    // One barrier per tile.  Wavefront (tile mod DG_WAVES) adds this tile into the
    // block's 2 x DG_ROWS running sums in the reference's order (lane 2r: the h of
    // row r, lane 2r+1: its energy terms) while the others compute the next tile
    // into the other buffer; a buffer is written again only after the next
    // barrier, which the adding wavefront reaches after its adds.
    __syncthreads();
    if (w == tile % DG_WAVES && l < 2 * DG_ROWS) {
      const int rr = l >> 1;
      if (blockIdx.x * DG_ROWS + rr < a.ny) {
        const double *src = tl[buf][rr][l & 1];
        const int n = min(DG_COLS, a.nx - i0);
        double acc = racc[l];
        for (int c = 0; c < n; c += 8) {     // columns past n hold -0
          double x[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) x[q] = src[c + q];
#pragma unroll
          for (int q = 0; q < 8; ++q) acc += x[q];
        }
        racc[l] = acc;
      }
    }
  }
  __syncthreads();
  if (t < 2 * DG_ROWS && blockIdx.x * DG_ROWS + (t >> 1) < a.ny)
    ((t & 1) ? a.row_e : a.row_v)[blockIdx.x * DG_ROWS + (t >> 1)] = racc[t];
  if (PROJ) return;
  // finite() covers the ghost ring too (swe.mjs:912-917): 2 ng full rows, then
  // the 2 ng ghost columns of each interior row
  const int band = a.ng * a.W;
  const int nghost = 2 * band + 2 * a.ng * a.ny;
  for (int gi = blockIdx.x * DG_THREADS + t; gi < nghost; gi += gridDim.x * DG_THREADS) {
    size_t k;
    if (gi < band) {
      k = gi;
    } else if (gi < 2 * band) {
      k = (size_t)(a.H - a.ng) * a.W + (gi - band);
    } else {
      const int r = gi - 2 * band, j = r / (2 * a.ng), c = r % (2 * a.ng);
      k = (size_t)(j + a.ng) * a.W + (c < a.ng ? c : a.W - 2 * a.ng + c);
    }
    if (!__builtin_isfinite(a.h[k]) || !__builtin_isfinite(a.hu[k]) || !__builtin_isfinite(a.hv[k])) ++nonfin;
  }
  // block reductions in a fixed order
  for (int off = 32; off >= 1; off >>= 1) {
    const DD o = {__shfl_down(vol.hi, off, 64), __shfl_down(vol.lo, off, 64)};
    const double osd = __shfl_down(m_sd, off, 64), osp = __shfl_down(m_sp, off, 64);
    const unsigned long long onf = __shfl_down(nonfin, off, 64);
    if (l + off < 64) {
      vol = dd_merge(vol, o);
      m_sd = fmax(m_sd, osd);
      m_sp = fmax(m_sp, osp);
      nonfin += onf;
    }
  }
  if (l == 0) {
    wred[w][0] = vol.hi; wred[w][1] = vol.lo; wred[w][2] = m_sd; wred[w][3] = m_sp;
    wnf[w] = nonfin;
  }
  __syncthreads();
  if (t == 0) {
    DD s = {wred[0][0], wred[0][1]};
    double sd = wred[0][2], sp = wred[0][3];
    unsigned long long nf = wnf[0];
    for (int q = 1; q < DG_THREADS / 64; ++q) {
      s = dd_merge(s, DD{wred[q][0], wred[q][1]});
      sd = fmax(sd, wred[q][2]);
      sp = fmax(sp, wred[q][3]);
      nf += wnf[q];
    }
    a.blk[2 * blockIdx.x] = s.hi;
    a.blk[2 * blockIdx.x + 1] = s.lo;
    // non-negative doubles order like their bit patterns
    if (sd > 0) atomicMax(&a.acc[0], (unsigned long long)__double_as_longlong(sd));
    if (sp > 0) atomicMax(&a.acc[1], (unsigned long long)__double_as_longlong(sp));
    if (nf) atomicAdd(&a.acc[2], nf);
  }
}

// The ordered tails.  Row sums are staged through LDS FC rows at a time; one
// lane adds them in the reference's order (energy in wavefront 0, the
// spherical volume in wavefront 1), the products ahead of the adds.  The
// blocks' double-double partials of the Cartesian volume are merged by a fixed
// tree (each thread a fixed run of blocks, then pairs in LDS), so the result
// depends on the grid size only.  Clears the accumulators for the next call.
constexpr int FC = 1024;
__global__ __launch_bounds__(DG_THREADS) void k_diag_final(DiagArgs a) {
  __shared__ double se[FC], sv[FC], sa[FC];
  __shared__ double thi[DG_THREADS], tlo[DG_THREADS];
  __shared__ double res[2];
  const int t = threadIdx.x;
  double e = 0.0, v = 0.0;
  for (int j0 = 0; j0 < a.ny; j0 += FC) {
    const int n = min(FC, a.ny - j0);
    const int n8 = (n + 7) & ~7;         // padded with -0 (times an area of 1)
    for (int q = t; q < n8; q += DG_THREADS) {
      const bool in = q < n;
      se[q] = in ? a.row_e[j0 + q] : -0.0;
      sv[q] = in ? a.row_v[j0 + q] : -0.0;
      sa[q] = (in && a.sphere) ? a.area[j0 + q + a.ng] : 1.0;    // row areas: the sphere only
    }
    __syncthreads();
    if (t == 0 || (t == 64 && a.sphere)) {
      const double *src = t == 0 ? se : sv;
      double acc = t == 0 ? e : v;
      if (a.sphere) {
        for (int c = 0; c < n8; c += 8) {   // swe.mjs:833, :876
          double x[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) x[q] = src[c + q] * sa[c + q];
#pragma unroll
          for (int q = 0; q < 8; ++q) acc += x[q];
        }
      } else {
        for (int c = 0; c < n8; c += 8) {   // swe.mjs:877
          double x[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) x[q] = src[c + q];
#pragma unroll
          for (int q = 0; q < 8; ++q) acc += x[q];
        }
      }
      if (t == 0) e = acc; else v = acc;
    }
    __syncthreads();
  }
  // Cartesian volume: a fixed tree over the blocks' partials
  DD s = {0.0, 0.0};
  const int per = (a.nblocks + DG_THREADS - 1) / DG_THREADS;
  for (int q = t * per; q < min(a.nblocks, (t + 1) * per); ++q) s = dd_merge(s, DD{a.blk[2 * q], a.blk[2 * q + 1]});
  thi[t] = s.hi;
  tlo[t] = s.lo;
  __syncthreads();
  for (int half = DG_THREADS / 2; half >= 1; half >>= 1) {
    if (t < half) {
      const DD m = dd_merge(DD{thi[t], tlo[t]}, DD{thi[t + half], tlo[t + half]});
      thi[t] = m.hi;
      tlo[t] = m.lo;
    }
    __syncthreads();
  }
  if (t == 0) res[0] = e;
  if (t == 64) res[1] = v;
  __syncthreads();
  if (t == 0) {
    const double hv = thi[0], lv = tlo[0];
    const double vc = __builtin_isfinite(hv) ? hv + lv : hv;
    a.out[0] = a.sphere ? res[1] : vc * a.dx * a.dy;                 // swe.mjs:826, :836
    a.out[1] = a.sphere ? res[0] : res[0] * a.dx * a.dy;             // swe.mjs:878
    a.out[2] = __longlong_as_double((long long)a.acc[0]);
    a.out[3] = __longlong_as_double((long long)a.acc[1]);
    a.out[4] = a.acc[2] == 0 ? 1.0 : 0.0;
    a.acc[0] = 0ull;
    a.acc[1] = 0ull;
    a.acc[2] = 0ull;
  }
}

// The step's mode projection s = sum_j coef_j * (row sum of b + h), rows in
// order (globe.mjs:1417-1424 with coef_j = (wRow_j / nx) * P_n(lat_j)), as one
// sample of the projection series -- when the step ran.
__global__ __launch_bounds__(DG_THREADS) void k_proj_final(DiagArgs a, const double *coef, Ctl *ctl,
                                                          double *vals, double *times, long long cap) {
  __shared__ double sr[FC], sc[FC];
  const int t = threadIdx.x;
  const bool ran = ctl->dt_last > 0;
  if (!ran) return;
  double acc = 0.0;
  for (int j0 = 0; j0 < a.ny; j0 += FC) {
    const int n = min(FC, a.ny - j0);
    const int n8 = (n + 7) & ~7;
    for (int q = t; q < n8; q += DG_THREADS) {
      const bool in = q < n;
      sr[q] = in ? a.row_v[j0 + q] : 0.0;
      sc[q] = in ? coef[j0 + q] : 0.0;
    }
    __syncthreads();
    if (t == 0) {
      for (int c = 0; c < n8; c += 8) {
        double x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = sc[c + q] * sr[c + q];
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (c + q < n) acc += x[q];
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    const long long k = ctl->proj_count;
    if (k < cap) { vals[k] = acc; times[k] = ctl->t; }
    ctl->proj_count = k + 1;
  }
}

}  // namespace

// The projection recorder's kernels for one step, on the state the step left
// (library-internal; enqueued after the step's finalize, graph-capturable).
hipError_t launch_projection(swe_ctx *c, const double *h, const double *b) {
  DiagArgs a{};
  a.h = h; a.b = b;
  a.row_v = c->diag; a.row_e = c->diag + c->P.ny;
  a.nx = c->P.nx; a.ny = c->P.ny; a.W = c->P.W; a.H = c->P.H; a.ng = 2;
  const int nblocks = (c->P.ny + DG_ROWS - 1) / DG_ROWS;
  hipLaunchKernelGGL(k_diag_rows<true>, dim3(nblocks), dim3(DG_THREADS), 0, c->stream, a);
  hipLaunchKernelGGL(k_proj_final, dim3(1), dim3(DG_THREADS), 0, c->stream, a, c->proj_coef, c->dctl,
                     c->proj_vals, c->proj_times, c->proj_cap);
  return hipGetLastError();
}

// the diagnostics' scratch (row sums, block partials, accumulators, results)
int ensure_diag(swe_ctx *c) {
  if (c->diag) return SWE_OK;
  const int ny = c->P.ny, nblocks = (ny + DG_ROWS - 1) / DG_ROWS;
  const size_t nd = 2 * (size_t)ny + 2 * (size_t)nblocks + 3 + DG_OUT;
  hipError_t e = hipMalloc(&c->diag, nd * sizeof(double));
  if (e != hipSuccess) {
    c->diag = nullptr;
    return ctx_fail(c, e == hipErrorOutOfMemory ? SWE_ENOMEM : SWE_EHIP, std::string("diagnostics: ") + hipGetErrorString(e));
  }
  HIPCHK(c, hipMemsetAsync(c->diag, 0, nd * sizeof(double), c->stream));
  c->diag_blocks = nblocks;
  return SWE_OK;
}
}  // namespace swe

using namespace swe;

extern "C" int swe_diagnostics(swe_ctx *c, double eta_ref, double out[SWE_D_COUNT]) {
  if (!c || !out) return ctx_fail(c, SWE_EINVAL, "null argument");
  if (c->mid) return ctx_fail(c, SWE_ESTATE, "a step is in progress (swe_step_begin without swe_step_end)");
  const int nx = c->P.nx, ny = c->P.ny, ng = 2;
  const int nblocks = (ny + DG_ROWS - 1) / DG_ROWS;
  // layout: row_v[ny], row_e[ny], blk[2 nblocks], acc[3], out[DG_OUT]
  int rc = ensure_diag(c);
  if (rc) return rc;
  DiagArgs a{};
  a.h = c->st[c->A][0]; a.hu = c->st[c->A][1]; a.hv = c->st[c->A][2];
  a.b = c->b; a.area = c->rows[R_AREA];
  a.row_v = c->diag; a.row_e = c->diag + ny;
  a.blk = c->diag + 2 * ny;
  a.acc = reinterpret_cast<unsigned long long *>(c->diag + 2 * ny + 2 * nblocks);
  a.out = c->diag + 2 * ny + 2 * nblocks + 3;
  a.nx = nx; a.ny = ny; a.W = c->P.W; a.H = c->P.H; a.ng = ng;
  a.sphere = c->P.sphere; a.nblocks = nblocks;
  a.min_depth = c->P.min_depth; a.eps4 = c->P.eps4; a.g = c->P.g; a.ref = eta_ref;
  // the Cartesian grid's dx, dy: every row of dxRow / dyRowN holds them (geometry.mjs:169-171)
  a.dx = c->hrows[SWE_ROW_DX].empty() ? 0.0 : c->hrows[SWE_ROW_DX][ng];
  a.dy = c->hrows[SWE_ROW_DYN].empty() ? 0.0 : c->hrows[SWE_ROW_DYN][ng];
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(k_diag_rows<false>, dim3(nblocks), dim3(DG_THREADS), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(DG_THREADS), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  double res[DG_OUT];
  HIPCHK(c, hipMemcpyAsync(res, a.out, sizeof(res), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->diag_ms = ms;
  std::memcpy(out, res, sizeof(double) * SWE_D_COUNT);
  return SWE_OK;
}

// ---- swe_diagnostics_noalloc: the same kernels on caller-owned fields, with
// their scratch in a static device array instead of an allocation -----------
//
// Layout of the array (doubles): [0, 3) the three accumulators (bit patterns of
// the two maxima, the non-finite count; they must start at zero, so every call
// clears them first), [3, 3 + DG_OUT) the results, then from 16 the per-row sums
// row_v and row_e (NA_ROWS each) and the blocks' volume partials (2 per block).
// Each GPU holds its own copy until the process ends, at its own address.
namespace swe {
namespace {
constexpr int NA_ROWS = SWE_DIAG_NOALLOC_MAX_ROWS;
constexpr int NA_BLOCKS = (NA_ROWS + DG_ROWS - 1) / DG_ROWS;
constexpr size_t NA_DOUBLES = 16 + 2 * (size_t)NA_ROWS + 2 * (size_t)NA_BLOCKS;
constexpr int NA_DEVICES = 64;

__global__ void k_clear_acc(unsigned long long *acc) {
  if (threadIdx.x < 3) acc[threadIdx.x] = 0ull;
}
}  // namespace
}  // namespace swe

static __device__ __align__(256) double scratch[swe::NA_DOUBLES];

namespace {
std::mutex scratch_mu;
double *scratch_dev[swe::NA_DEVICES];   // the array's address on each device, once looked up

// The current device's copy of `scratch`: its address is looked up once per
// device (with that device current) and kept; host code uses only this
// address, never the array's name, which on the host is a stand-in.
int scratch_on_device(double **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return swe::ctx_fail(nullptr, SWE_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= swe::NA_DEVICES) return swe::ctx_fail(nullptr, SWE_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return swe::ctx_fail(nullptr, SWE_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    scratch_dev[dev] = static_cast<double *>(p);
  }
  *out = scratch_dev[dev];
  return SWE_OK;
}
}  // namespace

extern "C" int swe_diagnostics_noalloc(int32_t nx, int32_t ny, int32_t kind, const double *b, const double *h,
                                       const double *hu, const double *hv, const double *area, double dx, double dy,
                                       double min_depth, double eps4, double g, double eta_ref,
                                       double out[SWE_D_COUNT]) {
  if (!b || !h || !hu || !hv || !out) return ctx_fail(nullptr, SWE_EINVAL, "swe_diagnostics_noalloc: null argument");
  if (nx < 2 || ny < 1) return ctx_fail(nullptr, SWE_EINVAL, "swe_diagnostics_noalloc: nx must be >= 2 and ny >= 1");
  if (ny > SWE_DIAG_NOALLOC_MAX_ROWS)
    return ctx_fail(nullptr, SWE_EINVAL, "swe_diagnostics_noalloc: ny = " + std::to_string(ny) + " exceeds its " +
                                             std::to_string(SWE_DIAG_NOALLOC_MAX_ROWS) + "-row scratch");
  if (kind != SWE_CARTESIAN && kind != SWE_SPHERE) return ctx_fail(nullptr, SWE_EINVAL, "swe_diagnostics_noalloc: unknown kind");
  if (kind == SWE_SPHERE && !area) return ctx_fail(nullptr, SWE_EINVAL, "swe_diagnostics_noalloc: the sphere needs its row areas");
  double *sp = nullptr;
  int rc = scratch_on_device(&sp);
  if (rc) return rc;
  const int ng = 2, nblocks = (ny + DG_ROWS - 1) / DG_ROWS;
  DiagArgs a{};
  a.h = h; a.hu = hu; a.hv = hv; a.b = b; a.area = area;
  a.acc = reinterpret_cast<unsigned long long *>(sp);
  a.out = sp + 3;
  a.row_v = sp + 16;
  a.row_e = sp + 16 + NA_ROWS;
  a.blk = sp + 16 + 2 * (size_t)NA_ROWS;
  a.nx = nx; a.ny = ny; a.W = nx + 2 * ng; a.H = ny + 2 * ng; a.ng = ng;
  a.sphere = kind == SWE_SPHERE; a.nblocks = nblocks;
  a.min_depth = min_depth; a.eps4 = eps4; a.g = g; a.ref = eta_ref; a.dx = dx; a.dy = dy;
  hipLaunchKernelGGL(k_clear_acc, dim3(1), dim3(64), 0, nullptr, a.acc);
  hipLaunchKernelGGL(k_diag_rows<false>, dim3(nblocks), dim3(DG_THREADS), 0, nullptr, a);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(DG_THREADS), 0, nullptr, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ctx_fail(nullptr, SWE_EHIP, std::string("swe_diagnostics_noalloc: ") + hipGetErrorString(e));
  double res[DG_OUT];
  e = hipMemcpy(res, a.out, sizeof(res), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return ctx_fail(nullptr, SWE_EHIP, std::string("swe_diagnostics_noalloc: ") + hipGetErrorString(e));
  std::memcpy(out, res, sizeof(double) * SWE_D_COUNT);
  return SWE_OK;
}
