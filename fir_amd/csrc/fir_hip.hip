// fir_hip.hip -- gfx950 kernels and the C-ABI of include/fir_hip.h.
//
// Plain (on-the-fly) path: every projector kernel maps one LOR to one lane.
// Consecutive lanes take consecutive tangential bins of one
// (segment, view, axial) row, i.e. neighbouring, nearly parallel LORs of one
// fan (reference ProjData.cc:310-319): their walks touch neighbouring voxels
// and their sinogram reads/writes are contiguous.
//   - forward: gather, FP64 walk, FP32 accumulator updated as
//     float(double(acc) + len * double(v)) -- the reference's
//     VolData::computeLineIntegral (VolData.cc:357-371) -- so bins are
//     bit-identical to the reference's.
//   - backprojection: FP32 atomic scatter (global_atomic_add_f32) of
//     float(len * double(v)) (VolData::projectLineIntegral, VolData.cc:373-385).
//   - OSEM: one fused kernel per sub-iteration: walk, line integral, +bias,
//     ratio, replay of the same walk (saved set-up state) for the scatter.
// Element-wise and Gaussian kernels are memory-bound; they keep the
// reference's FP32 operation order (no contraction).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"

using fir::Grid;
using fir::kEpsF;
using fir::Lor;
using fir::LorGeom;
using fir::Walk;
using namespace firi;

// ---------------------------------------------------------------------------
// Context helpers (declared in internal.h)
// ---------------------------------------------------------------------------

namespace firi
{
thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}


unsigned blocks_for(long long n, int block)
{
  return (unsigned)((n + block - 1) / block);
}

// Grid of element-wise kernels: grid-stride loops over at most this many
// blocks (8 per CU on 256 CUs).
unsigned ew_blocks(long long n)
{
  const long long b = (n + kBlock - 1) / kBlock;
  return (unsigned)std::max(1LL, std::min(b, 2048LL));
}

int make_grid(const fir_ctx* ctx, const fir_vol_desc* v, Grid* g)
{
  if (!v) return fail(FIR_E_ARG, "volume descriptor is NULL");
  if (v->nx <= 0 || v->ny <= 0 || v->nz <= 0)
    return fail(FIR_E_ARG,
                "Number of voxels must be greater than zero in each "
                "dimension");
  if (!(v->dx > 0 && v->dy > 0 && v->dz > 0))
    return fail(FIR_E_ARG,
                "Voxel extent must be greater than zero in each dimension");
  if ((long long)v->nx * v->ny * v->nz > 0x7fffffffLL)
    return fail(FIR_E_ARG, "volume has more than 2^31-1 voxels");
  const double ext[3] = {v->dx, v->dy, v->dz};
  const int n[3] = {v->nx, v->ny, v->nz};
  g->low[0] = v->ox - v->dx / 2;
  g->low[1] = v->oy - v->dy / 2;
  g->low[2] = v->oz - (v->nz * v->dz) / 2;
  for (int d = 0; d < 3; ++d)
  {
    g->high[d] = g->low[d] + n[d] * ext[d];
    if (ctx->walk_mode == FIR_WALK_REFERENCE)
    {
      const int t = (int)ext[d];
      if (t <= 0)
        return fail(FIR_E_ARG,
                    "voxel extent %.7f mm truncates to 0 in the reference "
                    "walk (which would not terminate); use the geometric "
                    "walk mode",
                    ext[d]);
      g->step[d] = (double)t;
    }
    else
      g->step[d] = ext[d];
    g->nm1[d] = n[d] - 1;
  }
  g->row = v->nx;
  g->slice = v->nx * v->ny;
  g->max_iter = v->nx + v->ny + v->nz + 8;
  return FIR_OK;
}

int need_lors(const fir_ctx* ctx)
{
  if (!ctx) return fail(FIR_E_ARG, "context is NULL");
  if (!ctx->d_xy) return fail(FIR_E_STATE, "scanner tables not set");
  if (!ctx->proj_set) return fail(FIR_E_STATE, "sinogram geometry not set");
  if (ctx->n_cpr != 2 * ctx->n_views)
    return fail(FIR_E_ARG,
                "Projection must have the same number of crystals per ring "
                "(%d) as the scanner",
                ctx->n_cpr);
  if (ctx->n_slices != 2 * ctx->n_rings - 1)
    return fail(FIR_E_ARG,
                "Projection must have the same number of rings (%d) as the "
                "scanner",
                (ctx->n_slices + 1) / 2);
  return FIR_OK;
}

int check_subsets(const fir_ctx* ctx, int n_subsets)
{
  if (ctx->uneven_subsets && n_subsets > 0 && n_subsets <= ctx->n_views)
    return FIR_OK;
  if (n_subsets <= 0 || ctx->n_views % n_subsets != 0)
    return fail(FIR_E_ARG, "Number of subsets must be a divisor of %d",
                ctx->n_views);
  return FIR_OK;
}

LorGeom lor_geom(const fir_ctx* ctx)
{
  LorGeom g;
  g.xy = ctx->d_xy;
  g.z = ctx->d_z;
  g.pairs = ctx->d_pairs;
  g.n_axial = ctx->d_n_axial;
  g.seg_base = ctx->d_seg_base;
  g.n_views = ctx->n_views;
  g.n_pairs = ctx->n_pairs;
  g.ax_prefix = ctx->d_ax_prefix;
  g.n_seg = ctx->n_seg;
  g.row_mode = ctx->row_mode;
  g.tang_major = ctx->tang_major;
  g.n_tang = ctx->n_tang;
  g.tang_off = ctx->tang_off;
  g.n_cpr = ctx->n_cpr;
  return g;
}

long long n_work(const fir_ctx* ctx, int n_subsets, int subset)
{
  const long long vpr = fir::views_in_subset(ctx->n_views, n_subsets, subset);
  return vpr * ctx->n_pairs * (long long)ctx->n_tang;
}

// Row order of the projector kernels: the (segment, axial) pairs sorted by
// segment then axial (the sinogram's own order), or by the centre slice
// s1 + s2 so that rows processed together cover a thin z-slab of the
// image and their gathers and scatters stay in L2.
std::vector<int4> pair_order(const fir_ctx* c, int order)
{
  std::vector<int4> pairs;
  int flat = 0;
  for (int k = 0; k < c->n_seg; ++k)
    for (int a = 0; a < c->n_axial[k]; ++a, ++flat)
      pairs.push_back(make_int4(k, a, c->slices[flat].x, c->slices[flat].y));
  if (order == FIR_ORDER_ZCENTER)
    std::stable_sort(pairs.begin(), pairs.end(),
                     [&](const int4& u, const int4& v) {
                       const int cu = u.z + u.w, cv = v.z + v.w;
                       if (cu != cv) return cu < cv;
                       const int su = std::abs(u.x - c->seg_off);
                       const int sv = std::abs(v.x - c->seg_off);
                       return su < sv;
                     });
  return pairs;
}

int upload_pairs(fir_ctx* c)
{
  const std::vector<int4> pairs = pair_order(c, c->lor_order);
  const std::vector<int4> bins = pair_order(c, FIR_ORDER_SEGMENT);
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  (void)hipFree(c->d_pairs);
  (void)hipFree(c->d_pairs_bins);
  c->d_pairs = nullptr;
  c->d_pairs_bins = nullptr;
  FIR_HIP_CHECK(hipMalloc(&c->d_pairs, sizeof(int4) * pairs.size()));
  FIR_HIP_CHECK(hipMalloc(&c->d_pairs_bins, sizeof(int4) * bins.size()));
  FIR_HIP_CHECK(hipMemcpy(c->d_pairs, pairs.data(), sizeof(int4) * pairs.size(),
                          hipMemcpyHostToDevice));
  FIR_HIP_CHECK(hipMemcpy(c->d_pairs_bins, bins.data(),
                          sizeof(int4) * bins.size(), hipMemcpyHostToDevice));
  c->n_pairs = (int)pairs.size();
  return FIR_OK;
}

// LorGeom whose pair table is in sinogram (segment, axial) order.
LorGeom lor_geom_bins(const fir_ctx* ctx)
{
  LorGeom g = lor_geom(ctx);
  g.pairs = ctx->d_pairs_bins;
  return g;
}

// Work items of one launch must fit the kernels' 32-bit index decode.
int check_work(long long n)
{
  if (n > 0x7fffffffLL)
    return fail(FIR_E_ARG, "more than 2^31-1 LORs in one launch");
  return FIR_OK;
}

} // namespace firi

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

namespace
{

// Forward projection: one LOR per lane (all bins, or one subset).
__global__ __launch_bounds__(kBlock) void k_forward(Grid g, LorGeom lg,
                                                    const float* __restrict__ img,
                                                    float* __restrict__ sino,
                                                    long long n, int n_subsets,
                                                    int subset)
{
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  Lor L;
  fir::lor_from_index(lg, t, n_subsets, subset, L);
  Walk w;
  float line = 0.0f;
  if (fir::walk_setup(g, L.e1, L.e2, w))
  {
    fir::walk_gather(
      g, w, [&](int idx) { return img[idx]; },
      [&](float v, double len, bool ok) {
        const float nl = (float)((double)line + len * (double)v);
        line = ok ? nl : line;
      });
  }
  sino[L.bin] = line;
}

// Backprojection scatter of one subset (or all LORs); sino == nullptr
// backprojects ones.
__global__ __launch_bounds__(kBlock) void k_backward(Grid g, LorGeom lg,
                                                     const float* __restrict__ sino,
                                                     fir::Scatter sc,
                                                     long long n, int n_subsets,
                                                     int subset)
{
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  Lor L;
  fir::lor_from_index(lg, t, n_subsets, subset, L);
  const float val = sino ? sino[L.bin] : 1.0f;
  if (val == 0.0f) return; // adding float(len * 0) changes nothing
  Walk w;
  if (!fir::walk_setup(g, L.e1, L.e2, w)) return;
  const double dv = (double)val;
  float* img = fir::scatter_target(sc);
  fir::walk_run(g, w, [&](int idx, double len) {
    unsafeAtomicAdd(img + idx, (float)(len * dv));
  });
}

// Fused OSEM projector pass (reconAlgos.cc:126-158).
__global__ __launch_bounds__(kBlock) void k_osem_project(
  Grid g, LorGeom lg, const float* __restrict__ x, const float* __restrict__ y,
  const float* __restrict__ bias, fir::Scatter sc, long long n,
  int n_subsets, int subset)
{
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  Lor L;
  fir::lor_from_index(lg, t, n_subsets, subset, L);
  Walk w;
  if (!fir::walk_setup(g, L.e1, L.e2, w)) return; // empty path: nothing to scatter
  float line = 0.0f;
  fir::walk_gather(
    g, w, [&](int idx) { return x[idx]; },
    [&](float v, double len, bool ok) {
      const float nl = (float)((double)line + len * (double)v);
      line = ok ? nl : line;
    });
  if (bias) line += bias[L.bin];
  if (!(line > kEpsF)) return;
  const float ratio = y[L.bin] / line;
  if (ratio == 0.0f) return;
  const double dr = (double)ratio;
  float* bp = fir::scatter_target(sc);
  fir::walk_run(g, w, [&](int idx, double len) {
    unsafeAtomicAdd(bp + idx, (float)(len * dr));
  });
}

// ---------------------------------------------------------------------------
// Family form of the walked OSEM projector
// ---------------------------------------------------------------------------
// The LORs of one (view, tangential bin) -- every ring pair of one line of
// the xy-plane -- cross the same pixels, in the same order: the family's
// sheet is those pixels times all slices (at most (nx + ny - 1) * nz
// voxels). One block per family stages the sheet of x in LDS, walks each
// LOR (one per thread) gathering from LDS (the same values in the same
// order: the line integrals are the same bit for bit), then reuses the LDS
// as the family's backprojection accumulator (LDS atomics) and adds its
// nonzero voxels to the image: one global atomic per touched sheet voxel
// instead of one per path element.

constexpr int kFamSheetFloats = 16384; // 64 KB of LDS
constexpr int kFamBlock = 1024;
constexpr int kFamMaxPx = 1024;

// OSEM: the projector pass (line integrals, ratio y / (line + bias),
// backprojection of the ratios). Otherwise the backprojection of sinogram y
// (of ones if y is NULL), as k_backward.
template <bool OSEM>
__global__ __launch_bounds__(kFamBlock) void k_project_fam(
  Grid g, LorGeom lg, const float* __restrict__ x, const float* __restrict__ y,
  const float* __restrict__ bias, fir::Scatter sc, int n_subsets, int subset,
  int n_tang, int n_pairs, int nz)
{
  __shared__ float sh[kFamSheetFloats];
  __shared__ int px[kFamMaxPx];
  __shared__ int n_px;
  const int view = (int)blockIdx.y * n_subsets + subset;
  const int tang = (int)blockIdx.x;
  const int nxy = g.slice;
  if (threadIdx.x == 0)
  {
    // the family's pixels: the walk of its line at the volume's mid-depth
    Lor L;
    fir::lor_from_index(lg, (long long)tang * n_pairs, lg.n_views, view, L);
    const double zc = 0.5 * (g.low[2] + g.high[2]);
    const double e1[3] = {L.e1[0], L.e1[1], zc};
    const double e2[3] = {L.e2[0], L.e2[1], zc};
    Walk w;
    int n = 0;
    if (fir::walk_setup(g, e1, e2, w))
    {
      int last = -1;
      fir::walk_run(g, w, [&](int idx, double) {
        const int p = idx % nxy;
        if (p != last && n < kFamMaxPx)
        {
          px[n++] = p;
          last = p;
        }
      });
    }
    n_px = n;
  }
  __syncthreads();
  const int L = n_px;
  if (L == 0 || L * nz > kFamSheetFloats) return; // host checks the bound
  const int n_st = L * nz;
  if (OSEM)
  {
    for (int i = threadIdx.x; i < n_st; i += kFamBlock)
    {
      const int z = i / L, p = i - z * L;
      sh[p * nz + z] = x[(long long)z * nxy + px[p]];
    }
    __syncthreads();
  }

  // Forward: one LOR per thread, the walk's pixels followed along the
  // sheet (rank r), values from LDS.
  const int pair = threadIdx.x;
  float ratio = 0.0f;
  Walk w;
  bool hit = false;
  if (pair < n_pairs)
  {
    Lor Lr;
    fir::lor_from_index(lg, (long long)tang * n_pairs + pair, lg.n_views, view,
                        Lr);
    hit = fir::walk_setup(g, Lr.e1, Lr.e2, w);
    if (hit && !OSEM) ratio = y ? y[Lr.bin] : 1.0f;
    if (hit && OSEM)
    {
      float line = 0.0f;
      int r = 0;
      Walk v = w;
      double prev = v.prev;
      for (int it = 0; prev < v.a_hi && it < g.max_iter; ++it)
      {
        const double next =
          fir::dmin(v.a_hi, fir::dmin(v.a[0], fir::dmin(v.a[1], v.a[2])));
        if ((unsigned)v.pos[0] <= (unsigned)g.nm1[0] &&
            (unsigned)v.pos[1] <= (unsigned)g.nm1[1] &&
            (unsigned)v.pos[2] <= (unsigned)g.nm1[2])
        {
          const int P = v.pos[0] + v.pos[1] * g.row;
          while (r < L && px[r] != P) ++r;
          const float val = r < L ? sh[r * nz + v.pos[2]]
                                  : x[P + (long long)v.pos[2] * nxy];
          if (r >= L) r = 0; // not on the sheet (not expected): restart
          line = (float)((double)line + (next - prev) * v.d12 * (double)val);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
          if (fir::dabs(v.a[d] - next) < fir::kEps)
          {
            v.a[d] += v.da[d];
            v.pos[d] += v.dir[d];
          }
        prev = next;
      }
      if (bias) line += bias[Lr.bin];
      if (line > kEpsF) ratio = y[Lr.bin] / line;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n_st; i += kFamBlock) sh[i] = 0.0f;
  __syncthreads();

  // Backprojection into the sheet.
  float* bp = fir::scatter_target(sc);
  if (hit && ratio != 0.0f)
  {
    const double dr = (double)ratio;
    int r = 0;
    double prev = w.prev;
    for (int it = 0; prev < w.a_hi && it < g.max_iter; ++it)
    {
      const double next =
        fir::dmin(w.a_hi, fir::dmin(w.a[0], fir::dmin(w.a[1], w.a[2])));
      if ((unsigned)w.pos[0] <= (unsigned)g.nm1[0] &&
          (unsigned)w.pos[1] <= (unsigned)g.nm1[1] &&
          (unsigned)w.pos[2] <= (unsigned)g.nm1[2])
      {
        const int P = w.pos[0] + w.pos[1] * g.row;
        while (r < L && px[r] != P) ++r;
        const float add = (float)((next - prev) * w.d12 * dr);
        if (r < L)
          atomicAdd(sh + r * nz + w.pos[2], add);
        else
        {
          unsafeAtomicAdd(bp + P + (long long)w.pos[2] * nxy, add);
          r = 0;
        }
      }
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (fir::dabs(w.a[d] - next) < fir::kEps)
        {
          w.a[d] += w.da[d];
          w.pos[d] += w.dir[d];
        }
      prev = next;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n_st; i += kFamBlock)
  {
    const int z = i / L, p = i - z * L;
    const float a = sh[p * nz + z];
    if (a != 0.0f) unsafeAtomicAdd(bp + (long long)z * nxy + px[p], a);
  }
}

// Whether the walked projectors can run family by family on this geometry.
bool family_ok(const fir_ctx* c, const fir_vol_desc* v)
{
  return c->walk_sheets && c->n_pairs <= kFamBlock &&
    (long long)(v->nx + v->ny - 1) * v->nz <= kFamSheetFloats &&
    v->nx + v->ny <= kFamMaxPx;
}

// One block per (view, tangential bin) family of the subset (n_subsets 0:
// every view).
template <bool OSEM>
void launch_family(fir_ctx* c, const Grid& g, const fir_vol_desc* v,
                   const float* x, const float* y, const float* bias,
                   const fir::Scatter& sc, int n_subsets, int subset)
{
  LorGeom lg = lor_geom(c);
  lg.row_mode = fir::ROWS_VIEW_MAJOR;
  lg.tang_major = 1;
  const int S = n_subsets > 0 ? n_subsets : 1;
  const int s = n_subsets > 0 ? subset : 0;
  const int vpr = fir::views_in_subset(c->n_views, S, s);
  hipLaunchKernelGGL(k_project_fam<OSEM>, dim3(c->n_tang, vpr),
                     dim3(kFamBlock), 0, c->stream, g, lg, x, y, bias, sc, S,
                     s, c->n_tang, c->n_pairs, v->nz);
}

// Poisson negative log-likelihood L = sum_b (l_b - y_b log l_b), with
// l = A x + r, and its gradient A^T (1 - y / l), fused per LOR: walk,
// line integral, weight, replay of the walk for the scatter. Bins with
// l <= 0 contribute l to L and weight 1 (their y log l term is dropped).
// L is reduced per block in double and added to *loss.
__global__ __launch_bounds__(kBlock) void k_poisson_grad(
  Grid g, LorGeom lg, const float* __restrict__ x, const float* __restrict__ y,
  const float* __restrict__ r, fir::Scatter sc,
  double* __restrict__ loss, long long n)
{
  __shared__ double part[kBlock];
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  double term = 0.0;
  if (t < n)
  {
    Lor L;
    fir::lor_from_index(lg, t, 0, 0, L);
    Walk w;
    const bool hit = fir::walk_setup(g, L.e1, L.e2, w);
    float line = 0.0f;
    if (hit)
      fir::walk_gather(
        g, w, [&](int idx) { return x[idx]; },
        [&](float v, double len, bool ok) {
          const float nl = (float)((double)line + len * (double)v);
          line = ok ? nl : line;
        });
    const float l = r ? line + r[L.bin] : line;
    const float yb = y[L.bin];
    float wgt;
    if (l > 0.0f)
    {
      term = (double)l - (double)yb * log((double)l);
      wgt = 1.0f - yb / l;
    }
    else
    {
      term = (double)l;
      wgt = 1.0f;
    }
    if (hit && wgt != 0.0f)
    {
      const double dw = (double)wgt;
      float* grad = fir::scatter_target(sc);
      fir::walk_run(g, w, [&](int idx, double len) {
        unsafeAtomicAdd(grad + idx, (float)(len * dw));
      });
    }
  }
  part[threadIdx.x] = term;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1)
  {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) unsafeAtomicAdd(loss, part[0]);
}

__global__ void k_zero_f64(double* p) { *p = 0.0; }

// Batched OSEM projector pass: NB reconstructions that share the image
// geometry and subset schedule. One walk per LOR serves all NB images: the
// walk (FP64 ALU work) is paid once, each element gathers NB voxels and
// scatters NB ratios.
constexpr int kMaxBatch = 8;

struct BatchPtrs
{
  const float* x[kMaxBatch];
  const float* y[kMaxBatch];
  const float* bias[kMaxBatch];
  float* bp[kMaxBatch];
  long long copy_stride; // per-XCD copies: image b's copy c at bp[b] + c * stride
  int copies;
};

template <int NB>
__global__ __launch_bounds__(kBlock) void k_osem_project_batch(
  Grid g, LorGeom lg, BatchPtrs p, long long n, int n_subsets, int subset)
{
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  Lor L;
  fir::lor_from_index(lg, t, n_subsets, subset, L);
  Walk w;
  if (!fir::walk_setup(g, L.e1, L.e2, w)) return;
  float line[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) line[b] = 0.0f;
  struct Vals
  {
    float v[NB];
  };
  fir::walk_gather(
    g, w,
    [&](int idx) {
      Vals r;
#pragma unroll
      for (int b = 0; b < NB; ++b) r.v[b] = p.x[b][idx];
      return r;
    },
    [&](const Vals& r, double len, bool ok) {
#pragma unroll
      for (int b = 0; b < NB; ++b)
      {
        const float nl = (float)((double)line[b] + len * (double)r.v[b]);
        line[b] = ok ? nl : line[b];
      }
    });
  double ratio[NB];
  bool any = false;
#pragma unroll
  for (int b = 0; b < NB; ++b)
  {
    float l = line[b];
    if (p.bias[b]) l += p.bias[b][L.bin];
    const float r = l > kEpsF ? p.y[b][L.bin] / l : 0.0f;
    ratio[b] = (double)r;
    any = any || r != 0.0f;
  }
  if (!any) return;
  const long long off =
    p.copies > 1 ? (long long)(__builtin_amdgcn_s_getreg((20) | (3 << 11)) %
                               (unsigned)p.copies) * p.copy_stride
                 : 0;
  fir::walk_run(g, w, [&](int idx, double len) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if (ratio[b] != 0.0)
        unsafeAtomicAdd(p.bp[b] + off + idx, (float)(len * ratio[b]));
  });
}

// Path dump for parity checks.
__global__ __launch_bounds__(kBlock) void k_path_dump(Grid g, LorGeom lg,
                                                      const int* ax_prefix,
                                                      int n_seg,
                                                      long long bin0,
                                                      long long n, int maxp,
                                                      int* __restrict__ counts,
                                                      int* __restrict__ idx,
                                                      double* __restrict__ len)
{
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  Lor L;
  fir::lor_from_bin(lg, ax_prefix, n_seg, bin0 + t, L);
  Walk w;
  if (!fir::walk_setup(g, L.e1, L.e2, w))
  {
    counts[t] = -1;
    return;
  }
  int c = 0;
  int* oi = idx + t * (long long)maxp;
  double* ol = len + t * (long long)maxp;
  fir::walk_run(g, w, [&](int i, double l) {
    if (c < maxp)
    {
      oi[c] = i;
      ol[c] = l;
    }
    ++c;
  });
  counts[t] = c;
}

// Element count of each LOR's path, reduced per block into one 64-bit
// counter.
__global__ __launch_bounds__(kBlock) void k_path_count(
  Grid g, LorGeom lg, long long n, int n_subsets, int subset,
  unsigned long long* __restrict__ total)
{
  __shared__ unsigned long long part[kBlock];
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  unsigned long long c = 0;
  if (t < n)
  {
    Lor L;
    fir::lor_from_index(lg, t, n_subsets, subset, L);
    Walk w;
    if (fir::walk_setup(g, L.e1, L.e2, w))
      fir::walk_run(g, w, [&](int, double) { ++c; });
  }
  part[threadIdx.x] = c;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1)
  {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(total, part[0]);
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ULL; }

// dst += sum of the private copies; the copies are cleared for next use.
__global__ void k_reduce_copies(float* __restrict__ dst,
                                float* __restrict__ copies, long long n,
                                int n_copies, long long stride)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
  {
    float acc = dst[i];
    for (int c = 0; c < n_copies; ++c)
    {
      acc += copies[c * stride + i];
      copies[c * stride + i] = 0.0f;
    }
    dst[i] = acc;
  }
}

__global__ void k_fill(float* __restrict__ a, long long n, float v)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    a[i] = v;
}

__global__ void k_mul_clamped(float* __restrict__ a, const float* __restrict__ b,
                              long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
  {
    const float x = a[i], y = b[i];
    a[i] = (x > kEpsF && y > kEpsF) ? x * y : 0.0f;
  }
}

__global__ void k_div_clamped(float* __restrict__ a, const float* __restrict__ b,
                              long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
  {
    const float x = a[i], y = b[i];
    a[i] = (x > kEpsF && y > kEpsF) ? x / y : 0.0f;
  }
}

__global__ void k_exp_clamped(float* __restrict__ a, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
  {
    const float x = a[i];
    a[i] = x > kEpsF ? expf(x) : 1.0f;
  }
}

__global__ void k_hu_to_mu(float* __restrict__ a, long long n)
{
  const float water = 0.0096f;
  const float s1 = water / 1000.0f;
  const float s2 = (0.015f - water) / 1000.0f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
  {
    float x = a[i];
    if (x <= -1000.0f)
      x = 0.0f;
    else if (x <= 0.0f)
    {
      x *= s1;
      x += water;
    }
    else
    {
      x *= s2;
      x += water;
    }
    a[i] = x;
  }
}

// Radial distance test of operations::cutCircle (operations.cc:206-214),
// computed in double as the reference does.
struct CutGeom
{
  double dx, dy, half_w, half_h;
  int nx, ny;
};

__device__ __forceinline__ double radius_of(const CutGeom& c, int i, int j)
{
  const double px = i * c.dx + c.dx / 2.0 - c.half_w;
  const double py = j * c.dy + c.dy / 2.0 - c.half_h;
  return sqrt(px * px + py * py);
}

CutGeom cut_geom(const fir_vol_desc* v)
{
  CutGeom c;
  c.dx = v->dx;
  c.dy = v->dy;
  c.half_w = (v->nx * v->dx) / 2.0;
  c.half_h = (v->ny * v->dy) / 2.0;
  c.nx = v->nx;
  c.ny = v->ny;
  return c;
}

__global__ void k_cut_circle(float* __restrict__ img, long long n, CutGeom c,
                             float cut)
{
  const long long plane = (long long)c.nx * c.ny;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
  {
    const long long r = i % plane;
    const int y = (int)(r / c.nx);
    const int x = (int)(r - (long long)y * c.nx);
    if (radius_of(c, x, y) > cut) img[i] = 0.0f;
  }
}

// OSEM update: bp /= sens (clamped), x *= bp (clamped), optional cut,
// optional bp reset (VolData.cc:305-343, reconAlgos.cc:161-186).
__global__ void k_osem_update(float* __restrict__ x, float* __restrict__ bp,
                              const float* __restrict__ sens, long long n,
                              CutGeom c, float cut, int apply_cut, int reset)
{
  const long long plane = (long long)c.nx * c.ny;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
  {
    float b = bp[i];
    const float s = sens[i];
    b = (b > kEpsF && s > kEpsF) ? b / s : 0.0f;
    float v = x[i];
    v = (v > kEpsF && b > kEpsF) ? v * b : 0.0f;
    if (apply_cut)
    {
      const long long r = i % plane;
      const int yy = (int)(r / c.nx);
      const int xx = (int)(r - (long long)yy * c.nx);
      if (radius_of(c, xx, yy) > cut) v = 0.0f;
    }
    x[i] = v;
    bp[i] = reset ? 0.0f : b;
  }
}

// One axis of the separable Gaussian (operations.cc:83-156): out = sum(k*v) /
// sum(k) over in-bounds taps, accumulated in tap order in FP32.
// axis 0: x, 1: y, 2: z. For axis 2 with `restore`, voxels at radius >=
// restore_r take keep[] (edge restore, operations.cc:158-185); with
// cut > 0 the circular cut follows (operations.cc:194-220).
constexpr int kGaussX = 32, kGaussY = kBlock / kGaussX;
constexpr int kGaussMaxH = 7; // taps per side loaded up front (15 taps)
constexpr int kGaussTapFloats = 192; // X, Y, Z taps, 64 floats each

// Sum over the in-bounds taps of one axis, in tap order: s = sum(k * v),
// w = sum(k) (operations.cc:83-156). The taps' loads come first (clamped to
// the image, so always in bounds) so that they are in flight together.
__device__ __forceinline__ void gauss_sum(const float* __restrict__ in,
                                          long long i, int coord, int lim,
                                          long long stride,
                                          const float* __restrict__ taps,
                                          int h, float* s_out, float* w_out)
{
  float s = 0.0f, w = 0.0f;
  if (h <= kGaussMaxH)
  {
    float t[2 * kGaussMaxH + 1];
#pragma unroll
    for (int q = 0; q <= 2 * kGaussMaxH; ++q)
      if (q <= 2 * h)
      {
        int cq = coord + q - h;
        cq = cq < 0 ? 0 : (cq >= lim ? lim - 1 : cq);
        t[q] = in[i + (long long)(cq - coord) * stride];
      }
#pragma unroll
    for (int q = 0; q <= 2 * kGaussMaxH; ++q)
      if (q <= 2 * h)
      {
        const int cq = coord + q - h;
        if (cq >= 0 && cq < lim)
        {
          const float kv = taps[q];
          s += kv * t[q];
          w += kv;
        }
      }
  }
  else
    for (int q = -h; q <= h; ++q)
    {
      const int cq = coord + q;
      if (cq >= 0 && cq < lim)
      {
        const float kv = taps[h + q];
        s += kv * in[i + q * stride];
        w += kv;
      }
    }
  *s_out = s;
  *w_out = w;
}

template <int AXIS, bool EPILOGUE>
__global__ __launch_bounds__(kBlock) void k_gauss(const float* __restrict__ in,
                                                  float* __restrict__ out,
                                                  const float* __restrict__ taps,
                                                  int h, int nx, int ny, int nz,
                                                  CutGeom c, int restore,
                                                  double restore_r, float cut)
{
  // One thread per voxel, its (x, y, z) taken from the 3D grid (no index
  // divisions): blocks of kGaussX x kGaussY voxels of one xy-plane.
  const int ii = blockIdx.x * kGaussX + (int)threadIdx.x % kGaussX;
  const int j = blockIdx.y * kGaussY + (int)threadIdx.x / kGaussX;
  const int k = blockIdx.z;
  const long long plane = (long long)nx * ny;
  if (ii < nx && j < ny)
  {
    const long long i = (long long)k * plane + (long long)j * nx + ii;
    const int coord = AXIS == 0 ? ii : (AXIS == 1 ? j : k);
    const int lim = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
    const long long stride = AXIS == 0 ? 1 : (AXIS == 1 ? nx : plane);
    float s, w;
    gauss_sum(in, i, coord, lim, stride, taps, h, &s, &w);
    float o;
    if (w > 0.0f)
      o = s / w;
    else
      o = out[i];
    if (EPILOGUE)
    {
      const double rad = radius_of(c, ii, j);
      if (restore && rad >= restore_r) o = out[i];
      if (cut > 0.0f && rad > cut) o = 0.0f;
    }
    out[i] = o;
  }
}

// The Gaussian's passes for frames whose rows are whole float4s (nx a
// multiple of 4, 16-byte aligned frames) and half-widths up to kGaussHV,
// H the pass's half-width: each thread computes four outputs from aligned
// float4 loads kept in registers. X: 4 consecutive x of one row from 3
// loads; Y (Z): 4 x by 4 consecutive y (z) from 4 + 2H loads; instead of
// 2H + 1 scalar gathers per output. Every output is summed as gauss_sum
// sums it (the same in-range taps in the same order, then s / w), so the
// results are the same bit for bit. The window's centre tap (1) is always
// in range, so w > 0 and no output falls back to its old value.
constexpr int kGaussHV = 4;
constexpr long long kGaussVecMinVoxels = 1LL << 19;

__device__ __forceinline__ float4 ld4(const float* p)
{
  return *reinterpret_cast<const float4*>(p);
}

__device__ __forceinline__ float comp4(const float4& a, int c)
{
  return c == 0 ? a.x : (c == 1 ? a.y : (c == 2 ? a.z : a.w));
}

template <int H>
__device__ __forceinline__ float gauss_out(const float* v, const float* tp,
                                           int coord, int lim)
{
  float s = 0.0f, w = 0.0f;
#pragma unroll
  for (int q = 0; q <= 2 * H; ++q)
  {
    const int cq = coord + q - H;
    if (cq >= 0 && cq < lim)
    {
      s += tp[q] * v[q];
      w += tp[q];
    }
  }
  return s / w;
}

// X pass of unit u (x-quad xq of row z * ny + y).
template <int H>
__device__ __forceinline__ void gauss_vx_unit(const float* __restrict__ in,
                                              float* __restrict__ out,
                                              const float* tp, int nx,
                                              long long row, int xq)
{
  static_assert(H <= kGaussHV, "window of three float4");
  const float* r = in + row * nx;
  const int x0 = 4 * xq;
  const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4 a = x0 >= 4 ? ld4(r + x0 - 4) : z4;
  const float4 b = ld4(r + x0);
  const float4 c = x0 + 4 < nx ? ld4(r + x0 + 4) : z4;
  const float v[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w,
                       c.x, c.y, c.z, c.w};
  float o[4];
#pragma unroll
  for (int g = 0; g < 4; ++g)
    o[g] = gauss_out<H>(v + 4 + g - H, tp, x0 + g, nx);
  *reinterpret_cast<float4*>(out + row * nx + x0) =
    make_float4(o[0], o[1], o[2], o[3]);
}

// Y or Z pass of the 4 x 4 patch at x-quad xq, axis positions a0..a0 + 3
// (a = y or z), `base` the patch's first column (the other coordinate
// applied), `stride` the axis stride. With EPI (the Z pass) the edge
// restore and the cut of k_gauss's epilogue, per x of the patch's row y.
template <int H, bool EPI>
__device__ __forceinline__ void gauss_vyz_unit(
  const float* __restrict__ in, float* __restrict__ out, const float* tp,
  long long base, long long stride, int a0, int lim, int x0, int y,
  const CutGeom& cg, int restore, double restore_r, float cut)
{
  const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 win[4 + 2 * H];
#pragma unroll
  for (int k = 0; k < 4 + 2 * H; ++k)
  {
    const int aa = a0 - H + k;
    win[k] = aa >= 0 && aa < lim ? ld4(in + base + aa * stride) : z4;
  }
  bool keep[4] = {false, false, false, false};
  bool zero[4] = {false, false, false, false};
  if (EPI)
#pragma unroll
    for (int c = 0; c < 4; ++c)
    {
      const double rad = radius_of(cg, x0 + c, y);
      keep[c] = restore && rad >= restore_r;
      zero[c] = cut > 0.0f && rad > cut;
    }
  const bool any_keep = keep[0] || keep[1] || keep[2] || keep[3];
#pragma unroll
  for (int g = 0; g < 4; ++g)
  {
    if (a0 + g >= lim) break;
    float o[4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
    {
      float v[2 * H + 1];
#pragma unroll
      for (int q = 0; q <= 2 * H; ++q) v[q] = comp4(win[g + q], c);
      o[c] = gauss_out<H>(v, tp, a0 + g, lim);
    }
    float* dst = out + base + (long long)(a0 + g) * stride;
    if (EPI)
    {
      if (any_keep)
      {
        const float4 old = ld4(dst);
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (keep[c]) o[c] = comp4(old, c);
      }
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (zero[c]) o[c] = 0.0f;
    }
    *reinterpret_cast<float4*>(dst) = make_float4(o[0], o[1], o[2], o[3]);
  }
}

template <int H>
__device__ __forceinline__ void gauss_taps_h(const float* __restrict__ taps,
                                             float* tp)
{
#pragma unroll
  for (int q = 0; q <= 2 * H; ++q) tp[q] = taps[q];
}

// Units per frame of the vectorised passes: X (x-quad, row), Y (x-quad,
// y-quad, z), Z (x-quad, y, z-quad).
__host__ __device__ __forceinline__ long long gauss_units(int axis, int nx,
                                                          int ny, int nz)
{
  const long long nxq = nx / 4;
  return axis == 0 ? nxq * ny * nz
                   : (axis == 1 ? nxq * ((ny + 3) / 4) * nz
                                : nxq * ny * ((nz + 3) / 4));
}

// One unit of pass AXIS over frame in -> out.
template <int AXIS, int H>
__device__ __forceinline__ void gauss_v_unit(
  const float* __restrict__ in, float* __restrict__ out, const float* tp,
  int nx, int ny, int nz, long long u, const CutGeom& cg, int restore,
  double restore_r, float cut)
{
  const int nxq = nx / 4;
  const int xq = (int)(u % nxq);
  const long long r = u / nxq;
  const long long plane = (long long)nx * ny;
  if (AXIS == 0)
    gauss_vx_unit<H>(in, out, tp, nx, r, xq);
  else if (AXIS == 1)
  {
    const int nyq = (ny + 3) / 4;
    const int yq = (int)(r % nyq), z = (int)(r / nyq);
    gauss_vyz_unit<H, false>(in, out, tp, z * plane + 4 * xq, nx, 4 * yq,
                             ny, 4 * xq, 0, cg, 0, 0.0, 0.0f);
  }
  else
  {
    const int y = (int)(r % ny), zq = (int)(r / ny);
    gauss_vyz_unit<H, true>(in, out, tp, (long long)y * nx + 4 * xq, plane,
                            4 * zq, nz, 4 * xq, y, cg, restore, restore_r,
                            cut);
  }
}

template <int AXIS, int H>
__global__ __launch_bounds__(kBlock) void k_gauss_v(
  const float* __restrict__ in, float* __restrict__ out,
  const float* __restrict__ taps, int nx, int ny, int nz, CutGeom cg,
  int restore, double restore_r, float cut)
{
  const long long u = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (u >= gauss_units(AXIS, nx, ny, nz)) return;
  float tp[2 * H + 1];
  gauss_taps_h<H>(taps, tp);
  gauss_v_unit<AXIS, H>(in, out, tp, nx, ny, nz, u, cg, restore, restore_r,
                        cut);
}

// Whether a frame takes the vectorised Gaussian passes (k_gauss_v).
bool gauss_vec_ok(int nx, const int* h)
{
  return nx % 4 == 0 && h[0] <= kGaussHV && h[1] <= kGaussHV &&
    h[2] <= kGaussHV && h[0] >= 0 && h[1] >= 0 && h[2] >= 0;
}

template <int AXIS>
void launch_gauss_v_axis(hipStream_t st, int h, const float* in, float* out,
                         const float* taps, const fir_vol_desc* v,
                         const CutGeom& cg, int restore, double rr, float cut)
{
  const long long n = gauss_units(AXIS, v->nx, v->ny, v->nz);
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
#define FIR_GAUSS_V(HH)                                                     \
  hipLaunchKernelGGL((k_gauss_v<AXIS, HH>), grid, dim3(kBlock), 0, st, in, \
                     out, taps, v->nx, v->ny, v->nz, cg, restore, rr, cut)
  switch (h)
  {
  case 0: FIR_GAUSS_V(0); break;
  case 1: FIR_GAUSS_V(1); break;
  case 2: FIR_GAUSS_V(2); break;
  case 3: FIR_GAUSS_V(3); break;
  default: FIR_GAUSS_V(4); break;
  }
#undef FIR_GAUSS_V
}

void launch_gauss_v(hipStream_t st, int axis, int h, const float* in,
                    float* out, const float* taps, const fir_vol_desc* v,
                    const CutGeom& cg, int restore, double rr, float cut)
{
  if (axis == 0)
    launch_gauss_v_axis<0>(st, h, in, out, taps, v, cg, restore, rr, cut);
  else if (axis == 1)
    launch_gauss_v_axis<1>(st, h, in, out, taps, v, cg, restore, rr, cut);
  else
    launch_gauss_v_axis<2>(st, h, in, out, taps, v, cg, restore, rr, cut);
}

// Multi-job OSEM steps: the element-wise phases of many images (different
// grids) in one launch each. A block takes one tile (kGaussX x kGaussY
// voxels of one slice) of one image; its image is found by a binary search
// over the images' first tiles. The arithmetic per voxel is k_osem_update's
// and k_gauss's.
__device__ __forceinline__ int step_find_image(const StepImage* imgs, int n,
                                               int tile)
{
  int lo = 0, hi = n - 1; // last image whose first tile is <= tile
  while (lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    if (imgs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

struct StepVoxel
{
  int ii, j, k;
  long long i;
  bool in;
};

__device__ __forceinline__ StepVoxel step_voxel(const StepImage& m, int tile)
{
  const int t = tile - m.tile0;
  const int bx = (m.nx + kGaussX - 1) / kGaussX;
  const int by = (m.ny + kGaussY - 1) / kGaussY;
  const int k = t / (bx * by);
  const int r = t - k * bx * by;
  const int tyy = r / bx, txx = r - tyy * bx;
  StepVoxel v;
  v.ii = txx * kGaussX + (int)threadIdx.x % kGaussX;
  v.j = tyy * kGaussY + (int)threadIdx.x / kGaussX;
  v.k = k;
  v.in = v.ii < m.nx && v.j < m.ny;
  v.i = ((long long)k * m.ny + v.j) * m.nx + v.ii;
  return v;
}

__device__ __forceinline__ CutGeom step_cut_geom(const StepImage& m)
{
  CutGeom c;
  c.dx = m.dx;
  c.dy = m.dy;
  c.half_w = (m.nx * m.dx) / 2.0;
  c.half_h = (m.ny * m.dy) / 2.0;
  c.nx = m.nx;
  c.ny = m.ny;
  return c;
}

// x = x * clamp(bp / sens), bp = 0 (k_osem_update with reset), the cut when
// apply_cut.
__global__ __launch_bounds__(kBlock) void k_step_update(
  const StepImage* __restrict__ imgs, int n_imgs, float cut, int apply_cut)
{
  const int tile = blockIdx.x;
  const StepImage& m = imgs[step_find_image(imgs, n_imgs, tile)];
  const StepVoxel v = step_voxel(m, tile);
  if (!v.in) return;
  float b = m.bp[v.i];
  const float s = m.sens[v.i];
  b = (b > kEpsF && s > kEpsF) ? b / s : 0.0f;
  float x = m.x[v.i];
  x = (x > kEpsF && b > kEpsF) ? x * b : 0.0f;
  if (apply_cut && radius_of(step_cut_geom(m), v.ii, v.j) > cut) x = 0.0f;
  m.x[v.i] = x;
  m.bp[v.i] = 0.0f;
}

// One axis of the Gaussian for every image (k_gauss): X: x -> t1,
// Y: t1 -> t2, Z: t2 -> x with the edge restore and the cut.
template <int AXIS>
__global__ __launch_bounds__(kBlock) void k_step_gauss(
  const StepImage* __restrict__ imgs, int n_imgs, int restore,
  double restore_r, float cut)
{
  const int tile = blockIdx.x;
  const StepImage& m = imgs[step_find_image(imgs, n_imgs, tile)];
  const StepVoxel v = step_voxel(m, tile);
  if (!v.in) return;
  const float* in = AXIS == 0 ? m.x : (AXIS == 1 ? m.t1 : m.t2);
  float* out = AXIS == 0 ? m.t1 : (AXIS == 1 ? m.t2 : m.x);
  const long long plane = (long long)m.nx * m.ny;
  const int coord = AXIS == 0 ? v.ii : (AXIS == 1 ? v.j : v.k);
  const int lim = AXIS == 0 ? m.nx : (AXIS == 1 ? m.ny : m.nz);
  const long long stride = AXIS == 0 ? 1 : (AXIS == 1 ? m.nx : plane);
  float s, w;
  gauss_sum(in, v.i, coord, lim, stride, m.taps + 64 * AXIS, m.h[AXIS], &s,
            &w);
  float o;
  if (w > 0.0f)
    o = s / w;
  else
    o = out[v.i];
  if (AXIS == 2)
  {
    const double rad = radius_of(step_cut_geom(m), v.ii, v.j);
    if (restore && rad >= restore_r) o = out[v.i];
    if (cut > 0.0f && rad > cut) o = 0.0f;
  }
  out[v.i] = o;
}

// One vectorised pass of the Gaussian (k_gauss_v) for every image: a block
// takes kBlock units of one image, found by a binary search over the
// images' first blocks of this pass.
template <int AXIS, int H>
__global__ __launch_bounds__(kBlock) void k_step_gauss_v(
  const StepImage* __restrict__ imgs, int n_imgs, int restore,
  double restore_r, float cut)
{
  const int b = blockIdx.x;
  int lo = 0, hi = n_imgs - 1; // last image whose first block is <= b
  while (lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    if (imgs[mid].vt0[AXIS] <= b) lo = mid;
    else hi = mid - 1;
  }
  const StepImage& m = imgs[lo];
  const long long u = (long long)(b - m.vt0[AXIS]) * kBlock + threadIdx.x;
  if (u >= gauss_units(AXIS, m.nx, m.ny, m.nz)) return;
  const float* in = AXIS == 0 ? m.x : (AXIS == 1 ? m.t1 : m.t2);
  float* out = AXIS == 0 ? m.t1 : (AXIS == 1 ? m.t2 : m.x);
  float tp[2 * H + 1];
  gauss_taps_h<H>(m.taps + 64 * AXIS, tp);
  gauss_v_unit<AXIS, H>(in, out, tp, m.nx, m.ny, m.nz, u, step_cut_geom(m),
                        AXIS == 2 ? restore : 0, restore_r,
                        AXIS == 2 ? cut : 0.0f);
}

// Taps passed by value (kernel arguments), so that no host-to-device copy
// and no device allocation is needed to place them.
struct GaussTaps
{
  float v[kGaussTapFloats];
};

__global__ void k_store_taps(GaussTaps t, float* __restrict__ dst)
{
  const int i = threadIdx.x;
  if (i < kGaussTapFloats) dst[i] = t.v[i];
}

} // namespace

// ---------------------------------------------------------------------------
// Static scratch of fir_convolve_noalloc: the taps, then the two intermediate
// images of the separable Gaussian, for frames of up to kNoAllocMaxVoxels
// voxels. Every device the library runs on holds one copy (about 64 MB) from
// the module's load until the process ends. Host code never passes this
// name to a kernel or a copy: it takes the copy's device address from
// hipGetSymbolAddress once per device (noalloc_setup).
// ---------------------------------------------------------------------------
namespace
{
constexpr long long kNoAllocMaxVoxels = 1LL << 23;
constexpr long long kNoAllocTapSlot = 256; // floats before the images
constexpr long long kNoAllocFloats = kNoAllocTapSlot + 2 * kNoAllocMaxVoxels;
constexpr int kNoAllocMaxDevices = 64;
} // namespace

static __device__ __align__(256) float scratch[kNoAllocFloats];

namespace
{
std::mutex g_noalloc_mu;
float* g_noalloc_ptr[kNoAllocMaxDevices] = {};

// Device address of this device's copy of `scratch`, looked up once per
// device; `device` must be current.
int noalloc_setup(int device)
{
  if (device < 0 || device >= kNoAllocMaxDevices)
    return fail(FIR_E_ARG, "device %d beyond the %d the static scratch "
                           "tracks", device, kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return FIR_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(FIR_E_HIP, "hipGetSymbolAddress of the static scratch on "
                           "device %d: %s", device, hipGetErrorString(e));
  g_noalloc_ptr[device] = static_cast<float*>(p);
  return FIR_OK;
}

float* noalloc_ptr(int device)
{
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  return device >= 0 && device < kNoAllocMaxDevices ? g_noalloc_ptr[device]
                                                    : nullptr;
}
} // namespace

namespace
{
// Scatter set-up: with bp_copies > 1, `n_images` images of n floats get
// bp_copies private copies each in the context's copies buffer (kept zero
// between launches); image b's copy c sits at b * n + c * n_images * n.
int ensure_copies(fir_ctx* c, size_t floats)
{
  if (c->copies_floats >= floats) return FIR_OK;
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (c->d_copies) c->retired.push_back(c->d_copies);
  c->d_copies = nullptr;
  c->copies_floats = 0;
  if (hipMalloc(&c->d_copies, floats * sizeof(float)) != hipSuccess)
    return fail(FIR_E_NOMEM, "allocation of %zu floats of scatter copies failed",
                floats);
  c->copies_floats = floats;
  hipLaunchKernelGGL(k_fill, dim3(ew_blocks((long long)floats)), dim3(kBlock),
                     0, c->stream, c->d_copies, (long long)floats, 0.0f);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

int scatter_for(fir_ctx* c, float* dst, long long n, int n_images,
                fir::Scatter* sc)
{
  if (c->bp_copies <= 1)
  {
    *sc = fir::Scatter{dst, 0, 1};
    return FIR_OK;
  }
  int rc = ensure_copies(c, (size_t)n * n_images * c->bp_copies);
  if (rc) return rc;
  *sc = fir::Scatter{c->d_copies, n * n_images, c->bp_copies};
  return FIR_OK;
}

// Add the copies of image `b` (of n_images) into dst and clear them.
int gather_copies(fir_ctx* c, float* dst, long long n, int n_images, int b)
{
  if (c->bp_copies <= 1) return FIR_OK;
  hipLaunchKernelGGL(k_reduce_copies, dim3(ew_blocks(n)), dim3(kBlock), 0,
                     c->stream, dst, c->d_copies + (long long)b * n, n,
                     c->bp_copies, n * n_images);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}
} // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

extern "C" {

const char* fir_last_error(void) { return g_err.c_str(); }

int fir_abi_version(void) { return 1; }

int fir_ctx_create(fir_ctx** out, int device)
{
  if (!out) return fail(FIR_E_ARG, "output pointer is NULL");
  *out = nullptr;
  int n = 0;
  FIR_HIP_CHECK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return fail(FIR_E_ARG, "device %d out of range (%d devices)", device, n);
  FIR_HIP_CHECK(hipSetDevice(device));
  fir_ctx* c = new fir_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess)
  {
    delete c;
    return fail(FIR_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  c->stream = c->own_stream;
  const int rc = noalloc_setup(device);
  if (rc)
  {
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return rc;
  }
  *out = c;
  return FIR_OK;
}

int fir_ctx_destroy(fir_ctx* c)
{
  if (!c) return FIR_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(c->d_xy);
  (void)hipFree(c->d_z);
  (void)hipFree(c->d_pairs);
  (void)hipFree(c->d_pairs_bins);
  (void)hipFree(c->d_ax_prefix);
  (void)hipFree(c->d_n_axial);
  (void)hipFree(c->d_seg_base);
  (void)hipFree(c->d_copies);
  (void)hipFree(c->d_scratch);
  for (auto& t : c->tap_sets) (void)hipFree(t.d);
  for (void* p : c->retired) (void)hipFree(p);
  (void)hipFree(c->d_counter);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return FIR_OK;
}

int fir_set_stream(fir_ctx* c, void* s)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  c->stream = (hipStream_t)s;
  return FIR_OK;
}

int fir_synchronize(fir_ctx* c)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  return FIR_OK;
}

int fir_set_scanner(fir_ctx* c, const double* xy, int n_cpr, const double* z,
                    int n_slices)
{
  if (!c || !xy || !z) return fail(FIR_E_ARG, "NULL argument");
  if (n_cpr <= 0 || n_slices <= 0)
    return fail(FIR_E_ARG, "empty scanner tables");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  (void)hipFree(c->d_xy);
  (void)hipFree(c->d_z);
  c->d_xy = nullptr;
  c->d_z = nullptr;
  FIR_HIP_CHECK(hipMalloc(&c->d_xy, sizeof(double2) * n_cpr));
  FIR_HIP_CHECK(hipMalloc(&c->d_z, sizeof(double) * n_slices));
  FIR_HIP_CHECK(hipMemcpy(c->d_xy, xy, sizeof(double2) * n_cpr,
                          hipMemcpyHostToDevice));
  FIR_HIP_CHECK(
    hipMemcpy(c->d_z, z, sizeof(double) * n_slices, hipMemcpyHostToDevice));
  c->n_cpr = n_cpr;
  c->n_slices = n_slices;
  ++c->generation;
  return FIR_OK;
}

// ProjHeader::check + ProjGeometry::fill + getCrystalAxialCoord tables.
int fir_set_proj(fir_ctx* c, int n_rings, int n_cpr, int span, int n_seg,
                 int n_tang)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (n_rings <= 0)
    return fail(FIR_E_ARG, "\"number of rings\" must be provided and > 0");
  if (n_cpr <= 0)
    return fail(FIR_E_ARG,
                "\"number of crystals per ring\" must be provided and > 0");
  if (n_cpr % 4 != 0)
    return fail(FIR_E_ARG,
                "\"number of crystals per ring\" must be a multiple of four");
  if (span <= 0) return fail(FIR_E_ARG, "\"segment span\" must be > 0");
  if (span % 2 != 1)
    return fail(FIR_E_ARG, "\"segment span\" must be an odd number");
  const int max_span = 2 * n_rings - 1;
  if (span > max_span)
    return fail(FIR_E_ARG,
                "For %d ring(s), \"segment span\" must not be > %d", n_rings,
                max_span);
  if (n_seg < 0)
    return fail(FIR_E_ARG, "\"number of segments\" must not be negative");
  if (n_seg % 2 != 1)
    return fail(FIR_E_ARG, "\"number of segments\" must be an odd number");
  int max_seg = max_span / span;
  max_seg -= 1 - max_seg % 2;
  if (n_seg > max_seg)
    return fail(FIR_E_ARG,
                "For %d ring(s) and a segment span of %d, \"number of "
                "segments\" must not be > %d",
                n_rings, span, max_seg);
  if (n_tang < 0)
    return fail(FIR_E_ARG,
                "\"number of tangential coordinates\" must not be negative");
  if (n_tang == 0)
    n_tang = n_cpr - 1;
  else if (n_tang > n_cpr - 1)
    return fail(FIR_E_ARG,
                "For %d crystals per ring, \"number of tangential "
                "coordinates\" must not be > %d",
                n_cpr, n_cpr - 1);

  const int seg_off = (n_seg - 1) / 2;
  const int n_views = n_cpr / 2;
  const int half = (span - 1) / 2;
  std::vector<int> na(n_seg);
  for (int s = -seg_off; s <= seg_off; ++s)
  {
    const int a = s < 0 ? -s : s;
    int len;
    if (span == 1)
      len = n_rings - a;
    else
    {
      len = 2 * n_rings - 1;
      if (a >= 1) len -= span + 1;
      if (a >= 2) len -= 2 * span * (a - 1);
    }
    na[s + seg_off] = len;
  }
  std::vector<int> pre(n_seg + 1, 0);
  std::vector<long long> base(n_seg, 0);
  for (int k = 0; k < n_seg; ++k)
  {
    pre[k + 1] = pre[k] + na[k];
    base[k] = (long long)pre[k] * n_views * n_tang;
  }
  const long long nbins = (long long)pre[n_seg] * n_views * n_tang;

  // slice pair of every (seg, axial): ProjData::getCrystalAxialCoord
  // (ProjData.cc:198-299)
  std::vector<int2> sl(pre[n_seg]);
  for (int s = -seg_off; s <= seg_off; ++s)
  {
    const int k = s + seg_off;
    for (int ax = 0; ax < na[k]; ++ax)
    {
      int s1, s2;
      if (span == 1)
      {
        if (s == 0)
          s1 = s2 = 2 * ax;
        else if (s > 0)
        {
          s1 = 2 * (ax + s);
          s2 = 2 * ax;
        }
        else
        {
          s1 = 2 * ax;
          s2 = 2 * (ax - s);
        }
      }
      else if (s == 0)
        s1 = s2 = ax;
      else
      {
        const int mid = (s < 0 ? -s : s) * span;
        if (ax < half)
        {
          const int rd = mid - half + ax;
          s1 = s > 0 ? 2 * rd : 0;
          s2 = s > 0 ? 0 : 2 * rd;
        }
        else
        {
          const int inv = na[k] - ax - 1;
          if (inv < half)
          {
            const int rd = mid - half + inv;
            const int smax = 2 * n_rings - 2;
            s1 = s > 0 ? smax : smax - 2 * rd;
            s2 = s > 0 ? smax - 2 * rd : smax;
          }
          else
          {
            const int adj = ax - half;
            s1 = s > 0 ? 2 * mid + adj : adj;
            s2 = s > 0 ? adj : 2 * mid + adj;
          }
        }
      }
      if (s1 < 0 || s2 < 0 || s1 > 2 * n_rings - 2 || s2 > 2 * n_rings - 2)
        return fail(FIR_E_ARG, "slice index out of range for segment %d", s);
      sl[pre[k] + ax] = make_int2(s1, s2);
    }
  }

  FIR_HIP_CHECK(hipSetDevice(c->device));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  (void)hipFree(c->d_n_axial);
  (void)hipFree(c->d_seg_base);
  (void)hipFree(c->d_ax_prefix);
  c->d_n_axial = nullptr;
  c->d_seg_base = nullptr;
  c->d_ax_prefix = nullptr;
  c->proj_set = false;
  FIR_HIP_CHECK(hipMalloc(&c->d_ax_prefix, sizeof(int) * pre.size()));
  FIR_HIP_CHECK(hipMemcpy(c->d_ax_prefix, pre.data(), sizeof(int) * pre.size(),
                          hipMemcpyHostToDevice));
  FIR_HIP_CHECK(hipMalloc(&c->d_n_axial, sizeof(int) * na.size()));
  FIR_HIP_CHECK(hipMalloc(&c->d_seg_base, sizeof(long long) * base.size()));
  FIR_HIP_CHECK(hipMemcpy(c->d_n_axial, na.data(), sizeof(int) * na.size(),
                          hipMemcpyHostToDevice));
  FIR_HIP_CHECK(hipMemcpy(c->d_seg_base, base.data(),
                          sizeof(long long) * base.size(),
                          hipMemcpyHostToDevice));
  c->n_rings = n_rings;
  c->span = span;
  c->n_seg = n_seg;
  c->n_tang = n_tang;
  c->seg_off = seg_off;
  c->tang_off = n_tang / 2;
  c->n_views = n_views;
  c->half_span = half;
  c->n_bins = nbins;
  c->n_axial = na;
  c->seg_base = base;
  c->slices = sl;
  int rc = upload_pairs(c);
  if (rc) return rc;
  c->proj_set = true;
  ++c->generation;
  return FIR_OK;
}

int64_t fir_proj_n_bins(const fir_ctx* c)
{
  return (c && c->proj_set) ? c->n_bins : 0;
}

int fir_set_walk_mode(fir_ctx* c, int mode)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (mode != FIR_WALK_REFERENCE && mode != FIR_WALK_GEOMETRIC)
    return fail(FIR_E_ARG, "unknown walk mode %d", mode);
  if (c->walk_mode != mode) ++c->generation;
  c->walk_mode = mode;
  return FIR_OK;
}

int fir_malloc(fir_ctx* c, void** p, size_t bytes)
{
  if (!c || !p) return fail(FIR_E_ARG, "NULL argument");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  if (hipMalloc(p, bytes) != hipSuccess)
  {
    *p = nullptr;
    return fail(FIR_E_NOMEM, "hipMalloc of %zu bytes failed", bytes);
  }
  return FIR_OK;
}

int fir_host_alloc(void** p, size_t bytes)
{
  if (!p) return fail(FIR_E_ARG, "NULL argument");
  *p = nullptr;
  if (hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess)
  {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(FIR_E_NOMEM, "hipHostMalloc of %zu bytes failed", bytes);
  }
  return FIR_OK;
}

int fir_host_free(void* p)
{
  if (!p) return FIR_OK;
  FIR_HIP_CHECK(hipHostFree(p));
  return FIR_OK;
}

int fir_free(fir_ctx* c, void* p)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  FIR_HIP_CHECK(hipFree(p));
  return FIR_OK;
}

int fir_memcpy_h2d(fir_ctx* c, void* dst, const void* src, size_t bytes)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  FIR_HIP_CHECK(
    hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  return FIR_OK;
}

int fir_memcpy_d2h(fir_ctx* c, void* dst, const void* src, size_t bytes)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  FIR_HIP_CHECK(
    hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  return FIR_OK;
}

int fir_memcpy_h2d_async(fir_ctx* c, void* dst, const void* src,
                         size_t bytes)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (bytes == 0) return FIR_OK;
  if (!dst || !src) return fail(FIR_E_ARG, "NULL buffer");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  FIR_HIP_CHECK(
    hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return FIR_OK;
}

int fir_memcpy_d2h_async(fir_ctx* c, void* dst, const void* src,
                         size_t bytes)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (bytes == 0) return FIR_OK;
  if (!dst || !src) return fail(FIR_E_ARG, "NULL buffer");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  FIR_HIP_CHECK(
    hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  return FIR_OK;
}

int fir_fill(fir_ctx* c, float* dst, int64_t n, float value)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (n <= 0) return FIR_OK;
  if (!dst) return fail(FIR_E_ARG, "NULL buffer");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_fill, dim3(ew_blocks(n)), dim3(kBlock), 0, c->stream,
                     dst, (long long)n, value);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

int fir_forward_subset(fir_ctx* c, const fir_vol_desc* v, const float* img,
                       float* sino, int subset, int n_subsets)
{
  int rc = need_lors(c);
  if (rc) return rc;
  if (!img || !sino) return fail(FIR_E_ARG, "NULL buffer");
  Grid g;
  if ((rc = make_grid(c, v, &g))) return rc;
  if (n_subsets > 0)
  {
    if ((rc = check_subsets(c, n_subsets))) return rc;
    if (subset < 0 || subset >= n_subsets)
      return fail(FIR_E_ARG, "subset %d out of range", subset);
  }
  const long long n = n_work(c, n_subsets, subset);
  if ((rc = check_work(n))) return rc;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_forward, dim3(blocks_for(n)), dim3(kBlock), 0,
                     c->stream, g, lor_geom(c), img, sino, n, n_subsets,
                     subset);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

int fir_forward(fir_ctx* c, const fir_vol_desc* v, const float* img,
                float* sino)
{
  return fir_forward_subset(c, v, img, sino, 0, 0);
}

int fir_backward_subset(fir_ctx* c, const fir_vol_desc* v, const float* sino,
                        float* img, int subset, int n_subsets)
{
  int rc = need_lors(c);
  if (rc) return rc;
  if (!img) return fail(FIR_E_ARG, "NULL image");
  Grid g;
  if ((rc = make_grid(c, v, &g))) return rc;
  if ((rc = check_subsets(c, n_subsets))) return rc;
  if (subset < 0 || subset >= n_subsets)
    return fail(FIR_E_ARG, "subset %d out of range", subset);
  const long long n = n_work(c, n_subsets, subset);
  if ((rc = check_work(n))) return rc;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  const long long nvox = (long long)v->nx * v->ny * v->nz;
  fir::Scatter sc;
  if ((rc = scatter_for(c, img, nvox, 1, &sc))) return rc;
  if (family_ok(c, v))
    launch_family<false>(c, g, v, nullptr, sino, nullptr, sc, n_subsets,
                         subset);
  else
    hipLaunchKernelGGL(k_backward, dim3(blocks_for(n)), dim3(kBlock), 0,
                       c->stream, g, lor_geom(c), sino, sc, n, n_subsets,
                       subset);
  FIR_CHECK_LAUNCH();
  return gather_copies(c, img, nvox, 1, 0);
}

int fir_backward(fir_ctx* c, const fir_vol_desc* v, const float* sino,
                 float* img, int n_subsets)
{
  int rc = need_lors(c);
  if (rc) return rc;
  if ((rc = check_subsets(c, n_subsets))) return rc;
  Grid g;
  if ((rc = make_grid(c, v, &g))) return rc;
  const long long nvox = (long long)v->nx * v->ny * v->nz;
  if ((rc = fir_fill(c, img, nvox * n_subsets, 0.0f))) return rc;
  for (int s = 0; s < n_subsets; ++s)
    if ((rc = fir_backward_subset(c, v, sino, img + nvox * s, s, n_subsets)))
      return rc;
  return FIR_OK;
}

int fir_sensitivity(fir_ctx* c, const fir_vol_desc* v, float* sens,
                    int n_subsets)
{
  return fir_backward(c, v, nullptr, sens, n_subsets);
}

int fir_osem_project(fir_ctx* c, const fir_vol_desc* v, const float* x,
                     const float* y, const float* bias, float* bp, int subset,
                     int n_subsets)
{
  int rc = need_lors(c);
  if (rc) return rc;
  if (!x || !y || !bp) return fail(FIR_E_ARG, "NULL buffer");
  Grid g;
  if ((rc = make_grid(c, v, &g))) return rc;
  if ((rc = check_subsets(c, n_subsets))) return rc;
  if (subset < 0 || subset >= n_subsets)
    return fail(FIR_E_ARG, "subset %d out of range", subset);
  const long long n = n_work(c, n_subsets, subset);
  if ((rc = check_work(n))) return rc;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  const long long nvox = (long long)v->nx * v->ny * v->nz;
  fir::Scatter sc;
  if ((rc = scatter_for(c, bp, nvox, 1, &sc))) return rc;
  if (family_ok(c, v))
    launch_family<true>(c, g, v, x, y, bias, sc, n_subsets, subset);
  else
    hipLaunchKernelGGL(k_osem_project, dim3(blocks_for(n)), dim3(kBlock), 0,
                       c->stream, g, lor_geom(c), x, y, bias, sc, n, n_subsets,
                       subset);
  FIR_CHECK_LAUNCH();
  return gather_copies(c, bp, nvox, 1, 0);
}

int fir_osem_update(fir_ctx* c, const fir_vol_desc* v, float* x, float* bp,
                    const float* sens_s, float cut, int apply_cut, int reset)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (!x || !bp || !sens_s) return fail(FIR_E_ARG, "NULL buffer");
  Grid g;
  int rc;
  if ((rc = make_grid(c, v, &g))) return rc;
  const long long n = (long long)v->nx * v->ny * v->nz;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_osem_update, dim3(ew_blocks(n)), dim3(kBlock), 0,
                     c->stream, x, bp, sens_s, n, cut_geom(v), cut,
                     (apply_cut && cut > 0.0f) ? 1 : 0, reset ? 1 : 0);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

namespace
{
// Gaussian taps of operations::convolve (operations.cc:31-75) for (fwhm,
// voxel size), computed on the host: X taps at [0, 64), Y at [64, 128), Z at
// [128, 192), h[d] the half-width per axis. X and Y use a float sigma, Z a
// double sigma; taps are computed in double, stored as float.
int gauss_taps_host(const fir_vol_desc* v, const float* fwhm, float* taps,
                    int* h)
{
  const float sx = fwhm[0] / 2.3548f;
  const float sy = fwhm[1] / 2.3548f;
  const double sz = fwhm[2] / 2.3548;
  h[0] = ((int)(6.0 * sx / v->dx) + 1) / 2;
  h[1] = ((int)(6.0 * sy / v->dy) + 1) / 2;
  h[2] = ((int)(6.0 * sz / v->dz) + 1) / 2;
  for (int d = 0; d < 3; ++d)
    if (h[d] > 31 || h[d] < 0)
      return fail(FIR_E_ARG, "Gaussian kernel wider than 63 taps");
  std::memset(taps, 0, sizeof(float) * kGaussTapFloats);
  for (int k = -h[0]; k <= h[0]; ++k)
    taps[h[0] + k] = (float)std::exp(-(k * k * v->dx * v->dx) / (2 * sx * sx));
  for (int k = -h[1]; k <= h[1]; ++k)
    taps[64 + h[1] + k] =
      (float)std::exp(-(k * k * v->dy * v->dy) / (2 * sy * sy));
  for (int k = -h[2]; k <= h[2]; ++k)
    taps[128 + h[2] + k] =
      (float)std::exp(-(k * k * v->dz * v->dz) / (2 * sz * sz));
  return FIR_OK;
}

// The context's tap set for (fwhm, voxel size), uploaded on first use.
int prepare_taps(fir_ctx* c, const fir_vol_desc* v, const float* fwhm,
                 const fir_ctx::TapSet** out)
{
  for (const auto& t : c->tap_sets)
    if (t.key[0] == fwhm[0] && t.key[1] == fwhm[1] && t.key[2] == fwhm[2] &&
        t.vox[0] == v->dx && t.vox[1] == v->dy && t.vox[2] == v->dz)
    {
      *out = &t;
      return FIR_OK;
    }
  float taps[kGaussTapFloats];
  fir_ctx::TapSet t;
  int rc = gauss_taps_host(v, fwhm, taps, t.h);
  if (rc) return rc;
  FIR_HIP_CHECK(hipMalloc(&t.d, sizeof(taps)));
  if (hipMemcpyAsync(t.d, taps, sizeof(taps), hipMemcpyHostToDevice,
                     c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
  {
    (void)hipFree(t.d);
    return fail(FIR_E_HIP, "upload of the Gaussian taps failed");
  }
  for (int d = 0; d < 3; ++d) t.key[d] = fwhm[d];
  t.vox[0] = v->dx;
  t.vox[1] = v->dy;
  t.vox[2] = v->dz;
  c->tap_sets.push_back(t);
  *out = &c->tap_sets.back();
  return FIR_OK;
}

int ensure_scratch(fir_ctx* c, size_t floats)
{
  if (c->scratch_floats >= floats) return FIR_OK;
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (c->d_scratch) c->retired.push_back(c->d_scratch);
  c->d_scratch = nullptr;
  c->scratch_floats = 0;
  if (hipMalloc(&c->d_scratch, floats * sizeof(float)) != hipSuccess)
    return fail(FIR_E_NOMEM, "scratch allocation of %zu floats failed",
                floats);
  c->scratch_floats = floats;
  return FIR_OK;
}

// Gaussian of one frame plus edge restore and cut (fused in the Z pass):
// X pass img -> t1, Y pass t1 -> t2, Z pass t2 -> img. Every voxel of t1
// and t2 is written before it is read, so they need no clearing. taps is a
// device array laid out as gauss_taps_host's, h its half-widths.
int convolve_frame(fir_ctx* c, const fir_vol_desc* v, float* img,
                   const float* taps, const int* h, float* t1, float* t2,
                   const float* fwhm, float restore_cut, float final_cut)
{
  const CutGeom cg = cut_geom(v);
  const float fw = fwhm[0] > fwhm[1] ? fwhm[0] : fwhm[1];
  const int restore = restore_cut > 0.0f ? 1 : 0;
  const double rr = (double)restore_cut - 5.0 * fw;
  // (small frames keep the one-voxel-per-thread passes: four outputs per
  // thread leave too few threads to hide the loads' latency)
  if (gauss_vec_ok(v->nx, h) &&
      (long long)v->nx * v->ny * v->nz >= kGaussVecMinVoxels &&
      ((reinterpret_cast<uintptr_t>(img) | reinterpret_cast<uintptr_t>(t1) |
        reinterpret_cast<uintptr_t>(t2)) & 15u) == 0)
  {
    launch_gauss_v(c->stream, 0, h[0], img, t1, taps, v, cg, 0, 0.0, 0.0f);
    launch_gauss_v(c->stream, 1, h[1], t1, t2, taps + 64, v, cg, 0, 0.0,
                   0.0f);
    launch_gauss_v(c->stream, 2, h[2], t2, img, taps + 128, v, cg, restore,
                   rr, final_cut);
    FIR_CHECK_LAUNCH();
    return FIR_OK;
  }
  const dim3 nb((v->nx + kGaussX - 1) / kGaussX,
                (v->ny + kGaussY - 1) / kGaussY, v->nz);
  hipLaunchKernelGGL((k_gauss<0, false>), nb, dim3(kBlock), 0,
                     c->stream, img, t1, taps, h[0], v->nx,
                     v->ny, v->nz, cg, 0, 0.0, 0.0f);
  hipLaunchKernelGGL((k_gauss<1, false>), nb, dim3(kBlock), 0,
                     c->stream, t1, t2, taps + 64, h[1], v->nx,
                     v->ny, v->nz, cg, 0, 0.0, 0.0f);
  // Z pass writes img in place: each voxel reads its own img value (the
  // pre-convolution copy) only for the edge restore.
  hipLaunchKernelGGL((k_gauss<2, true>), nb, dim3(kBlock), 0, c->stream,
                     t2, img, taps + 128, h[2], v->nx, v->ny,
                     v->nz, cg, restore, rr, final_cut);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

// Gaussian of n_frames frames through the context's tap set and scratch.
int convolve_ctx(fir_ctx* c, const fir_vol_desc* v, float* img, int n_frames,
                 const float* fwhm, float restore_cut, float final_cut)
{
  const long long n = (long long)v->nx * v->ny * v->nz;
  const fir_ctx::TapSet* ts = nullptr;
  int rc;
  if ((rc = prepare_taps(c, v, fwhm, &ts))) return rc;
  const float* taps = ts->d;
  const int h[3] = {ts->h[0], ts->h[1], ts->h[2]};
  if ((rc = ensure_scratch(c, 2 * (size_t)n))) return rc;
  for (int f = 0; f < n_frames; ++f)
    if ((rc = convolve_frame(c, v, img + n * f, taps, h, c->d_scratch,
                             c->d_scratch + n, fwhm, restore_cut, final_cut)))
      return rc;
  return FIR_OK;
}
} // namespace

} // extern "C"

namespace firi
{
int step_image_tiles(int nx, int ny, int nz)
{
  return ((nx + kGaussX - 1) / kGaussX) * ((ny + kGaussY - 1) / kGaussY) * nz;
}

int step_image_vtiles(int nx, int ny, int nz, int axis)
{
  return (int)((gauss_units(axis, nx, ny, nz) + kBlock - 1) / kBlock);
}

bool gauss_vec_shape(int nx, const int* h)
{
  return gauss_vec_ok(nx, h);
}

int gauss_taps_for(fir_ctx* c, const fir_vol_desc* v, const float* fwhm,
                   const float** taps, int h[3])
{
  const fir_ctx::TapSet* ts = nullptr;
  const int rc = prepare_taps(c, v, fwhm, &ts);
  if (rc) return rc;
  *taps = ts->d;
  for (int d = 0; d < 3; ++d) h[d] = ts->h[d];
  return FIR_OK;
}

template <int AXIS>
void launch_step_gauss_v(hipStream_t st, const StepImage* imgs, int n_imgs,
                         int blocks, int h, int restore, double restore_r,
                         float cut)
{
#define FIR_STEP_GAUSS_V(HH)                                                \
  hipLaunchKernelGGL((k_step_gauss_v<AXIS, HH>), dim3(blocks), dim3(kBlock), \
                     0, st, imgs, n_imgs, restore, restore_r, cut)
  switch (h)
  {
  case 0: FIR_STEP_GAUSS_V(0); break;
  case 1: FIR_STEP_GAUSS_V(1); break;
  case 2: FIR_STEP_GAUSS_V(2); break;
  case 3: FIR_STEP_GAUSS_V(3); break;
  default: FIR_STEP_GAUSS_V(4); break;
  }
#undef FIR_STEP_GAUSS_V
}

int launch_step_elementwise(hipStream_t st, const StepImage* imgs, int n_imgs,
                            int n_tiles, bool gauss, float cut, int restore,
                            double restore_r, const int* vtiles,
                            const int* vh)
{
  if (n_tiles <= 0) return FIR_OK;
  hipLaunchKernelGGL(k_step_update, dim3(n_tiles), dim3(kBlock), 0, st, imgs,
                     n_imgs, cut, (!gauss && cut > 0.0f) ? 1 : 0);
  FIR_CHECK_LAUNCH();
  if (!gauss) return FIR_OK;
  if (vtiles)
  {
    launch_step_gauss_v<0>(st, imgs, n_imgs, vtiles[0], vh[0], 0, 0.0, 0.0f);
    launch_step_gauss_v<1>(st, imgs, n_imgs, vtiles[1], vh[1], 0, 0.0, 0.0f);
    launch_step_gauss_v<2>(st, imgs, n_imgs, vtiles[2], vh[2], restore,
                           restore_r, cut);
    FIR_CHECK_LAUNCH();
    return FIR_OK;
  }
  hipLaunchKernelGGL(k_step_gauss<0>, dim3(n_tiles), dim3(kBlock), 0, st,
                     imgs, n_imgs, 0, 0.0, 0.0f);
  hipLaunchKernelGGL(k_step_gauss<1>, dim3(n_tiles), dim3(kBlock), 0, st,
                     imgs, n_imgs, 0, 0.0, 0.0f);
  hipLaunchKernelGGL(k_step_gauss<2>, dim3(n_tiles), dim3(kBlock), 0, st,
                     imgs, n_imgs, restore, restore_r, cut);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}
} // namespace firi

extern "C" {

int fir_convolve(fir_ctx* c, const fir_vol_desc* v, float* img, int n_frames,
                 const float* fwhm, float cut)
{
  if (!c || !fwhm) return fail(FIR_E_ARG, "NULL argument");
  if (!img) return fail(FIR_E_ARG, "NULL image");
  Grid g;
  int rc;
  if ((rc = make_grid(c, v, &g))) return rc;
  if (!(fwhm[0] > 0.0 && fwhm[1] > 0.0 && fwhm[2] > 0.0)) return FIR_OK;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  return convolve_ctx(c, v, img, n_frames, fwhm, cut, 0.0f);
}

int fir_convolve_noalloc(fir_ctx* c, const fir_vol_desc* v, float* img,
                         int n_frames, const float* fwhm, float cut)
{
  if (!c || !fwhm) return fail(FIR_E_ARG, "NULL argument");
  if (!img) return fail(FIR_E_ARG, "NULL image");
  Grid g;
  int rc;
  if ((rc = make_grid(c, v, &g))) return rc;
  if (!(fwhm[0] > 0.0 && fwhm[1] > 0.0 && fwhm[2] > 0.0)) return FIR_OK;
  const long long n = (long long)v->nx * v->ny * v->nz;
  if (n > kNoAllocMaxVoxels)
    return fail(FIR_E_ARG, "fir_convolve_noalloc: %lld voxels per frame "
                           "exceed the static scratch's %lld",
                n, kNoAllocMaxVoxels);
  GaussTaps t;
  int h[3];
  if ((rc = gauss_taps_host(v, fwhm, t.v, h))) return rc;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  float* base = noalloc_ptr(c->device);
  if (!base)
    return fail(FIR_E_STATE, "static scratch of device %d not set up",
                c->device);
  float* t1 = base + kNoAllocTapSlot;
  float* t2 = t1 + n;
  hipLaunchKernelGGL(k_store_taps, dim3(1), dim3(kGaussTapFloats), 0,
                     c->stream, t, base);
  FIR_CHECK_LAUNCH();
  for (int f = 0; f < n_frames; ++f)
    if ((rc = convolve_frame(c, v, img + n * f, base, h, t1, t2, fwhm, cut,
                             0.0f)))
      return rc;
  return FIR_OK;
}

int fir_cut_circle(fir_ctx* c, const fir_vol_desc* v, float* img,
                   int n_frames, float cut)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (!img) return fail(FIR_E_ARG, "NULL image");
  Grid g;
  int rc;
  if ((rc = make_grid(c, v, &g))) return rc;
  if (!(cut > 0.0f)) return FIR_OK;
  const long long n = (long long)v->nx * v->ny * v->nz * n_frames;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_cut_circle, dim3(ew_blocks(n)), dim3(kBlock), 0,
                     c->stream, img, n, cut_geom(v), cut);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

int fir_osem_subiter(fir_ctx* c, const fir_vol_desc* v, float* x,
                     const float* y, const float* bias, const float* sens_s,
                     float* bp, int subset, int n_subsets, int do_convolve,
                     const float* fwhm, float cut, int reset_bp)
{
  int rc = fir_osem_project(c, v, x, y, bias, bp, subset, n_subsets);
  if (rc) return rc;
  const bool conv = do_convolve && fwhm && fwhm[0] > 0.0 && fwhm[1] > 0.0 &&
    fwhm[2] > 0.0;
  if ((rc = fir_osem_update(c, v, x, bp, sens_s, cut, conv ? 0 : 1,
                            reset_bp)))
    return rc;
  if (conv)
  {
    FIR_HIP_CHECK(hipSetDevice(c->device));
    if ((rc = convolve_ctx(c, v, x, 1, fwhm, cut, cut))) return rc;
  }
  return FIR_OK;
}

int fir_mul_clamped(fir_ctx* c, float* a, const float* b, int64_t n)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (n <= 0) return FIR_OK;
  if (!a || !b) return fail(FIR_E_ARG, "NULL buffer");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_mul_clamped, dim3(ew_blocks(n)), dim3(kBlock), 0,
                     c->stream, a, b, (long long)n);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

int fir_div_clamped(fir_ctx* c, float* a, const float* b, int64_t n)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (n <= 0) return FIR_OK;
  if (!a || !b) return fail(FIR_E_ARG, "NULL buffer");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_div_clamped, dim3(ew_blocks(n)), dim3(kBlock), 0,
                     c->stream, a, b, (long long)n);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

int fir_exp_clamped(fir_ctx* c, float* a, int64_t n)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (n <= 0) return FIR_OK;
  if (!a) return fail(FIR_E_ARG, "NULL buffer");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_exp_clamped, dim3(ew_blocks(n)), dim3(kBlock), 0,
                     c->stream, a, (long long)n);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

int fir_hu_to_mu(fir_ctx* c, float* a, int64_t n)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (n <= 0) return FIR_OK;
  if (!a) return fail(FIR_E_ARG, "NULL buffer");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_hu_to_mu, dim3(ew_blocks(n)), dim3(kBlock), 0,
                     c->stream, a, (long long)n);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

int fir_path_dump(fir_ctx* c, const fir_vol_desc* v, int64_t bin0,
                  int64_t nb, int maxp, int* counts, int* idx, double* len)
{
  int rc = need_lors(c);
  if (rc) return rc;
  Grid g;
  if ((rc = make_grid(c, v, &g))) return rc;
  if (bin0 < 0 || nb < 0 || bin0 + nb > c->n_bins)
    return fail(FIR_E_ARG, "bin range out of bounds");
  if (maxp <= 0) return fail(FIR_E_ARG, "max_per_lor must be > 0");
  if (!counts || !idx || !len) return fail(FIR_E_ARG, "NULL buffer");
  if (nb == 0) return FIR_OK;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  if ((rc = check_work(c->n_bins))) return rc;
  hipLaunchKernelGGL(k_path_dump, dim3(blocks_for(nb)), dim3(kBlock), 0,
                     c->stream, g, lor_geom_bins(c), c->d_ax_prefix, c->n_seg,
                     (long long)bin0,
                     (long long)nb, maxp, counts, idx, len);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

int fir_path_count(fir_ctx* c, const fir_vol_desc* v, int subset,
                   int n_subsets, int64_t* total)
{
  int rc = need_lors(c);
  if (rc) return rc;
  if (!total) return fail(FIR_E_ARG, "NULL output");
  Grid g;
  if ((rc = make_grid(c, v, &g))) return rc;
  if (n_subsets > 0)
  {
    if ((rc = check_subsets(c, n_subsets))) return rc;
    if (subset < 0 || subset >= n_subsets)
      return fail(FIR_E_ARG, "subset %d out of range", subset);
  }
  FIR_HIP_CHECK(hipSetDevice(c->device));
  if (!c->d_counter)
    FIR_HIP_CHECK(hipMalloc(&c->d_counter, sizeof(unsigned long long)));
  const long long n = n_work(c, n_subsets, subset);
  if ((rc = check_work(n))) return rc;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, c->stream, c->d_counter);
  hipLaunchKernelGGL(k_path_count, dim3(blocks_for(n)), dim3(kBlock), 0,
                     c->stream, g, lor_geom(c), n, n_subsets, subset,
                     c->d_counter);
  FIR_CHECK_LAUNCH();
  unsigned long long h = 0;
  FIR_HIP_CHECK(hipMemcpyAsync(&h, c->d_counter, sizeof(h),
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  *total = (int64_t)h;
  return FIR_OK;
}

int fir_poisson_nll_grad(fir_ctx* c, const fir_vol_desc* v, const float* x,
                         const float* y, const float* r, float* grad,
                         double* loss)
{
  int rc = need_lors(c);
  if (rc) return rc;
  if (!x || !y || !grad || !loss) return fail(FIR_E_ARG, "NULL buffer");
  Grid g;
  if ((rc = make_grid(c, v, &g))) return rc;
  const long long n = c->n_bins;
  const long long nvox = (long long)v->nx * v->ny * v->nz;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_fill, dim3(ew_blocks(nvox)), dim3(kBlock), 0,
                     c->stream, grad, nvox, 0.0f);
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(1), 0, c->stream, loss);
  if ((rc = check_work(n))) return rc;
  fir::Scatter sc;
  if ((rc = scatter_for(c, grad, nvox, 1, &sc))) return rc;
  hipLaunchKernelGGL(k_poisson_grad, dim3(blocks_for(n)), dim3(kBlock), 0,
                     c->stream, g, lor_geom(c), x, y, r, sc, loss, n);
  FIR_CHECK_LAUNCH();
  return gather_copies(c, grad, nvox, 1, 0);
}

int fir_osem_project_batch(fir_ctx* c, const fir_vol_desc* v, int n_images,
                           const float* const* x, const float* const* y,
                           const float* const* bias, float* const* bp,
                           int subset, int n_subsets)
{
  int rc = need_lors(c);
  if (rc) return rc;
  if (n_images <= 0) return fail(FIR_E_ARG, "n_images must be > 0");
  if (!x || !y || !bp) return fail(FIR_E_ARG, "NULL pointer array");
  for (int i = 0; i < n_images; ++i)
    if (!x[i] || !y[i] || !bp[i]) return fail(FIR_E_ARG, "NULL buffer");
  Grid g;
  if ((rc = make_grid(c, v, &g))) return rc;
  if ((rc = check_subsets(c, n_subsets))) return rc;
  if (subset < 0 || subset >= n_subsets)
    return fail(FIR_E_ARG, "subset %d out of range", subset);
  const long long n = n_work(c, n_subsets, subset);
  if ((rc = check_work(n))) return rc;
  const long long nvox = (long long)v->nx * v->ny * v->nz;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  for (int i0 = 0; i0 < n_images; i0 += kMaxBatch)
  {
    const int nb = std::min(kMaxBatch, n_images - i0);
    BatchPtrs p;
    std::memset(&p, 0, sizeof(p));
    fir::Scatter sc;
    if ((rc = scatter_for(c, nullptr, nvox, nb, &sc))) return rc;
    for (int b = 0; b < nb; ++b)
    {
      p.x[b] = x[i0 + b];
      p.y[b] = y[i0 + b];
      p.bias[b] = bias ? bias[i0 + b] : nullptr;
      p.bp[b] = sc.copies > 1 ? sc.base + (long long)b * nvox : bp[i0 + b];
    }
    p.copy_stride = sc.stride;
    p.copies = sc.copies;
    const dim3 grid(blocks_for(n)), block(kBlock);
    const LorGeom lg = lor_geom(c);
    switch (nb)
    {
    case 1: hipLaunchKernelGGL(k_osem_project_batch<1>, grid, block, 0, c->stream, g, lg, p, n, n_subsets, subset); break;
    case 2: hipLaunchKernelGGL(k_osem_project_batch<2>, grid, block, 0, c->stream, g, lg, p, n, n_subsets, subset); break;
    case 3: hipLaunchKernelGGL(k_osem_project_batch<3>, grid, block, 0, c->stream, g, lg, p, n, n_subsets, subset); break;
    case 4: hipLaunchKernelGGL(k_osem_project_batch<4>, grid, block, 0, c->stream, g, lg, p, n, n_subsets, subset); break;
    case 5: hipLaunchKernelGGL(k_osem_project_batch<5>, grid, block, 0, c->stream, g, lg, p, n, n_subsets, subset); break;
    case 6: hipLaunchKernelGGL(k_osem_project_batch<6>, grid, block, 0, c->stream, g, lg, p, n, n_subsets, subset); break;
    case 7: hipLaunchKernelGGL(k_osem_project_batch<7>, grid, block, 0, c->stream, g, lg, p, n, n_subsets, subset); break;
    default: hipLaunchKernelGGL(k_osem_project_batch<8>, grid, block, 0, c->stream, g, lg, p, n, n_subsets, subset); break;
    }
    FIR_CHECK_LAUNCH();
    for (int b = 0; b < nb; ++b)
      if ((rc = gather_copies(c, bp[i0 + b], nvox, nb, b))) return rc;
  }
  return FIR_OK;
}

int fir_set_uneven_subsets(fir_ctx* c, int allow)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  c->uneven_subsets = allow != 0;
  return FIR_OK;
}

int fir_set_lor_order(fir_ctx* c, int order)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  const int pair_order_id = order & 0xff;
  const int rows = order & (FIR_ORDER_PAIR_MAJOR | FIR_ORDER_VIEW_MAJOR);
  if (order & ~(0xff | FIR_ORDER_PAIR_MAJOR | FIR_ORDER_VIEW_MAJOR |
                FIR_ORDER_TANG_MAJOR))
    return fail(FIR_E_ARG, "unknown LOR order flags 0x%x", order);
  if (pair_order_id != FIR_ORDER_SEGMENT && pair_order_id != FIR_ORDER_ZCENTER)
    return fail(FIR_E_ARG, "unknown LOR order %d", order);
  if (rows == (FIR_ORDER_PAIR_MAJOR | FIR_ORDER_VIEW_MAJOR))
    return fail(FIR_E_ARG, "FIR_ORDER_PAIR_MAJOR and FIR_ORDER_VIEW_MAJOR "
                           "exclude each other");
  if (rows == 0 && pair_order_id != FIR_ORDER_SEGMENT)
    return fail(FIR_E_ARG, "segment-major rows need FIR_ORDER_SEGMENT pairs");
  c->lor_order = pair_order_id;
  c->row_mode = rows == FIR_ORDER_PAIR_MAJOR ? fir::ROWS_PAIR_MAJOR
              : rows == FIR_ORDER_VIEW_MAJOR ? fir::ROWS_VIEW_MAJOR
                                             : fir::ROWS_SEGMENT_VIEW;
  c->tang_major = (order & FIR_ORDER_TANG_MAJOR) ? 1 : 0;
  if (!c->proj_set) return FIR_OK;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  return upload_pairs(c);
}

int fir_set_walk_sheets(fir_ctx* c, int enable)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  c->walk_sheets = enable ? 1 : 0;
  return FIR_OK;
}

int fir_set_scatter_copies(fir_ctx* c, int copies)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (copies != 1 && copies != 2 && copies != 4 && copies != 8)
    return fail(FIR_E_ARG, "scatter copies must be 1, 2, 4 or 8");
  c->bp_copies = copies;
  return FIR_OK;
}

} // extern "C"
